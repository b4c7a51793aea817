"""Network front end on the host: file/string parsing, the ORDER grammar and the order search.

The search is checked against the reference's own known answers (tests/search_tree_test.cpp:
costs 32 / 1536 / 284 / 120 and their order strings, the 5-tensor chain, the 21-tensor iPEPS
network, the 64-tensor limit), against an independent recomputation of the cost of the tree it
returns, and against exhaustive search on small random networks (greedy never beats the optimum
and equals it on chains of matrices with a monotone dimension profile).
"""
import itertools

import numpy as np
import pytest

from cytnx_amd.network import (Network, parse_order, parse_tout, search_order, tokenize_order,
                               tree_str)

DMRG = ["psi: -1,-2,-3,-4", "L: -5,-1,0", "R: -7,-4,3", "M1: -5,-6,-2,1", "M2: -6,-7,-3,2",
        "TOUT: 0,1;2,3"]


def _ids(n):
    return [str(i) for i in range(n)]


def _search(labels, shapes):
    tree, cost = search_order(labels, shapes)
    return tree_str(tree, _ids(len(labels))), cost


def _tree_cost(tree, labels, shapes):
    """Cost of a given tree under prod(shape1)*prod(shape2)/prod(common) (exact integers)."""
    dims = {l: s for ls, ss in zip(labels, shapes) for l, s in zip(ls, ss)}

    def rec(t):
        if isinstance(t, int):
            return list(labels[t]), 0
        la, ca = rec(t[0])
        lb, cb = rec(t[1])
        common = [l for l in lb if l in la]
        c = int(np.prod([dims[l] for l in la])) * int(np.prod([dims[l] for l in lb]))
        c //= int(np.prod([dims[l] for l in common])) if common else 1
        out = [l for l in la if l not in common] + [l for l in lb if l not in common]
        return out, c + ca + cb

    return rec(tree)


def test_search_reference_known_answers():
    assert _search([["i", "j"], ["j", "k"], ["k", "i"]], [[2, 3], [3, 4], [4, 2]]) == \
        ("(2,(0,1))", 32.0)
    assert _search([["i", "j"], ["j", "k", "m"], ["k", "l"], ["l", "m", "n"]],
                   [[2, 10], [10, 4, 8], [4, 2], [2, 8, 7]]) == ("((2,3),(0,1))", 1536.0)
    assert _search([["x", "y"], ["y", "z"], ["z", "w"], ["w", "v"]],
                   [[10, 10], [10, 3], [3, 4], [4, 2]]) == ("(0,(1,(2,3)))", 284.0)
    assert _search([["i", "j"], ["k", "l"]], [[2, 3], [4, 5]]) == ("(0,1)", 120.0)


def test_search_chain_and_ipeps_contract_fully():
    labels = [["a", "b"], ["b", "c"], ["c", "d"], ["d", "e"], ["e", "f"]]
    shapes = [[2, 3], [3, 4], [4, 5], [5, 6], [6, 7]]
    tree, cost = search_order(labels, shapes)
    out, c2 = _tree_cost(tree, labels, shapes)
    assert sorted(out) == ["a", "f"] and c2 == cost
    net = [["1", "2"], ["34", "35"], ["41", "42"], ["19", "16"], ["1", "20", "3", "4"],
           ["20", "34", "21", "22"], ["35", "38", "36", "37"], ["38", "41", "39", "40"],
           ["42", "33", "31", "32"], ["33", "19", "17", "18"], ["16", "13", "14", "15"],
           ["13", "2", "5", "6"], ["4", "24", "10", "6", "8"], ["3", "23", "9", "5", "7"],
           ["22", "37", "27", "24", "25"], ["21", "36", "26", "23", "25"],
           ["27", "40", "32", "29", "30"], ["26", "39", "31", "28", "30"],
           ["10", "29", "18", "15", "12"], ["9", "28", "17", "14", "11"], ["7", "8", "11", "12"]]
    d2 = {"7", "8", "11", "12", "25", "30"}
    shapes = [[2 if l in d2 else 8 for l in ls] for ls in net]
    tree, cost = search_order(net, shapes)
    from cytnx_amd.network import leaves
    assert sorted(leaves(tree)) == list(range(21))
    out, c2 = _tree_cost(tree, net, shapes)
    assert out == []
    assert abs(c2 - cost) <= 1e-6 * c2  # single-precision accumulation


def test_search_limits():
    with pytest.raises(ValueError):
        search_order([["a", "b"]], [[2, 2]])
    with pytest.raises(ValueError):
        search_order([[str(i), str(i + 1)] for i in range(65)], [[2, 2]] * 65)


def _exhaustive(labels, shapes):
    n = len(labels)
    dims = {l: s for ls, ss in zip(labels, shapes) for l, s in zip(ls, ss)}
    best = {}
    for i in range(n):
        best[1 << i] = (0, tuple(labels[i]))
    for mask in range(1, 1 << n):
        if mask in best:
            continue
        cand = None
        sub = (mask - 1) & mask
        while sub:
            o = mask ^ sub
            if sub in best and o in best:
                ca, la = best[sub]
                cb, lb = best[o]
                common = [l for l in lb if l in la]
                c = int(np.prod([dims[l] for l in la])) * int(np.prod([dims[l] for l in lb]))
                c //= int(np.prod([dims[l] for l in common])) if common else 1
                lab = tuple([l for l in la if l not in common] + [l for l in lb if l not in common])
                if cand is None or c + ca + cb < cand[0]:
                    cand = (c + ca + cb, lab)
            sub = (sub - 1) & mask
        best[mask] = cand
    return best[(1 << n) - 1][0]


def test_search_costs_against_exhaustive_on_random_networks():
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(3, 7))
        # random connected network: a spanning path plus random extra edges, each a label
        edges = [(i, i + 1) for i in range(n - 1)]
        for a, b in itertools.combinations(range(n), 2):
            if b > a + 1 and rng.random() < 0.3:
                edges.append((a, b))
        labels = [[] for _ in range(n)]
        dims = {}
        for k, (a, b) in enumerate(edges):
            lab = "e%d" % k
            labels[a].append(lab)
            labels[b].append(lab)
            dims[lab] = int(rng.integers(2, 9))
        for i in range(n):
            if rng.random() < 0.5:
                lab = "o%d" % i
                labels[i].append(lab)
                dims[lab] = int(rng.integers(2, 9))
        shapes = [[dims[l] for l in ls] for ls in labels]
        tree, cost = search_order(labels, shapes)
        _, exact = _tree_cost(tree, labels, shapes)
        assert abs(exact - cost) <= 1e-6 * exact
        assert exact >= _exhaustive(labels, shapes)


def test_greedy_is_optimal_on_monotone_matrix_chain():
    # decreasing bond dimensions: contracting from the cheap end is optimal and greedy finds it
    labels = [["a", "b"], ["b", "c"], ["c", "d"], ["d", "e"]]
    shapes = [[64, 32], [32, 16], [16, 8], [8, 4]]
    tree, cost = search_order(labels, shapes)
    assert cost == _exhaustive(labels, shapes)


def test_parse_order_left_associative_and_grouping():
    n2p = {"A": 0, "B": 1, "C": 2, "D": 3}
    assert parse_order(tokenize_order("B,C,A"), n2p) == ((1, 2), 0)
    assert parse_order(tokenize_order("(A,(B,C))"), n2p) == (0, (1, 2))
    assert parse_order(tokenize_order(" ( (A , B) , (C,D) ) "), n2p) == ((0, 1), (2, 3))
    assert parse_order(tokenize_order("A,(B,C),D"), n2p) == ((0, (1, 2)), 3)
    for bad in ["(A,B", "A;B", "AB", "A:B"]:
        with pytest.raises(ValueError):
            parse_order(tokenize_order(bad), n2p)
    with pytest.raises(ValueError):
        parse_order(tokenize_order("(A,X)"), n2p)
    with pytest.raises(ValueError):
        parse_order(tokenize_order("(A,B))(,C"), n2p)


def test_parse_tout_rowrank():
    assert parse_tout("0,1;2,3") == (["0", "1", "2", "3"], 2)
    assert parse_tout("a,b,e") == (["a", "b", "e"], 1)
    assert parse_tout(";-2, 1") == (["-2", "1"], 0)
    assert parse_tout("a ;") == (["a"], 1)
    with pytest.raises(ValueError):
        parse_tout("a;b;c")
    with pytest.raises(ValueError):
        parse_tout("a,,b")


def test_fromstring_dmrg_network():
    net = Network()
    net.FromString(DMRG)
    assert net.names == ["psi", "L", "R", "M1", "M2"]
    assert net.label_arr[3] == ["-5", "-6", "-2", "1"]
    assert net.tout == ["0", "1", "2", "3"] and net.tout_rowrank == 2
    assert net.getOrder() == "No optimal or user specified order found."
    assert tree_str(net.contraction_tree(), net.names) == "((((psi,L),R),M1),M2)"
    net.setOrder(False, "(L,(psi,(M1,(M2,R))))")
    assert net.getOrder() == "(L,(psi,(M1,(M2,R))))"
    assert net.contraction_tree() == (1, (0, (3, (4, 2))))
    net.setOrder(False, "")
    assert net.contraction_tree() == ((((0, 1), 2), 3), 4)


def test_fromstring_comments_spaces_and_order_line():
    net = Network()
    net.FromString(["# comment", "A: a, b ,c  # trailing", "", "  B : c,d", "C: d,e",
                    "ORDER : B,C,A", "TOUT: a,b;e"])
    assert net.label_arr == [["a", "b", "c"], ["c", "d"], ["d", "e"]]
    assert net.getOrder() == "B,C,A"
    assert net.contraction_tree() == ((1, 2), 0)


def test_fromfile(tmp_path):
    p = tmp_path / "proj.net"
    p.write_text("\n".join(DMRG) + "\n")
    net = Network(str(p))
    assert net.names == ["psi", "L", "R", "M1", "M2"]


@pytest.mark.parametrize("lines", [
    ["A: a,b"],                                             # fewer than two tensors
    ["A: a,b", "A: b,c", "TOUT: a;c"],                      # duplicated name
    ["A: a,b", "B: b,c", "TOUT: a;d"],                      # TOUT does not match
    ["A: a,b", "B: b,c"],                                   # TOUT missing
    ["A: a,b", "B: b,c", "C: c,d", "ORDER: (A,B)", "TOUT: a;d"],   # C not in ORDER
    ["A: a,b", "B: b,c", "ORDER: (A,X)", "TOUT: a;c"],     # unknown name in ORDER
    ["A a,b", "B: b,c", "TOUT: a;c"],                       # no ':'
    ["A: a,b:c", "B: b,c", "TOUT: a;c"],                    # two ':'
    ["A B: a,b", "C: b,c", "TOUT: a;c"],                    # space in name
    ["A: ", "B: b,c", "TOUT: c"],                           # empty labels
    ["A: a,,b", "B: b,c", "TOUT: a;c"],                     # empty label
])
def test_fromstring_errors(lines):
    with pytest.raises(ValueError):
        Network().FromString(lines)


def test_put_and_launch_errors_without_tensors():
    net = Network()
    net.FromString(DMRG)
    with pytest.raises(ValueError):
        net.Launch()
    with pytest.raises(ValueError):
        net.getOptimalOrder()
    with pytest.raises(ValueError):
        net.PutUniTensor("X", None)
    with pytest.raises(ValueError):
        net.PutUniTensors(["psi"], [])


def test_optimal_order_of_dmrg_network_uses_declared_labels():
    class Shape:
        def __init__(self, s):
            self._s = s

        def shape(self):
            return self._s

        def rank(self):
            return len(self._s)

    chi, D, d = 64, 5, 2
    net = Network()
    net.FromString(DMRG)
    net.tensors = [Shape([chi, d, d, chi]), Shape([D, chi, chi]), Shape([D, chi, chi]),
                   Shape([D, D, d, d]), Shape([D, D, d, d])]
    line = net.getOptimalOrder()
    tree, cost = search_order(net.label_arr, [t.shape() for t in net.tensors])
    assert line == tree_str(tree, net.names)
    _, exact = _tree_cost(tree, net.label_arr, [t.shape() for t in net.tensors])
    assert exact == cost
    # the left-to-right order ((((psi,L),R),M1),M2) costs more than the searched one
    _, default = _tree_cost(((((0, 1), 2), 3), 4), net.label_arr,
                            [t.shape() for t in net.tensors])
    assert cost <= default


def _cpu_dense(shape, labels, name=""):
    from cytnx_amd.unitensor import Bond, UniTensor
    T = UniTensor([Bond(int(s)) for s in shape], labels, 1, 3, -1)
    return T.set_name_(name)


def test_contracts_validation():
    """Contracts' argument checks (src/UniTensor_base.cpp:717-746), before any data moves."""
    from cytnx_amd.network import Contracts
    from cytnx_amd.unitensor import Contract
    A = _cpu_dense([2, 3], ["a", "b"])
    B = _cpu_dense([3, 4], ["b", "c"])
    with pytest.raises(ValueError, match="more than 2 TNs"):
        Contracts([A])
    with pytest.raises(ValueError, match="more than 2 TNs"):
        Contract([A])
    # a given order with optimal=False names the tensors by their own names
    with pytest.raises(ValueError, match="missing tensor name at index 0"):
        Contracts([A, B], "(A,B)", False)
    B.set_name_("B")
    A.set_name_("A")
    with pytest.raises(ValueError):
        Contracts([A, B], "(A,X)", False)  # an order naming a tensor that is not in the list
    with pytest.raises(TypeError):
        Contract(A, B, False, False, True)
    with pytest.raises(TypeError):
        Contract(A, B, cache=True)


def test_contracts_optimal_with_order_warns():
    """optimal=True with an order: the order is dropped with the reference's warning."""
    from cytnx_amd import network
    A = _cpu_dense([2, 3], ["a", "b"])
    B = _cpu_dense([3, 4], ["b", "c"])
    seen = {}

    def fake_launch(self):
        seen["tout"], seen["rr"], seen["order"] = list(self.tout), self.tout_rowrank, \
            self.order_line
        return None

    orig = network.Network.Launch
    network.Network.Launch = fake_launch
    try:
        with pytest.warns(UserWarning, match="Setting Optimal = true"):
            network.Contracts([B, A], "(T1,T0)", True)
    finally:
        network.Network.Launch = orig
    # sorted uncontracted labels, rowrank 1, the searched order over T0, T1
    assert seen["tout"] == ["a", "c"] and seen["rr"] == 1
    assert seen["order"] == "(T0,T1)"
