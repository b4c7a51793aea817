"""Label-network contraction: the reference's RegularNetwork (src/RegularNetwork.cpp).

Network file / string syntax (RegularNetwork::FromString, src/RegularNetwork.cpp:381-590):

    name: l1, l2, ...          one line per tensor (labels of its bonds, in order)
    TOUT: k1, k2; b1, b2       output labels; the part before ';' is the row space (rowrank),
                               without ';' the first half of the labels
    ORDER: ((A,B),C)           optional contraction order; ',' pairs left-associatively,
                               parentheses group
    # ...                      comments (whole line or end of line)

At least two tensors; the labels that appear exactly once must be the TOUT labels.  Launch
relabels every tensor to its declared labels, contracts pairwise in post order of the
contraction tree (the ORDER tree, or left to right over the declaration order by default,
src/contraction_tree.cpp:6-26) with Contract, then permutes the result to TOUT with rowrank =
the number of row labels (RegularNetwork::Launch, src/RegularNetwork.cpp:901-1007).  Each
pairwise Contract is the block-sparse MFMA contraction (or the dense Tensordot) on the GPU.

setOrder(optimal=True) runs the reference's order search (OptimalTreeSolver::solve,
src/search_tree.cpp:120-290): per connected component, repeatedly contract the adjacent pair of
live nodes with the smallest accumulated cost, where the pair (t1, t2) costs
prod(shape1) * prod(shape2) / prod(contracted dims) + cost(t1) + cost(t2) in single precision
(get_cost, src/search_tree.cpp:9-28) and ties go to the lower node index (leaves 0..n-1, merged
nodes numbered on from n in creation order); a pair formed by a merge puts the older node on
the left.  Disconnected components are joined left to right at the end.  The search runs over
the network's declared labels with the shapes of the tensors put in.
"""
from __future__ import annotations

import heapq
import re
import warnings
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from .unitensor import Contract, UniTensor

Tree = Union[int, Tuple["Tree", "Tree"]]

_NO_ORDER = "No optimal or user specified order found."


def _split_labels(text: str, what: str, lineno: int) -> List[str]:
    parts = [p.strip() for p in text.split(",")]
    if len(parts) == 1 and parts[0] == "":
        return []
    for p in parts:
        if not p:
            raise ValueError("[ERROR][Network][Fromfile] line:%d Invalid labels for %s line."
                             % (lineno, what))
    return parts


def parse_tout(text: str, lineno: int = 0) -> Tuple[List[str], int]:
    """(labels, rowrank) of a TOUT line (ParseToutLineInternal, src/RegularNetwork.cpp:40-99)."""
    halves = text.split(";")
    if len(halves) > 2:
        raise ValueError("[ERROR][Network] line:%d Invalid TOUT line." % lineno)
    if len(halves) == 1:
        labels = _split_labels(text, "TOUT", lineno)
        return labels, len(labels) // 2
    ket = _split_labels(halves[0], "TOUT", lineno)
    bra = _split_labels(halves[1], "TOUT", lineno)
    return ket + bra, len(ket)


def tokenize_order(text: str, lineno: int = 0) -> List[str]:
    """Names, '(' , ')' and ',' of an ORDER line (ParseOrderLineInternal,
    src/RegularNetwork.cpp:16-38)."""
    if any(c in text for c in "\t;\n:"):
        raise ValueError("[ERROR][Network][Fromfile] line:%d invalid ORDER line format." % lineno)
    if not any(c in text for c in "(),"):
        raise ValueError("[ERROR][Network][Fromfile] line:%d invalid ORDER line format. tensors "
                         "should be seperate by delimiter ',' (comma), and/or wrapped with '(' "
                         "and ')'" % lineno)
    if text.count("(") != text.count(")"):
        raise ValueError("[ERROR][Network][Fromfile] parentheses mismatch.")
    toks = [t.strip() for t in re.split(r"([(),])", text)]
    return [t for t in toks if t]


def parse_order(tokens: Sequence[str], name2pos: Dict[str, int]) -> Tree:
    """Contraction tree of ORDER tokens: expr := term (',' term)* (left-associative),
    term := name | '(' expr ')' (ContractionTree::build_contraction_tree_by_tokens,
    src/contraction_tree.cpp:28-123)."""
    pos = 0

    def term() -> Tree:
        nonlocal pos
        if pos >= len(tokens):
            raise ValueError("[ERROR][Network] incomplete ORDER line")
        t = tokens[pos]
        if t == "(":
            pos += 1
            e = expr()
            if pos >= len(tokens) or tokens[pos] != ")":
                raise ValueError("[ERROR][Network][Fromfile] parentheses mismatch.")
            pos += 1
            return e
        if t in (")", ","):
            raise ValueError("[ERROR][Network] invalid ORDER line near '%s'" % t)
        if t not in name2pos:
            raise ValueError("[ERROR][ContractionTree][build_contraction_order_by_token] tokens "
                             "contain invalid TN name: %s ,which is not previously defined." % t)
        pos += 1
        return name2pos[t]

    def expr() -> Tree:
        nonlocal pos
        left = term()
        while pos < len(tokens) and tokens[pos] == ",":
            pos += 1
            left = (left, term())
        return left

    tree = expr()
    if pos != len(tokens):
        raise ValueError("[ERROR][Network] invalid ORDER line near '%s'" % tokens[pos])
    return tree


def leaves(tree: Tree) -> List[int]:
    return [tree] if isinstance(tree, int) else leaves(tree[0]) + leaves(tree[1])


def tree_str(tree: Tree, names: Sequence[str]) -> str:
    if isinstance(tree, int):
        return str(names[tree])
    return "(%s,%s)" % (tree_str(tree[0], names), tree_str(tree[1], names))


def _check_order_names(tree: Tree, names: Sequence[str]) -> None:
    used = leaves(tree)
    for i, n in enumerate(names):
        if i not in used:
            raise ValueError("[ERROR][Network][Fromfile] TN: <%s> defined but is not used in "
                             "ORDER line" % n)
    if len(used) != len(set(used)):
        raise ValueError("[ERROR] a TN appears more than once in the ORDER line")


class _Node:
    __slots__ = ("labels", "shape", "cost", "tree")

    def __init__(self, labels, shape, cost, tree):
        self.labels, self.shape, self.cost, self.tree = labels, shape, cost, tree


def _pair_cost(a: _Node, b: _Node) -> np.float32:
    """get_cost in single precision (src/search_tree.cpp:9-28)."""
    c = np.float32(1.0)
    for s in a.shape:
        c = np.float32(c * np.float32(s))
    for s in b.shape:
        c = np.float32(c * np.float32(s))
    for lab, s in zip(b.labels, b.shape):
        if lab in a.labels:
            c = np.float32(c / np.float32(s))
    return np.float32(np.float32(c + a.cost) + b.cost)


def _merge(a: _Node, b: _Node) -> _Node:
    common = set(a.labels) & set(b.labels)
    labs = [l for l in a.labels if l not in common] + [l for l in b.labels if l not in common]
    shp = [s for l, s in zip(a.labels, a.shape) if l not in common] + \
        [s for l, s in zip(b.labels, b.shape) if l not in common]
    return _Node(labs, shp, _pair_cost(a, b), (a.tree, b.tree))


def search_order(labels: Sequence[Sequence[str]], shapes: Sequence[Sequence[int]]
                 ) -> Tuple[Tree, float]:
    """(tree, accumulated cost) of the greedy pairwise order (module docstring)."""
    n = len(labels)
    if n < 2:
        raise ValueError("[ERROR][SearchTree] need at least 2 nodes.")
    if n > 64:
        raise ValueError("[ERROR][OptimalTreeSolver] solve() supports at most 64 tensors, got %d."
                         % n)
    nodes = [_Node(list(labels[i]), [int(s) for s in shapes[i]], np.float32(0.0), i)
             for i in range(n)]
    adj = [set() for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if set(labels[i]) & set(labels[j]):
                adj[i].add(j)
                adj[j].add(i)
    # connected components in order of their lowest leaf
    seen = [False] * n
    comps = []
    for s in range(n):
        if seen[s]:
            continue
        comp, stack = [], [s]
        while stack:
            v = stack.pop()
            if seen[v]:
                continue
            seen[v] = True
            comp.append(v)
            stack.extend(sorted(adj[v], reverse=True))
        comps.append(comp)
    results = []
    for comp in comps:
        alive = {i: nodes[i] for i in comp}
        heap = []
        cs = sorted(comp)
        for x in range(len(cs)):
            for y in range(x + 1, len(cs)):
                if cs[y] in adj[cs[x]]:
                    heapq.heappush(heap, (float(_pair_cost(alive[cs[x]], alive[cs[y]])),
                                          cs[x], cs[y]))
        while len(alive) > 1:
            _, i, j = heapq.heappop(heap)
            if i not in alive or j not in alive:
                continue
            a, b = alive.pop(i), alive.pop(j)
            m = len(nodes)
            merged = _merge(a, b)
            nodes.append(merged)
            adj.append(adj[i] | adj[j])
            for k in sorted(alive):
                if k in adj[m]:
                    adj[k].add(m)
                    heapq.heappush(heap, (float(_pair_cost(merged, alive[k])), k, m))
            alive[m] = merged
        results.append(next(iter(alive.values())))
    root = results[0]
    for other in results[1:]:
        root = _Node(root.labels + other.labels, root.shape + other.shape,
                     _pair_cost(root, other), (root.tree, other.tree))
    return root.tree, float(root.cost)


class Network:
    """RegularNetwork: FromString / Fromfile / Contract, PutUniTensor(s), RmUniTensor(s),
    setOrder, getOrder, getOptimalOrder, Launch."""

    def __init__(self, fname: Optional[str] = None):
        self.clear()
        if fname is not None:
            self.Fromfile(fname)

    def clear(self) -> None:
        self.names: List[str] = []
        self.label_arr: List[List[str]] = []
        self.name2pos: Dict[str, int] = {}
        self.tout: List[str] = []
        self.tout_rowrank = 0
        self.order: Optional[Tree] = None
        self.order_line = ""
        self.tensors: List[Optional[UniTensor]] = []

    # ---------------------------------------------------------------- construction
    def _register(self, name: str, labels: List[str]) -> None:
        if name in self.name2pos:
            raise ValueError("[ERROR][Network][Fromfile] tensor name: [%s] has already exist. "
                             "Cannot have duplicated tensor name in a network." % name)
        self.name2pos[name] = len(self.names)
        self.names.append(name)
        self.label_arr.append(labels)

    def _finish(self, order_tokens: Optional[List[str]], fill_tout: bool) -> None:
        if len(self.names) < 2:
            raise ValueError("[ERROR][Network][Fromfile] invalid network file. Should have at "
                             "least 2 tensors defined.")
        counts: Dict[str, int] = {}
        for labels in self.label_arr:
            for lab in labels:
                counts[lab] = counts.get(lab, 0) + 1
        expected = sorted(l for l, c in counts.items() if c == 1)
        if fill_tout and not self.tout:
            self.tout = expected
            self.tout_rowrank = 0
        if len(self.tout) != len(expected) or set(self.tout) != set(expected):
            raise ValueError("[ERROR][Network][Fromfile] The TOUT contains labels that does not "
                             "match with the delcartion from TNs. The reduced labels [rank:%d] "
                             "should be: %s; the TOUT [rank%d] specified is: %s"
                             % (len(expected), " ".join(expected), len(self.tout),
                                " ".join(self.tout)))
        if order_tokens is not None:
            tree = parse_order(order_tokens, self.name2pos)
            _check_order_names(tree, self.names)
            self.order = tree
        self.tensors = [None] * len(self.names)

    def FromString(self, contents: Sequence[str]) -> None:
        self.clear()
        order_tokens = None
        for i, raw in enumerate(contents):
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            line = line.split("#", 1)[0]
            if ":" not in line:
                raise ValueError("[ERROR][Network][Fromfile] invalid network description at "
                                 "line: %d. should contain ':'" % i)
            parts = line.split(":")
            if len(parts) != 2:
                raise ValueError("[ERROR][Network] invalid line in network file at line: %d." % i)
            name, content = parts[0].strip(), parts[1].strip()
            if not name:
                raise ValueError("[ERROR][Network][Fromfile] invalid tensor name at line: %d" % i)
            if any(c in name for c in " ;,\t\n"):
                raise ValueError("[ERROR] invalid Tensor name at line %d" % i)
            if name == "ORDER":
                if content:
                    self.order_line = content
                    order_tokens = tokenize_order(content, i)
            elif name == "TOUT":
                if content:
                    self.tout, self.tout_rowrank = parse_tout(content, i)
            else:
                if not content:
                    raise ValueError("[ERROR][Network][Fromfile] invalid tensor labelsat line %d. "
                                     "cannot have empty labels for input tensor." % i)
                labels = _split_labels(content, "TN", i)
                if not labels:
                    raise ValueError("[ERROR][Network][Fromfile] line:%d Invalid TN line." % i)
                self._register(name, labels)
        self._finish(order_tokens, fill_tout=False)

    def Fromfile(self, fname: str) -> None:
        with open(fname) as f:
            self.FromString(f.read().splitlines())

    @staticmethod
    def Contract(tensors: Sequence[UniTensor], Tout: str, alias: Sequence[str] = (),
                 contract_order: str = "") -> "Network":
        """Network over the tensors' own labels (Network::Contract / Contract_plan,
        include/Network.hpp:366-374, src/RegularNetwork.cpp:237-380); default names are
        '<name>_usr' for named tensors and 'uname_T<i>' otherwise."""
        if len(tensors) < 2:
            raise ValueError("[ERROR][Network] invalid network. Should have at least 2 tensors "
                             "defined.")
        if contract_order and not alias:
            raise ValueError("[ERRPR] conraction_order need to be specify using alias name, so "
                             "alias name have to be assigned!")
        if alias and len(alias) != len(tensors):
            raise ValueError("[ERROR] alias of UniTensor need to be assigned for all utensors.")
        net = Network()
        if Tout:
            net.tout, net.tout_rowrank = parse_tout(Tout)
        for i, T in enumerate(tensors):
            if alias:
                name = alias[i]
            else:
                nm = getattr(T, "_name", "")
                name = nm + "_usr" if nm else "uname_T%d" % i
            net._register(name, T.labels())
        order_tokens = None
        if contract_order:
            net.order_line = contract_order
            order_tokens = tokenize_order(contract_order)
        net._finish(order_tokens, fill_tout=True)
        for i, T in enumerate(tensors):
            net.tensors[i] = T
        return net

    def _index(self, name: Union[str, int], who: str) -> int:
        if isinstance(name, (int, np.integer)):
            if not 0 <= int(name) < len(self.names):
                raise ValueError("[ERROR][RegularNetwork][%s] index=%d out of range." % (who, name))
            return int(name)
        if name not in self.name2pos:
            raise ValueError("[ERROR][RegularNetwork][%s] Cannot find the tensor name: [%s] in "
                             "current network." % (who, name))
        return self.name2pos[name]

    def PutUniTensor(self, name: Union[str, int], utensor: UniTensor,
                     label_order: Sequence[str] = ()) -> None:
        idx = self._index(name, "PutUniTensor")
        if label_order:
            utensor = utensor.permute(list(label_order))
        if utensor.rank() != len(self.label_arr[idx]):
            raise ValueError("[ERROR][RegularNetwork][PutUniTensor] tensor name: [%s], the rank "
                             "of input UniTensor does not match the definition in network file."
                             % self.names[idx])
        self.tensors[idx] = utensor

    def PutUniTensors(self, names: Sequence[str], utensors: Sequence[UniTensor]) -> None:
        if len(names) != len(utensors):
            raise ValueError("[ERROR][RegularNetwork][PutUniTensors] total number of names does "
                             "not match number of input UniTensors.")
        for n, t in zip(names, utensors):
            self.PutUniTensor(n, t)

    def RmUniTensor(self, name: Union[str, int]) -> None:
        self.tensors[self._index(name, "RmUniTensor")] = None

    def RmUniTensors(self, names: Sequence[str]) -> None:
        for n in names:
            self.RmUniTensor(n)

    # ---------------------------------------------------------------- order
    def _default_tree(self) -> Tree:
        tree: Tree = 0
        for i in range(1, len(self.names)):
            tree = (tree, i)
        return tree

    def _check_set(self) -> None:
        for i, t in enumerate(self.tensors):
            if t is None:
                raise ValueError("[ERROR][Network] tensor [%s] is not set (PutUniTensor)"
                                 % self.names[i])

    def getOptimalOrder(self) -> str:
        self._check_set()
        tree, _ = search_order(self.label_arr, [t.shape() for t in self.tensors])
        return tree_str(tree, self.names)

    def setOrder(self, optimal: bool = False, contract_order: str = "") -> None:
        if optimal and contract_order:
            warnings.warn("[WARNING][setOrder][RegularNetwork] Setting Optimal = true while "
                          "specifying the order, will find the optimal order instead.")
        if optimal:
            line = self.getOptimalOrder()
            self.order_line = line
            self.order = parse_order(tokenize_order(line, 999999), self.name2pos)
        elif contract_order:
            tree = parse_order(tokenize_order(contract_order, 999999), self.name2pos)
            _check_order_names(tree, self.names)
            self.order_line = contract_order
            self.order = tree
        else:
            self.order_line = ""
            self.order = None

    def getOrder(self) -> str:
        return self.order_line if self.order_line else _NO_ORDER

    def contraction_tree(self) -> Tree:
        """The tree Launch follows: the ORDER / setOrder tree, else left to right."""
        return self.order if self.order is not None else self._default_tree()

    # ---------------------------------------------------------------- launch
    def Launch(self) -> UniTensor:
        self._check_set()
        base = [t.relabel(labs) for t, labs in zip(self.tensors, self.label_arr)]

        def run(node: Tree) -> UniTensor:
            if isinstance(node, int):
                return base[node]
            return Contract(run(node[0]), run(node[1]))

        out = run(self.contraction_tree())
        if self.tout:
            out = out.permute(self.tout, self.tout_rowrank)
        return out

    def __repr__(self) -> str:
        lines = ["==== Network ===="]
        for i, n in enumerate(self.names):
            st = "o" if self.tensors and self.tensors[i] is not None else "x"
            lines.append("[%s] %s : %s" % (st, n, " ".join(self.label_arr[i])))
        lines.append("TOUT : %s ; %s" % (" ".join(self.tout[:self.tout_rowrank]),
                                         " ".join(self.tout[self.tout_rowrank:])))
        lines.append("ORDER : %s" % self.order_line)
        lines.append("=================")
        return "\n".join(lines)

    PrintNet = __repr__


def Contracts(TNs: Sequence[UniTensor], order: str = "", optimal: bool = True) -> UniTensor:
    """Contract a list of UniTensors over their shared labels (Contracts and the list form of
    Contract, src/UniTensor_base.cpp:711-756, pybind/unitensor_py.cpp:2661-2671).

    A network over the tensors' own labels: named T0, T1, ... when the order is searched
    (optimal=True, which overrides a given order with a warning) or when no order is given;
    with optimal=False and an order, the tensors' names (UniTensor.name()) are the names the
    order refers to.  The output's labels are the uncontracted labels in sorted order with
    rowrank 1 (the reference builds the network with no TOUT and output rowrank -1, so the
    filled-in TOUT keeps one row label); a full contraction gives a rank-0 tensor."""
    TNs = list(TNs)
    if len(TNs) < 2:
        raise ValueError("[ERROR][Contracts] should have more than 2 TNs to contract.")
    orderin = order
    if optimal or not order:
        if order:
            warnings.warn("[WARNING][Contracts] Setting Optimal = true while specifying the "
                          "order, will find the optimal order instead. To use the desired order "
                          "please set Optimal = false.")
            orderin = ""
        names = ["T%d" % i for i in range(len(TNs))]
    else:
        names = []
        for i, T in enumerate(TNs):
            if T.name() == "":
                raise ValueError("[ERROR][Contracts] Unable to contract with the specified "
                                 "order, missing tensor name at index %d." % i)
            names.append(T.name())
    net = Network()
    for name, T in zip(names, TNs):
        net._register(name, T.labels())
    order_tokens = None
    if orderin:
        net.order_line = orderin
        order_tokens = tokenize_order(orderin)
    net._finish(order_tokens, fill_tout=True)
    net.tout_rowrank = 1 if net.tout else 0
    net.PutUniTensors(names, TNs)
    net.setOrder(optimal, orderin)
    return net.Launch()


__all__ = ["Network", "Contracts", "search_order", "parse_order", "tokenize_order",
           "parse_tout"]
