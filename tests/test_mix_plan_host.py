"""Host checks of the two-site mix plan (cytnx_amd.heff.mix_plan_arrays).

The task arrays are executed by a numpy restatement of the mix launch's semantics (slices are
strided matrix products, outputs their weighted sums) and compared with the dense closed form
Y[l', s1', s2', w'', r] = sum W12[w, s1, s1', w'', s2, s2'] L[w, l, l'] psi[l, s1, s2, r],
restricted to Y's blocks.  A wrong stride, offset, coefficient index or slice/output pairing
changes Y at 1e-14 relative or more."""
import numpy as np
import pytest
import torch

from cytnx_amd import heisenberg
from cytnx_amd.heff import (MIX_SLICES, check_mix_bounds, mix_plan_arrays, psi_is_left,
                             subset_blocks)
from cytnx_amd.unitensor import UniTensor


def _meta(bonds, labels):
    return UniTensor(bonds, labels, 1, 3, -1, _alloc=False)


def _fill_from_dense(T: UniTensor, D: np.ndarray) -> np.ndarray:
    flat = np.zeros(int(sum(np.prod(s.astype(np.int64)) for s in T._shapes)))
    offs = [np.concatenate([[0], np.cumsum(b.getDegeneracies())[:-1]]).astype(int)
            for b in T.bonds()]
    for bi, t in enumerate(T._itoi):
        shp = [int(s) for s in T._shapes[bi]]
        sl = tuple(slice(offs[a][int(q)], offs[a][int(q)] + shp[a]) for a, q in enumerate(t))
        o = int(T._offsets[bi])
        flat[o:o + int(np.prod(shp))] = D[sl].ravel()
    return flat


def _run_mix(arrays, A, B, ny):
    items, slices, outs, coef = arrays
    Y = np.full(ny, np.nan)
    s_at = o_at = 0
    for M, N, ns, no in items:
        X = []
        for s in range(s_at, s_at + ns):
            a_off, b_off, sam, sak, sbk, sbn, K = (int(v) for v in slices[s])
            Am = np.array([[A[a_off + m * sam + k * sak] for k in range(K)] for m in range(M)])
            Bm = np.array([[B[b_off + k * sbk + n * sbn] for n in range(N)] for k in range(K)])
            X.append(Am @ Bm if K else np.zeros((M, N)))
        for o in range(o_at, o_at + no):
            c_off, ldc = (int(v) for v in outs[o])
            val = np.zeros((M, N))
            for s in range(ns):
                val += coef[o, s] * X[s]
            for m in range(M):
                Y[c_off + m * ldc:c_off + m * ldc + N] = val[m]
        s_at += ns
        o_at += no
    assert s_at == len(slices) and o_at == len(outs)
    return Y


@pytest.mark.parametrize("chi", [8, 20])
def test_mix_plan_matches_dense_closed_form(chi):
    psi, L, M1, M2, R = heisenberg.two_site_problem(chi, seed=chi, hermitian=True)
    # L permuted to (l, w, l'), laid out afresh
    Lb = L.bonds()
    Lp = _meta([Lb[1], Lb[0], Lb[2]], ["-1", "-5", "0"])
    Ld = L.to_dense()                                   # (w, l, l')
    A = _fill_from_dense(Lp, np.transpose(Ld, (1, 0, 2)))
    B = psi.storage().numpy()
    W = np.einsum("awsx,wvty->asxvty", M1.to_dense(), M2.to_dense())  # (w,s1,s1',w'',s2,s2')
    Yb = [Lb[2], M1.bonds()[3], M2.bonds()[3], M2.bonds()[1], psi.bonds()[3]]
    Y = _meta(Yb, ["0", "1", "2", "k0", "k1"])
    arrays = mix_plan_arrays(Lp, {"l": "-1", "w": "-5", "lp": "0"},
                             psi, {"l": "-1", "s1": "-2", "s2": "-3", "r": "-4"}, W,
                             Y, {"lp": "0", "s1p": "1", "s2p": "2", "w2": "k0", "r": "k1"})
    items, slices, outs, coef = arrays
    assert items[:, 2].max() <= MIX_SLICES
    ny = int(sum(np.prod(s.astype(np.int64)) for s in Y._shapes))
    got = _run_mix(arrays, A, B, ny)
    assert not np.isnan(got).any()  # every element of Y written
    Pd = psi.to_dense()
    dense = np.einsum("asxvty,alp,lstr->pxyvr", W, Ld, Pd, optimize=True)
    want = _fill_from_dense(Y, dense)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)


@pytest.mark.parametrize("chi", [8, 20])
def test_mix_plan_pruned_layout(chi):
    """Y blocks no slice reaches are zero in the dense closed form; the plan on the pruned
    layout (those blocks left out, the rest packed) still reproduces every kept block."""
    psi, L, M1, M2, R = heisenberg.two_site_problem(chi, seed=chi + 1, hermitian=True)
    Lb = L.bonds()
    Lp = _meta([Lb[1], Lb[0], Lb[2]], ["-1", "-5", "0"])
    Ld = L.to_dense()
    A = _fill_from_dense(Lp, np.transpose(Ld, (1, 0, 2)))
    B = psi.storage().numpy()
    W = np.einsum("awsx,wvty->asxvty", M1.to_dense(), M2.to_dense())
    Y = _meta([Lb[2], M1.bonds()[3], M2.bonds()[3], M2.bonds()[1], psi.bonds()[3]],
              ["0", "1", "2", "k0", "k1"])
    labs = ({"l": "-1", "w": "-5", "lp": "0"}, {"l": "-1", "s1": "-2", "s2": "-3", "r": "-4"},
            {"lp": "0", "s1p": "1", "s2p": "2", "w2": "k0", "r": "k1"})
    live = np.zeros(Y.Nblocks(), dtype=bool)
    mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Y, labs[2], live)
    assert 0 < live.sum() < Y.Nblocks()
    dense = np.einsum("asxvty,alp,lstr->pxyvr", W, Ld, psi.to_dense(), optimize=True)
    full = _fill_from_dense(Y, dense)
    for bi in np.nonzero(~live)[0]:
        o = int(Y._offsets[bi])
        assert not np.any(full[o:o + int(np.prod(Y._shapes[bi].astype(np.int64)))])
    Yl = subset_blocks(Y, live)
    arrays = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Yl, labs[2])
    ny = int(sum(np.prod(s.astype(np.int64)) for s in Yl._shapes))
    check_mix_bounds(arrays, A.size, B.size, ny)
    got = _run_mix(arrays, A, B, ny)
    assert not np.isnan(got).any()
    want = _fill_from_dense(Yl, dense)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    assert arrays[0][:, 2].min() >= 1  # every kept item has a slice


def test_mix_plan_xxz_item_bounds():
    """For the XXZ MPO the items hold at most 8 slices (8 = 6 from the three degenerate MPO
    states of charge 0 times two spin pairs, plus one each from charges +2 and -2)."""
    psi, L, M1, M2, R = heisenberg.two_site_problem(48, seed=2, hermitian=True)
    Lb = L.bonds()
    Lp = _meta([Lb[1], Lb[0], Lb[2]], ["-1", "-5", "0"])
    W = np.einsum("awsx,wvty->asxvty", M1.to_dense(), M2.to_dense())
    Y = _meta([Lb[2], M1.bonds()[3], M2.bonds()[3], M2.bonds()[1], psi.bonds()[3]],
              ["0", "1", "2", "k0", "k1"])
    items, slices, outs, coef = mix_plan_arrays(
        Lp, {"l": "-1", "w": "-5", "lp": "0"},
        psi, {"l": "-1", "s1": "-2", "s2": "-3", "r": "-4"}, W,
        Y, {"lp": "0", "s1p": "1", "s2p": "2", "w2": "k0", "r": "k1"})
    assert items[:, 2].max() == 8
    # the weights are W12's 30 nonzero elements, each used by every item it reaches
    assert np.count_nonzero(W) == 30
    assert set(np.unique(coef[coef != 0]).round(12)) <= set(np.unique(W[W != 0]).round(12))


@pytest.mark.parametrize("chi", [8, 20])
def test_mix_plan_transposed_layout(chi):
    """Y laid out (w'', r, l', s1', s2'): l' is contiguous, the items compute the transposed
    tiles X[r, l'] with psi as the left operand, and Y still equals the dense closed form."""
    psi, L, M1, M2, R = heisenberg.two_site_problem(chi, seed=chi + 7, hermitian=True)
    Lb = L.bonds()
    Lp = _meta([Lb[1], Lb[0], Lb[2]], ["-1", "-5", "0"])
    Ld = L.to_dense()
    Lflat = _fill_from_dense(Lp, np.transpose(Ld, (1, 0, 2)))
    P = psi.storage().numpy()
    W = np.einsum("awsx,wvty->asxvty", M1.to_dense(), M2.to_dense())
    Y = _meta([M2.bonds()[1], psi.bonds()[3], Lb[2], M1.bonds()[3], M2.bonds()[3]],
              ["k0", "k1", "0", "1", "2"])
    assert np.all(Y._shapes[:, 3] * Y._shapes[:, 4] == 1)
    labY = {"lp": "0", "s1p": "1", "s2p": "2", "w2": "k0", "r": "k1"}
    assert psi_is_left(Y, labY)
    labs = ({"l": "-1", "w": "-5", "lp": "0"}, {"l": "-1", "s1": "-2", "s2": "-3", "r": "-4"})
    live = np.zeros(Y.Nblocks(), dtype=bool)
    mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Y, labY, live)
    Yl = subset_blocks(Y, live)
    arrays = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Yl, labY)
    ny = int(sum(np.prod(s.astype(np.int64)) for s in Yl._shapes))
    check_mix_bounds(arrays, P.size, Lflat.size, ny)  # A = psi, B = Lp
    with pytest.raises(ValueError):  # B (Lp) reads reach past a one-element buffer
        check_mix_bounds(arrays, P.size, 1, ny)
    got = _run_mix(arrays, P, Lflat, ny)
    assert not np.isnan(got).any()
    dense = np.einsum("asxvty,alp,lstr->vrpxy", W, Ld, psi.to_dense(), optimize=True)
    want = _fill_from_dense(Yl, dense)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    # the same items as the r-contiguous layout, with M and N exchanged
    Yr = _meta([Lb[2], M1.bonds()[3], M2.bonds()[3], M2.bonds()[1], psi.bonds()[3]],
               ["0", "1", "2", "k0", "k1"])
    assert not psi_is_left(Yr, labY)
    live_r = np.zeros(Yr.Nblocks(), dtype=bool)
    mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Yr, labY, live_r)
    items_r = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, subset_blocks(Yr, live_r), labY)[0]
    assert sorted(map(tuple, items_r[:, [1, 0, 2, 3]])) == sorted(map(tuple, arrays[0]))


@pytest.mark.parametrize("chi", [8, 20])
@pytest.mark.parametrize("transposed", [False, True])
def test_mix_plan_single_pass_equals_two_passes(chi, transposed):
    """mix_plan_arrays(..., read=mask) plans the pruned layout in one pass: the arrays and the
    kept-block mask equal those of the two-pass form (live mask, subset_blocks, plan again)."""
    psi, L, M1, M2, R = heisenberg.two_site_problem(chi, seed=chi + 3, hermitian=True)
    Lb = L.bonds()
    Lp = _meta([Lb[1], Lb[0], Lb[2]], ["-1", "-5", "0"])
    W = np.einsum("awsx,wvty->asxvty", M1.to_dense(), M2.to_dense())
    if transposed:
        Y = _meta([M2.bonds()[1], psi.bonds()[3], Lb[2], M1.bonds()[3], M2.bonds()[3]],
                  ["k0", "k1", "0", "1", "2"])
    else:
        Y = _meta([Lb[2], M1.bonds()[3], M2.bonds()[3], M2.bonds()[1], psi.bonds()[3]],
                  ["0", "1", "2", "k0", "k1"])
    labs = ({"l": "-1", "w": "-5", "lp": "0"}, {"l": "-1", "s1": "-2", "s2": "-3", "r": "-4"},
            {"lp": "0", "s1p": "1", "s2p": "2", "w2": "k0", "r": "k1"})
    rng = np.random.default_rng(chi)
    for read in (np.ones(Y.Nblocks(), dtype=bool), rng.random(Y.Nblocks()) < 0.8):
        live = np.zeros(Y.Nblocks(), dtype=bool)
        mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Y, labs[2], live)
        two = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, subset_blocks(Y, live & read),
                              labs[2])
        keep = np.zeros(Y.Nblocks(), dtype=bool)
        one = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Y, labs[2], keep, read)
        assert np.array_equal(keep, live & read)
        for a, b in zip(one, two):
            assert a.shape == b.shape and np.array_equal(a, b)


def _mix_args(chi, seed, transposed):
    psi, L, M1, M2, R = heisenberg.two_site_problem(chi, seed=seed, hermitian=True)
    Lb = L.bonds()
    Lp = _meta([Lb[1], Lb[0], Lb[2]], ["-1", "-5", "0"])
    W = np.einsum("awsx,wvty->asxvty", M1.to_dense(), M2.to_dense())
    yb = [Lb[2], M1.bonds()[3], M2.bonds()[3], M2.bonds()[1], psi.bonds()[3]]
    yl = ["0", "1", "2", "k0", "k1"]
    if transposed:  # (w'', r, l', s1', s2'): l' contiguous
        yb = [yb[3], yb[4], yb[0], yb[1], yb[2]]
        yl = [yl[3], yl[4], yl[0], yl[1], yl[2]]
    Y = _meta(yb, yl)
    labs = ({"l": "-1", "w": "-5", "lp": "0"}, {"l": "-1", "s1": "-2", "s2": "-3", "r": "-4"},
            {"lp": "0", "s1p": "1", "s2p": "2", "w2": "k0", "r": "k1"})
    return Lp, psi, W, Y, labs


@pytest.mark.parametrize("transposed", [False, True])
@pytest.mark.parametrize("chi,seed", [(8, 3), (20, 4), (64, 5), (160, 6)])
def test_native_mix_plan_equals_python_plan(chi, seed, transposed):
    """The library's C++ planner (cxa_heff2_mix_plan, what Heff2 uses) gives the same task
    arrays as the Python restatement above, element for element, with and without the read
    mask (the pruned layout) and for both Y leg orders."""
    from cytnx_amd.heff import mix_plan_arrays_native
    Lp, psi, W, Y, labs = _mix_args(chi, seed, transposed)
    live_py = np.zeros(Y.Nblocks(), dtype=bool)
    live_nt = np.zeros(Y.Nblocks(), dtype=bool)
    a_py = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Y, labs[2], live_py)
    a_nt = mix_plan_arrays_native(Lp, labs[0], psi, labs[1], W, Y, labs[2], live_nt)
    for x, y in zip(a_py, a_nt):
        assert x.shape == y.shape and np.array_equal(x, y)
    assert np.array_equal(live_py, live_nt)
    # with a read mask: every other reached block read
    read = live_py.copy()
    read[np.nonzero(read)[0][::2]] = False
    k_py = np.zeros(Y.Nblocks(), dtype=bool)
    k_nt = np.zeros(Y.Nblocks(), dtype=bool)
    b_py = mix_plan_arrays(Lp, labs[0], psi, labs[1], W, Y, labs[2], k_py, read)
    b_nt = mix_plan_arrays_native(Lp, labs[0], psi, labs[1], W, Y, labs[2], k_nt, read)
    for x, y in zip(b_py, b_nt):
        assert x.shape == y.shape and np.array_equal(x, y)
    assert np.array_equal(k_py, k_nt)


def test_native_mix_plan_rejects_more_than_eight_slices():
    """An environment with a wider MPO bond (charge-0 group of 6 states) and an all-nonzero
    two-site MPO gives items of more than CXA_MIX_SLICES slices: both planners refuse (Heff2
    then falls back to the three-launch chain)."""
    from cytnx_amd.heff import MixUnsupported, mix_plan_arrays_native
    from cytnx_amd.unitensor import BD_BRA, Bond
    Lp, psi, W, Y, labs = _mix_args(20, 9, False)
    wide = Bond(BD_BRA, [[0], [2], [-2]], [6, 2, 2])
    Lb = Lp.bonds()
    Lw = _meta([Lb[0], wide, Lb[2]], ["-1", "-5", "0"])
    W10 = np.ones((10,) + W.shape[1:])
    with pytest.raises(MixUnsupported):
        mix_plan_arrays(Lw, labs[0], psi, labs[1], W10, Y, labs[2])
    with pytest.raises(MixUnsupported):
        mix_plan_arrays_native(Lw, labs[0], psi, labs[1], W10, Y, labs[2])
