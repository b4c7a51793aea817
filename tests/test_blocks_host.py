"""Host-side block bookkeeping (no device): the C++ zero-flux enumerator against the oracle's
brute force and against the block lists stored in the reference's fixture files."""
import glob
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import REF_FIXTURES
from cytnx_amd import fileio
from cytnx_amd.unitensor import BD_BRA, BD_KET, Bond, enumerate_blocks


@pytest.mark.parametrize("seed", range(12))
def test_enumeration_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    rank = int(rng.integers(1, 6))
    bonds, obonds = [], []
    for _ in range(rank):
        nq = int(rng.integers(1, 5))
        q = [int(x) for x in rng.integers(-3, 4, size=nq)]
        d = [int(x) for x in rng.integers(1, 4, size=nq)]
        t = BD_KET if rng.random() < 0.5 else BD_BRA
        bonds.append(Bond(t, [[x] for x in q], d))
        obonds.append(po.OBond(t, q, d))
    got = enumerate_blocks(bonds).astype(int).tolist()
    assert got == po.enumerate_blocks(obonds)


def test_enumeration_matches_fixture_block_lists():
    n = 0
    for path in sorted(glob.glob(os.path.join(REF_FIXTURES, "*.cytnx"))):
        ut = fileio.read_unitensor(path)
        if ut.uten_type != fileio.UTEN_BLOCK:
            continue
        bonds = [Bond(b.btype, [q[0] for q in b.qnums], b.degs) for b in ut.bonds]
        assert enumerate_blocks(bonds).astype(int).tolist() == ut.itoi, path
        n += 1
    assert n >= 10


def test_enumeration_two_site_chi1024_profile():
    # the U(1) two-site wave function (l, s1, s2, r) at chi=1024: 31 sectors per virtual bond
    from cytnx_amd.heisenberg import gaussian_sectors
    q, d = gaussian_sectors(1024)
    assert sum(d) == 1024 and len(q) == 31
    vl = Bond(BD_KET, [[x] for x in q], d)
    ph = Bond(BD_KET, [[1], [-1]], [1, 1])
    vr = Bond(BD_BRA, [[x] for x in q], d)
    got = enumerate_blocks([vl, ph, ph, vr])
    ref = po.enumerate_blocks([po.OBond(BD_KET, q, d), po.OBond(BD_KET, [1, -1], [1, 1]),
                               po.OBond(BD_KET, [1, -1], [1, 1]), po.OBond(BD_BRA, q, d)])
    assert got.astype(int).tolist() == ref
