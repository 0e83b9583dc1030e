"""Oracle indexing: Morton keys, leaf coordinates, P2P and M2L lists.

Pins (reference paths relative to its source tree):
  - UTests/utestMorton.cpp:15-60 known answers and round trips;
  - a closed-form bit interleave written independently of the oracle;
  - brute-force neighbour / interaction definitions and their reference order
    (Src/Containers/FTreeCoordinate.hpp:189-264);
  - full-level interaction counts (6*2^l-8)^3-(3*2^l-2)^3 and (3*2^L-2)^3;
  - the FTestKernels/ValidateFMMAlgo counting invariant
    (Src/Components/FTestKernels.hpp:138): through near and far field every
    target sees every other particle exactly once.
"""
import itertools

import numpy as np
import pytest


def interleave(x, y, z):
    k = 0
    for b in range(21):
        k |= ((z >> b) & 1) << (3 * b)
        k |= ((y >> b) & 1) << (3 * b + 1)
        k |= ((x >> b) & 1) << (3 * b + 2)
    return k


def test_morton_known_answers(oracle):
    # UTests/utestMorton.cpp:41-60
    assert oracle.morton(0, 0, 0) == 0
    assert oracle.morton(1, 1, 1) == 7
    assert oracle.morton(0, 1, 1) == 3
    assert oracle.morton(2, 2, 2) == 7 << 3
    assert oracle.morton(1, 2, 4) == 84


@pytest.mark.parametrize("xyz", [(5, 1, 7), (2, 8, 3), (51, 11, 47)])
def test_morton_round_trip(oracle, xyz):
    # UTests/utestMorton.cpp:15-36
    assert oracle.demorton(oracle.morton(*xyz)) == xyz


def test_morton_closed_form(oracle):
    rng = np.random.default_rng(7)
    for x, y, z in rng.integers(0, 1 << 19, size=(200, 3)):
        assert oracle.morton(x, y, z) == interleave(int(x), int(y), int(z))


def test_leaf_coordinates_edges(oracle):
    h = 4
    c = np.array([0.5, 0.5, 0.5])
    # a particle on the far face maps into the last box (FCoordinateComputer.hpp:19-21)
    st, k = oracle.particle_keys(np.array([1.0]), np.array([1.0]), np.array([1.0]), c, 1.0, h)
    assert st == 0 and k[0] == interleave(7, 7, 7)
    # on an inner leaf boundary: the upper leaf
    st, k = oracle.particle_keys(np.array([0.125]), np.array([0.0]), np.array([0.25]), c, 1.0, h)
    assert st == 0 and k[0] == interleave(1, 0, 2)
    # outside the box: an error, not a key
    st, _ = oracle.particle_keys(np.array([1.0 + 1e-12]), np.array([0.5]), np.array([0.5]), c, 1.0, h)
    assert st == 2
    st, _ = oracle.particle_keys(np.array([-1e-300]), np.array([0.5]), np.array([0.5]), c, 1.0, h)
    assert st == 2


def _random_tree(oracle, n, h, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        pts = np.clip(rng.normal(0.3, 0.12, size=(n, 3)), 0, 1)
    else:
        pts = rng.random((n, 3))
    x, y, z = pts.T.copy()
    return x, y, z, oracle.Tree(x, y, z, [0.5, 0.5, 0.5], 1.0, h)


@pytest.mark.parametrize("clustered", [False, True])
def test_p2p_list_brute_force(oracle, clustered):
    h = 4
    x, y, z, t = _random_tree(oracle, 300, h, 11, clustered)
    keys = t.cell_keys(h - 1)
    coords = np.array([oracle.demorton(k) for k in keys])
    off, nbr, pos = t.p2p_list()
    for i in range(len(keys)):
        d = coords - coords[i]
        cheb = np.abs(d).max(axis=1)
        expect = [j for j in range(len(keys)) if cheb[j] == 1]
        # reference order: ascending position ((dx+1)*3+(dy+1))*3+dz+1
        epos = [((d[j, 0] + 1) * 3 + (d[j, 1] + 1)) * 3 + d[j, 2] + 1 for j in expect]
        order = np.argsort(epos, kind="stable")
        got = list(nbr[off[i]:off[i + 1]])
        assert got == [expect[o] for o in order]
        assert list(pos[off[i]:off[i + 1]]) == [epos[o] for o in order]
        assert len(got) <= 26


@pytest.mark.parametrize("clustered", [False, True])
def test_m2l_list_brute_force(oracle, clustered):
    h = 5
    x, y, z, t = _random_tree(oracle, 600, h, 5, clustered)
    for level in range(2, h):
        keys = t.cell_keys(level)
        coords = np.array([oracle.demorton(k) for k in keys])
        off, src, pos = t.m2l_list(level)
        for i in range(len(keys)):
            d = coords - coords[i]
            pd = (coords >> 1) - (coords[i] >> 1)
            pcheb = np.abs(pd).max(axis=1)
            cheb = np.abs(d).max(axis=1)
            cand = [j for j in range(len(keys)) if pcheb[j] == 1 and cheb[j] > 1]
            # reference order: parent offset (x, y, z lexicographic), then cousin index
            sortkey = [((pd[j, 0] + 1) * 9 + (pd[j, 1] + 1) * 3 + (pd[j, 2] + 1)) * 8 + (keys[j] & 7) for j in cand]
            order = np.argsort(sortkey, kind="stable")
            exp = [cand[o] for o in order]
            got = list(src[off[i]:off[i + 1]])
            assert got == exp
            epos = [((d[j, 0] + 3) * 7 + (d[j, 1] + 3)) * 7 + d[j, 2] + 3 for j in exp]
            assert list(pos[off[i]:off[i + 1]]) == epos
            assert len(got) <= 189


def test_full_level_counts(oracle):
    # one particle at the centre of every leaf of a height-5 tree
    h = 5
    m = 1 << (h - 1)
    g = (np.arange(m) + 0.5) / m
    x, y, z = (a.ravel().copy() for a in np.meshgrid(g, g, g, indexing="ij"))
    t = oracle.Tree(x, y, z, [0.5, 0.5, 0.5], 1.0, h)
    for level in range(2, h):
        s = 1 << level
        off, src, pos = t.m2l_list(level)
        assert len(src) == (6 * s - 8) ** 3 - (3 * s - 2) ** 3
        assert np.diff(off).max() == (189 if level >= 3 else 56)
    off, nbr, pos = t.p2p_list()
    assert len(nbr) + len(x) == (3 * m - 2) ** 3
    assert np.diff(off).max() == 26


@pytest.mark.parametrize("clustered", [False, True])
def test_counting_invariant(oracle, clustered):
    """ValidateFMMAlgo pattern: near + far field reach every other particle once."""
    h = 5
    n = 700
    x, y, z, t = _random_tree(oracle, n, h, 3, clustered)
    L = h - 1
    start, count = t.leaf_ranges()
    # particles below each cell of every level
    below = {L: count.astype(np.int64)}
    keys = {L: t.cell_keys(L)}
    for level in range(L - 1, 1, -1):
        keys[level] = t.cell_keys(level)
        parent_of_child = np.searchsorted(keys[level], keys[level + 1] >> 3)
        below[level] = np.bincount(parent_of_child, weights=below[level + 1], minlength=len(keys[level])).astype(np.int64)
    far = {}
    for level in range(2, h):
        off, src, pos = t.m2l_list(level)
        per = np.array([below[level][src[off[i]:off[i + 1]]].sum() for i in range(len(keys[level]))])
        far[level] = per
    off, nbr, pos = t.p2p_list()
    for i in range(len(keys[L])):
        total = count[i] - 1 + count[nbr[off[i]:off[i + 1]]].sum()
        k = keys[L][i]
        for level in range(L, 1, -1):
            idx = np.searchsorted(keys[level], k >> (3 * (L - level)))
            total += far[level][idx]
        assert total == n - 1
