"""The group tree's particle blocks (scalfmm_amd.group_tree) on the CPU: the
byte layout of GroupTree/Core/FGroupOfParticles.hpp and the leaf grouping of
GroupTree/Core/FGroupTree.hpp:178-268, checked against the reference's own
rules and against the oracle's leaf keys (pinned to the reference's Morton
code).  The evaluation through the blocks is in test_gpu_group_tree.py."""
import numpy as np
import pytest

from scalfmm_amd import group_tree as gt


def _particles(n, seed=5):
    r = np.random.default_rng(seed)
    return r.random(n), r.random(n), r.random(n), r.random(n) - 0.5


def test_round_to_upper_particles():
    # the reference's comment (FGroupOfParticles.hpp:53-58) with 8 doubles per 64 bytes:
    # a multiple of the pack still gains a whole pack
    assert [gt.round_to_upper_particles(n) for n in (0, 1, 7, 8, 63, 64)] == [8, 8, 8, 16, 64, 72]


def test_morton_keys_match_oracle(oracle, fixture20k):
    d = fixture20k["data"]
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    for h in (3, 5, 7):
        keys = gt.morton_keys(x, y, z, fixture20k["width"], fixture20k["center"], h)
        st, ref = oracle.particle_keys(x, y, z, fixture20k["center"], fixture20k["width"], h)
        assert st == 0
        np.testing.assert_array_equal(keys, ref)


def test_morton_keys_faces(oracle):
    # corners and the upper faces (the last leaf, FCoordinateComputer.hpp:19-21)
    x = np.array([0.0, 1.0, 0.5, 1.0, 0.25])
    y = np.array([0.0, 1.0, 1.0, 0.0, 0.75])
    z = np.array([0.0, 1.0, 0.0, 0.5, 0.125])
    keys = gt.morton_keys(x, y, z, 1.0, (0.5, 0.5, 0.5), 4)
    st, ref = oracle.particle_keys(x, y, z, (0.5, 0.5, 0.5), 1.0, 4)
    np.testing.assert_array_equal(keys, ref)
    assert keys[1] == 8 ** 3 - 1 and keys[0] == 0
    with pytest.raises(gt.FmmError):
        gt.morton_keys(np.array([1.5]), np.array([0.5]), np.array([0.5]), 1.0, (0.5, 0.5, 0.5), 4)


def test_block_header_bytes():
    # field offsets of BlockHeader / LeafHeader (alignas 8, long long FSize and MortonIndex)
    g = gt.FGroupOfParticles(40, 97, 3, 11, nvals=2)
    raw = g.getRawBuffer()
    i8 = lambda off: int(raw[off:off + 8].view(np.int64)[0])  # noqa: E731
    assert (i8(0), i8(8)) == (40, 97)
    assert int(raw[16:20].view(np.int32)[0]) == 3
    nalloc = gt.round_to_upper_particles(11 + 7 * 3)
    assert i8(24) == nalloc == 40
    assert i8(32) == 192  # 64 + 3 * 24 = 136, next 64-byte boundary
    assert i8(40) == i8(48) == 8 * nalloc
    assert i8(56) == 11
    assert all(i8(64 + 24 * k) == -1 for k in range(3))  # leaves not yet allocated
    off = g.newLeaf(41, 0, 5, 0)
    assert off == 8
    assert (i8(64), i8(72), i8(80)) == (41, 5, 0)
    assert g.getRawAttributesBuffer().size == 4 * 2 * nalloc


def test_tree_blocks_hold_sorted_particles():
    x, y, z, q = _particles(3000)
    h, per_block = 5, 37
    t = gt.FGroupTree(h, 1.0, (0.5, 0.5, 0.5), per_block, x, y, z, q)
    keys = gt.morton_keys(x, y, z, 1.0, (0.5, 0.5, 0.5), h)
    leaf_keys, counts = np.unique(keys, return_counts=True)
    assert t.getNbParticleGroup() == -(-len(leaf_keys) // per_block)
    seen_keys, total = [], 0
    order = t.original_index
    for b in range(t.getNbParticleGroup()):
        g = t.getParticleGroup(b)
        nl = g.getNumberOfLeavesInBlock()
        assert nl == min(per_block, len(leaf_keys) - b * per_block)
        assert g.getStartingIndex() == g.getLeafMortonIndex(0)
        assert g.getEndingIndex() == g.getLeafMortonIndex(nl - 1) + 1
        nalloc = int(g.header["nbParticlesAllocatedInGroup"][0])
        assert nalloc == gt.round_to_upper_particles(g.getNbParticlesInGroup() + 7 * nl)
        off = 0
        for leaf in range(nl):
            assert int(g.leaf_headers["offSet"][leaf]) == off
            c = g.getLeaf(leaf)
            n = c.getNbParticles()
            src = order[total:total + n]
            # the leaf's particles, in input order (stable sort)
            assert np.all(np.diff(src) > 0)
            assert np.all(keys[src] == g.getLeafMortonIndex(leaf))
            px, py, pz = c.getPositions()
            np.testing.assert_array_equal(px, x[src])
            np.testing.assert_array_equal(pz, z[src])
            np.testing.assert_array_equal(c.getPhysicalValues(), q[src])
            off = gt.round_to_upper_particles(off + n)
            total += n
            seen_keys.append(g.getLeafMortonIndex(leaf))
    assert total == len(x)
    np.testing.assert_array_equal(seen_keys, leaf_keys)


def test_tree_several_physical_values():
    x, y, z, q = _particles(500, seed=9)
    q2 = np.stack([q, 2 * q + 1])
    t = gt.FGroupTree(4, 1.0, (0.5, 0.5, 0.5), 10, x, y, z, q2)
    assert t.nvals == 2
    got = [[], []]
    t.forEachLeaf(lambda c: [got[v].append(c.getPhysicalValues(v).copy()) for v in range(2)])
    for v in range(2):
        np.testing.assert_array_equal(np.concatenate(got[v]), q2[v][t.original_index])
    # results gathered back to the input order through the attribute arrays
    for g in t.groups:
        for k, a in enumerate(g.attributes):
            a[:] = k
    pot, fx, fy, fz = t.results(idx_vals=1)
    assert np.all(pot == 1) and np.all(fx == 3) and np.all(fy == 5) and np.all(fz == 7)


def test_cell_blocks_layout(oracle, fixture20k):
    """Cell blocks of every level (FGroupTree.hpp:224-356, FGroupOfCells.hpp:
    125-168): 24-byte header, Morton indexes, 24-byte FBasicCellPOD per cell
    (FBasicCellPOD.hpp:9-13) with the coordinates of the Morton index
    (checked against the oracle's decode); nbElementsPerBlock cells per block;
    the cells of a level are the oracle's cell keys; FChebCellPOD buffers of
    nvals x 2 p^3 doubles per cell (FChebCellPOD.hpp:11-21)."""
    d = fixture20k["data"]
    h, per, order, nvals = 5, 100, 4, 2
    q = np.stack([d[:, 3], -d[:, 3]])
    t = gt.FGroupTree(h, fixture20k["width"], fixture20k["center"], per, d[:, 0], d[:, 1], d[:, 2], q, order=order)
    ot = oracle.Tree(d[:, 0], d[:, 1], d[:, 2], fixture20k["center"], fixture20k["width"], h)
    assert sorted(t.cell_groups) == list(range(1, h))
    for level in range(1, h):
        blocks = [t.getCellGroup(level, i) for i in range(t.getNbCellGroupAtLevel(level))]
        keys = np.concatenate([g.cell_indexes for g in blocks])
        if level >= 2:
            np.testing.assert_array_equal(keys, ot.cell_keys(level))
        for i, g in enumerate(blocks):
            n = g.getNumberOfCellsInBlock()
            assert n == per or i == len(blocks) - 1
            assert g.raw.nbytes == 24 + 8 * n + 24 * n
            assert g.getStartingIndex() == g.cell_indexes[0] and g.getEndingIndex() == g.cell_indexes[-1] + 1
            np.testing.assert_array_equal(g.cells["mortonIndex"], g.cell_indexes)
            assert g.multipoles.size == n * nvals * 2 * order ** 3 == g.locals.size
            for c in (0, n - 1):
                coord = oracle.demorton(int(g.cell_indexes[c]))
                assert list(g.cells["coordinates"][c]) == list(coord)
            assert np.all(g.cells["level"] == 0)
    t.getCellGroup(2, 0).getLocal(0, 1)[:] = 3.0
    t.resetCells()
    assert not t.getCellGroup(2, 0).locals.any()
