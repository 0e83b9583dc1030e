"""The group tree's cell blocks as the boundary (hfmm_get_cell_blocks /
hfmm_set_cell_blocks) on the GPU, in the flow of
Tests/GroupTree/testBlockedChebyshev.cpp:84-200: the 20k particle file, an
FGroupTree of 5 levels with 250 cells per block and FChebCellPOD cells of
order 6, FGroupSeqAlgorithm::execute().

Bars:
  - after execute(), the first p^3 values of every cell's multipole and local
    buffers (FChebCellPOD.hpp:85-100) equal the oracle's expansions of that
    cell: M within 1e-13, L within 1e-12 (relative, per level);
  - the blocks' Morton indexes and POD coordinates equal the oracle tree's
    cell keys and their coordinates;
  - the second half of each slot (FChebKernel's compressed copy) is left as
    the caller wrote it;
  - seeding: multipoles written into a fresh tree's blocks, then
    execute(M2L | L2L | L2P) reproduce the far field of a full run bit for
    bit (the reference's passes read the cells as the tree holds them);
  - the global-SVD kernel writes its compressed expansions behind the nodal
    ones: B^T M with the kernel's own basis B;
  - blocks naming a cell the tree does not have, or with indexes out of
    order, are rejected.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H, ORDER, PER_BLOCK = 5, 6, 250


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def data(oracle):
    return oracle.read_bfma(os.path.join(ROOT, "tests", "golden", "unitCubeRef20kDouble.bfma"))


def _tree(f, order=ORDER):
    from scalfmm_amd import group_tree as gt
    d = f["data"]
    return gt.FGroupTree(H, f["width"], f["center"], PER_BLOCK, d[:, 0], d[:, 1], d[:, 2], d[:, 3], order=order)


def _cells(tree, level, which):
    """(keys, first p^3 values per cell) of a level, in block order."""
    nn = ORDER ** 3
    keys, vals = [], []
    for i in range(tree.getNbCellGroupAtLevel(level)):
        g = tree.getCellGroup(level, i)
        for c in range(g.getNumberOfCellsInBlock()):
            keys.append(g.getCellMortonIndex(c))
            v = g.getMultipole(c) if which == "M" else g.getLocal(c)
            vals.append(v[:nn].copy())
    return np.array(keys), np.array(vals)


def test_cell_blocks_hold_the_oracle_expansions(oracle, data):
    from scalfmm_amd import FChebSymKernel
    from scalfmm_amd.group_tree import FGroupSeqAlgorithm, morton_to_coordinate
    f = data
    d = f["data"]
    tree = _tree(f)
    # the compressed half of every slot belongs to the caller
    for lv in range(2, H):
        for i in range(tree.getNbCellGroupAtLevel(lv)):
            g = tree.getCellGroup(lv, i)
            for c in range(g.getNumberOfCellsInBlock()):
                g.getMultipole(c)[ORDER ** 3:] = 7.0
    kernel = FChebSymKernel(H, f["width"], f["center"], ORDER)
    FGroupSeqAlgorithm(tree, kernel).execute()
    # one full execution (a further run.execute() would add into the kept cells)
    run = oracle.Run(oracle.Cheb(ORDER), d[:, 0], d[:, 1], d[:, 2], d[:, 3], f["center"], f["width"], H)
    for lv in range(2, H):
        keys, M = _cells(tree, lv, "M")
        _, L = _cells(tree, lv, "L")
        assert np.array_equal(keys, run.tree.cell_keys(lv)), lv
        OM, OL = run.expansions(lv)
        em, el = _rel(M, OM), _rel(L, OL)
        print(f"  level {lv}: {len(keys)} cells, M rel {em:.2e}, L rel {el:.2e}", flush=True)
        assert em <= 1e-13 and el <= 1e-12, (lv, em, el)
        g = tree.getCellGroup(lv, 0)
        assert np.all(g.getMultipole(0)[ORDER ** 3:] == 7.0)
        coords = g.cells["coordinates"][0].tolist()
        assert coords == morton_to_coordinate(g.getCellMortonIndex(0), lv)
    # leaves: blocks of 250 cells, intervals [first, last + 1)
    g = tree.getCellGroup(H - 1, 0)
    assert g.getNumberOfCellsInBlock() == PER_BLOCK
    assert g.getStartingIndex() == g.getCellMortonIndex(0)
    assert g.getEndingIndex() == g.getCellMortonIndex(PER_BLOCK - 1) + 1


def test_seeded_multipoles_reproduce_the_far_field(data):
    from scalfmm_amd import FChebSymKernel, FFmmFarField, FFmmL2L, FFmmL2P, FFmmM2L
    from scalfmm_amd.group_tree import FGroupSeqAlgorithm
    f = data
    full = _tree(f)
    FGroupSeqAlgorithm(full, FChebSymKernel(H, f["width"], f["center"], ORDER)).execute(FFmmFarField)
    seeded = _tree(f)
    for lv in range(2, H):
        for i in range(full.getNbCellGroupAtLevel(lv)):
            seeded.getCellGroup(lv, i).multipoles[:] = full.getCellGroup(lv, i).multipoles
    FGroupSeqAlgorithm(seeded, FChebSymKernel(H, f["width"], f["center"], ORDER)).execute(FFmmM2L | FFmmL2L | FFmmL2P)
    a, b = full.results(), seeded.results()
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    for lv in range(2, H):
        _, L1 = _cells(full, lv, "L")
        _, L2 = _cells(seeded, lv, "L")
        assert np.array_equal(L1, L2), lv


def test_global_kernel_writes_compressed_expansions(oracle, data):
    from scalfmm_amd import FChebKernel
    from scalfmm_amd.group_tree import FGroupSeqAlgorithm
    f = data
    order = 5
    tree = _tree(f, order=order)
    kernel = FChebKernel(H, f["width"], f["center"], order)
    FGroupSeqAlgorithm(tree, kernel).execute()
    nn = order ** 3
    Bg = kernel.engine.global_operators()[1]
    rank = Bg.shape[1]
    for lv in range(2, H):
        g = tree.getCellGroup(lv, 0)
        m = g.getMultipole(0)
        assert np.allclose(m[nn:nn + rank], Bg.T @ m[:nn], rtol=1e-12, atol=1e-12 * np.abs(m[nn:nn + rank]).max())


def test_bad_cell_blocks_are_rejected(data):
    """A cell the tree does not have, and indexes out of order."""
    from scalfmm_amd import FChebSymKernel, FmmError
    from scalfmm_amd.group_tree import FGroupSeqAlgorithm
    f = data
    tree = _tree(f)
    kernel = FChebSymKernel(H, f["width"], f["center"], ORDER)
    FGroupSeqAlgorithm(tree, kernel).execute()
    g = tree.getCellGroup(3, 0)
    n = g.getNumberOfCellsInBlock()
    for case in ("unknown", "order"):
        raw = g.raw.copy()
        hdr = raw[:24].view(np.int64)
        idx = raw[24:24 + 8 * n].view(np.int64)
        if case == "unknown":
            hdr[1] = 8 ** 3 + 6  # the interval reaches past the level's keys
            idx[-1] = 8 ** 3 + 5
        else:
            idx[0], idx[1] = idx[1], idx[0]
        with pytest.raises(FmmError):
            kernel.engine.get_cell_blocks(3, [raw], [g.multipoles.copy()], [g.locals.copy()])
