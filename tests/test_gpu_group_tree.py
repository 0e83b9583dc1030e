"""The group tree's particle blocks as the boundary (hfmm_evaluate_particle_blocks)
on the GPU, in the flow of Tests/GroupTree/testBlockedChebyshev.cpp:84-200: a
20k particle file, an FGroupTree of 5 levels with 250 leaves per block,
FChebSymKernel of order 6, execute(), results read per leaf out of the blocks'
attribute buffers and compared with a direct sum.

Bars: bitwise equal to the engine's particle-array path (the blocks hold the
particles in the order the engine sorts them to); within 1e-12 (relative L2)
of the oracle; the direct-sum errors of the oracle's own FMM on this file
(potential 9.6e-8, forces 1.1e-5 relative L2) doubled; results added, never
overwritten; cells kept between executions on one tree as the reference's
tree keeps them (split executions compose; without resetCells() a repeated
execution matches the oracle's repeated execution within 1e-12); bad headers
and misplaced particles rejected with the attribute buffers untouched."""
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
    f = oracle.read_bfma(os.path.join(ROOT, "tests", "golden", "unitCubeRef20kDouble.bfma"))
    return f


def _tree(f, per_block=PER_BLOCK, q=None, h=H):
    from scalfmm_amd import group_tree as gt
    d = f["data"]
    return gt.FGroupTree(h, f["width"], f["center"], per_block, d[:, 0], d[:, 1], d[:, 2],
                         d[:, 3] if q is None else q)


def _kernel(f, precision=None, h=H):
    from scalfmm_amd import FChebSymKernel, _abi
    return FChebSymKernel(h, f["width"], f["center"], ORDER,
                          precision=_abi.PRECISION_FP64 if precision is None else precision)


def _engine_path(f, ops=None, levels=None, q=None, h=H):
    from scalfmm_amd.fmm import NEAR_AND_FAR, Engine
    d = f["data"]
    e = Engine(h, f["width"], f["center"], ORDER)
    e.set_particles(d[:, 0], d[:, 1], d[:, 2], d[:, 3] if q is None else q)
    e.build_tree()
    e.evaluate(NEAR_AND_FAR if ops is None else ops, levels)
    r = e.results()
    e.close()
    return r


def test_blocked_chebyshev_flow(oracle, data):
    from scalfmm_amd import group_tree as gt
    t = _tree(data)
    print(f"  {t.getNbParticleGroup()} particle blocks", flush=True)
    gt.FGroupSeqAlgorithm(t, _kernel(data)).execute()
    got = t.results()
    ref = _engine_path(data)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    d = data["data"]
    orc = oracle.fmm(d[:, 0], d[:, 1], d[:, 2], d[:, 3], data["center"], data["width"], H, ORDER)
    for k in range(4):
        assert _rel(got[k], orc[k]) < 1e-12
    assert oracle.accuracy(d[:, 4], got[0])["rel_l2"] < 2 * 9.6e-8
    for k in range(3):
        assert oracle.accuracy(d[:, 5 + k], got[1 + k])["rel_l2"] < 2 * 1.1e-5
    # per leaf, as the reference's check walks forEachCellLeaf
    pots = []
    t.forEachLeaf(lambda c: pots.append(c.getPotentials().copy()))
    np.testing.assert_array_equal(np.concatenate(pots), got[0][t.original_index])


def test_blocks_executions_keep_cells(oracle, data):
    """The tree keeps its cells between executions (passes add into them) until
    resetCells(); the containers add every execution's results."""
    from scalfmm_amd import group_tree as gt
    t = _tree(data, per_block=97)
    kernel = _kernel(data)
    alg = gt.FGroupSeqAlgorithm(t, kernel)
    alg.execute()
    once = [a.copy() for a in t.results()]
    # cells reset: the second execution adds the same values again
    t.resetCells()
    alg.execute()
    for a, b in zip(t.results(), once):
        np.testing.assert_array_equal(a, 2 * b)
    # no reset: the third execution starts from the cells the second left, as the
    # oracle's second execution on one tree does
    alg.execute()
    d = data["data"]
    run = oracle.Run(oracle.Cheb(ORDER), d[:, 0], d[:, 1], d[:, 2], d[:, 3], data["center"], data["width"], H)
    acc = run.execute()
    for a, b, o in zip(t.results(), acc, once):
        assert _rel(a, b + o) < 1e-12
    # blocks handed over in reverse order: a new particle order, so a new tree
    # with fresh cells; the same particles give the same sums
    t2 = _tree(data, per_block=97)
    rev = t2.groups[::-1]
    kernel.engine.evaluate_particle_blocks([g.raw for g in rev], [g.attributes_raw for g in rev], 1)
    for a, b in zip(t2.results(), once):
        np.testing.assert_array_equal(a, b)


def test_blocks_split_executions(data):
    """execute(P2M|M2M) then execute(M2L|L2L|L2P) on one tree is the far field."""
    from scalfmm_amd import _abi, group_tree as gt
    t = _tree(data)
    alg = gt.FGroupSeqAlgorithm(t, _kernel(data))
    alg.execute(_abi.P2M | _abi.M2M)
    alg.execute(_abi.M2L | _abi.L2L | _abi.L2P)
    ref = _engine_path(data, _abi.FAR_FIELD)
    for a, b in zip(t.results(), ref):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("ops_name,levels", [("NEAR_FIELD", None), ("FAR_FIELD", None), ("NEAR_AND_FAR", (3, 5)),
                                             ("NEAR_AND_FAR", (2, 4)), ("M2L_L2P", None)])
def test_blocks_partial_operations(data, ops_name, levels):
    from scalfmm_amd import _abi, group_tree as gt
    ops = _abi.P2M | _abi.M2M | _abi.M2L | _abi.L2P if ops_name == "M2L_L2P" else getattr(_abi, ops_name)
    t = _tree(data)
    alg = gt.FGroupSeqAlgorithm(t, _kernel(data))
    if levels is None:
        alg.execute(ops)
    else:
        alg.execute(ops, *levels)
    ref = _engine_path(data, ops, levels)
    for a, b in zip(t.results(), ref):
        np.testing.assert_array_equal(a, b)


def test_blocks_several_physical_values(data):
    from scalfmm_amd import group_tree as gt
    d = data["data"]
    q2 = np.stack([d[:, 3], np.cos(7 * d[:, 0])])
    t = _tree(data, q=q2)
    gt.FGroupSeqAlgorithm(t, _kernel(data)).execute()
    for v in range(2):
        ref = _engine_path(data, q=q2[v])
        for a, b in zip(t.results(idx_vals=v), ref):
            assert _rel(a, b) < 1e-14


def test_blocks_mixed_precision(data):
    from scalfmm_amd import _abi, group_tree as gt
    t = _tree(data)
    gt.FGroupSeqAlgorithm(t, _kernel(data, _abi.PRECISION_MIXED)).execute()
    ref = _engine_path(data)
    got = t.results()
    assert _rel(got[0], ref[0]) < 1e-6
    for k in range(1, 4):
        assert _rel(got[k], ref[k]) < 1e-5


def test_blocks_rejected(data):
    from scalfmm_amd import FmmError, group_tree as gt
    kernel = _kernel(data)
    # a particle moved into another leaf
    t = _tree(data)
    g = t.getParticleGroup(1)
    x = g.getLeaf(3).getPositions()[0]
    x[0] = 0.999 if x[0] < 0.5 else 0.001
    with pytest.raises(FmmError, match="particle block 1: particle"):
        gt.FGroupSeqAlgorithm(t, kernel).execute()
    assert all(np.all(g.attributes_raw == 0) for g in t.groups)
    # counts that do not add up
    t = _tree(data)
    t.getParticleGroup(0).header["nbParticlesInGroup"] += 1
    with pytest.raises(FmmError, match="particle block 0: leaf particle counts"):
        gt.FGroupSeqAlgorithm(t, kernel).execute()
    # leaves out of order
    t = _tree(data)
    lh = t.getParticleGroup(2).leaf_headers
    lh["mindex"][[4, 5]] = lh["mindex"][[5, 4]]
    with pytest.raises(FmmError, match="particle block 2: leaf 5"):
        gt.FGroupSeqAlgorithm(t, kernel).execute()
    # a leaf running past the block's arrays
    t = _tree(data)
    g = t.getParticleGroup(0)
    g.leaf_headers["offSet"][-1] = int(g.header["nbParticlesAllocatedInGroup"][0])
    with pytest.raises(FmmError, match="outside the block"):
        gt.FGroupSeqAlgorithm(t, kernel).execute()
    assert all(np.all(g.attributes_raw == 0) for g in t.groups)
    # working levels outside 2 <= upper <= lower <= height
    t = _tree(data)
    with pytest.raises(FmmError, match="working levels"):
        gt.FGroupSeqAlgorithm(t, kernel).execute(upper_level=4, lower_level=3)
    # no blocks: nothing to do
    t.groups = []
    gt.FGroupSeqAlgorithm(t, kernel).execute()
