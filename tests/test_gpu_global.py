"""FChebKernel's M2L scheme on the GPU (hfmm_set_m2l_scheme(HFMM_M2L_GLOBAL_SVD)):
one global SVD of the 316 far-field operators
(Kernels/Chebyshev/FChebM2LHandler.hpp:216-523) and the compressed passes of
Kernels/Chebyshev/FChebKernel.hpp:93-200, against the oracle's restatement.

Bars: the rank equal to the oracle's; singular values within 1e-12
(relative); U C_t B^T at every far position within 1e-12 of the oracle's (the
bases themselves are fixed only up to signs); FMM values within 1e-12
(relative L2) of the oracle's FChebKernel FMM; the reference's unit test
(UTests/utestChebyshev.cpp:51-66 through FUKernelTester.hpp:56-170: order 6,
DirectDouble.bfma, 4 levels) with its gates; split executions on kept cells
equal one execution bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def og5(oracle):
    return oracle.ChebGlobal(5)


def _engine(h, width, center, order):
    from scalfmm_amd import _abi
    from scalfmm_amd.fmm import Engine
    e = Engine(h, width, center, order)
    e.set_m2l_scheme(_abi.M2L_GLOBAL_SVD)
    return e


def test_global_operators_match_oracle(og5):
    e = _engine(3, 1.0, (0.5, 0.5, 0.5), 5)
    U, B, Cm, S1, S2 = e.global_operators()
    oU, oB, oC, oS1, oS2 = og5.operators()
    print(f"  rank {U.shape[1]} (oracle {og5.rank})", flush=True)
    assert U.shape[1] == og5.rank
    assert _rel(S1, oS1) < 1e-12 and _rel(S2, oS2) < 1e-12
    for pos in range(343):
        if not oC[pos].any():
            assert not Cm[pos].any()
            continue
        A, oA = U @ Cm[pos] @ B.T, oU @ oC[pos] @ oB.T
        assert _rel(A, oA) < 1e-12, pos
    e.close()


@pytest.mark.parametrize("ops_name", ["all", "far", "near"])
def test_global_fmm_matches_oracle(oracle, og5, fixture20k, ops_name):
    ops = {"all": oracle.ALL, "far": oracle.ALL & ~oracle.P2P, "near": oracle.P2P}[ops_name]
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    e = _engine(5, w, c, 5)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(ops)
    got = e.results()
    ref = oracle.fmm_global(og5, x, y, z, q, c, w, 5, ops=ops)
    errs = [_rel(a, b) for a, b in zip(got, ref[:4])]
    print(f"  {ops_name}: rel L2 vs oracle {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-12
    e.close()


def test_global_working_levels_and_kept_cells(oracle, og5, fixture20k):
    from scalfmm_amd import _abi
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    e = _engine(5, w, c, 5)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(oracle.ALL, (3, 5))
    got = e.results()
    ref = oracle.fmm_global(og5, x, y, z, q, c, w, 5, levels=(3, 5))
    assert max(_rel(a, b) for a, b in zip(got, ref[:4])) < 1e-12
    far = oracle.ALL & ~oracle.P2P
    e.evaluate(far)
    one = [a.copy() for a in e.results()]
    e.set_cell_mode(_abi.CELLS_KEPT)
    e.evaluate(_abi.P2M | _abi.M2M)
    a = [v.copy() for v in e.results()]
    e.evaluate(_abi.M2L | _abi.L2L | _abi.L2P)
    for u, v, o in zip(a, e.results(), one):
        np.testing.assert_array_equal(u + v, o)
    e.close()


def test_utest_chebyshev_cheb_kernel(oracle):
    """TestChebKernel: FChebKernel of order 6 over the classic octree of 4
    levels on DirectDouble.bfma, with the FUKernelTester gates."""
    from scalfmm_amd import FChebKernel, octree
    from scalfmm_amd.fma import FFmaGenericLoader
    loader = FFmaGenericLoader(os.path.join(ROOT, "tests", "golden", "DirectDouble.bfma"))
    d = loader.data
    n = loader.getNumberOfParticles()
    tree = octree.FOctree(4, 2, loader.getBoxWidth(), loader.getCenterOfBox())
    tree.insert_many(d[:, 0], d[:, 1], d[:, 2], d[:, 3], np.arange(n))
    kernel = FChebKernel(4, loader.getBoxWidth(), loader.getCenterOfBox(), 6)
    octree.FFmmAlgorithmThread(tree, kernel).execute()
    pot = np.zeros(n)
    forces = np.zeros((n, 3))

    def visit(leaf):
        t = leaf.getTargets()
        idx = t.getIndexes()
        pot[idx] = t.getPotentials()
        forces[idx, 0], forces[idx, 1], forces[idx, 2] = t.getForcesX(), t.getForcesY(), t.getForcesZ()
    tree.forEachLeaf(visit)
    acc = [oracle.accuracy(d[:, 4], pot)] + [oracle.accuracy(d[:, 5 + k], forces[:, k]) for k in range(3)]
    print(f"  pot rel L2 {acc[0]['rel_l2']:.2e}, forces {max(a['rel_l2'] for a in acc[1:]):.2e}", flush=True)
    assert acc[0]["rel_l2"] < 9e-3 and acc[0]["rms"] < 9e-3
    assert all(a["rel_l2"] < 9e-2 and a["rms"] < 9e-2 for a in acc[1:])
    energy = float(np.dot(pot, d[:, 3]))
    energy_d = float(np.dot(d[:, 4], d[:, 3]))
    assert abs(energy - energy_d) / abs(energy_d) < 9e-3
    g6 = oracle.ChebGlobal(6)
    ref = oracle.fmm_global(g6, d[:, 0], d[:, 1], d[:, 2], d[:, 3], loader.getCenterOfBox(), loader.getBoxWidth(), 4)
    assert _rel(pot, ref[0]) < 1e-12
    for k in range(3):
        assert _rel(forces[:, k], ref[1 + k]) < 1e-12


def test_dense_fmm_matches_oracle(oracle, fixture20k):
    """FChebDenseKernel (HFMM_M2L_DENSE) against the oracle's dense operators."""
    from scalfmm_amd import _abi
    from scalfmm_amd.fmm import Engine
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    e = Engine(5, w, c, 5)
    e.set_m2l_scheme(_abi.M2L_DENSE)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    ref = oracle.fmm_global(oracle.ChebGlobal(5, dense=True), x, y, z, q, c, w, 5)
    errs = [_rel(a, b) for a, b in zip(got, ref[:4])]
    print(f"  dense: rel L2 vs oracle {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-12
    e.close()


def test_utest_chebyshev_dense_kernel(oracle):
    """TestChebDenseKernel (UTests/utestChebyshev.cpp:33-48): order 6, the
    classic octree of 4 levels, DirectDouble.bfma, the FUKernelTester gates."""
    from scalfmm_amd import FChebDenseKernel, octree
    from scalfmm_amd.fma import FFmaGenericLoader
    loader = FFmaGenericLoader(os.path.join(ROOT, "tests", "golden", "DirectDouble.bfma"))
    d = loader.data
    n = loader.getNumberOfParticles()
    tree = octree.FOctree(4, 2, loader.getBoxWidth(), loader.getCenterOfBox())
    tree.insert_many(d[:, 0], d[:, 1], d[:, 2], d[:, 3], np.arange(n))
    octree.FFmmAlgorithm(tree, FChebDenseKernel(4, loader.getBoxWidth(), loader.getCenterOfBox(), 6)).execute()
    pot = np.zeros(n)
    fx = np.zeros(n)

    def visit(leaf):
        t = leaf.getTargets()
        pot[t.getIndexes()] = t.getPotentials()
        fx[t.getIndexes()] = t.getForcesX()
    tree.forEachLeaf(visit)
    assert oracle.accuracy(d[:, 4], pot)["rel_l2"] < 9e-3
    assert oracle.accuracy(d[:, 5], fx)["rel_l2"] < 9e-2
    ref = oracle.fmm_global(oracle.ChebGlobal(6, dense=True), d[:, 0], d[:, 1], d[:, 2], d[:, 3],
                            loader.getCenterOfBox(), loader.getBoxWidth(), 4)
    assert _rel(pot, ref[0]) < 1e-12 and _rel(fx, ref[1]) < 1e-12
