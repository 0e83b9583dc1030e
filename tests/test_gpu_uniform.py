"""FUnifKernel on the GPU (hfmm_set_m2l_scheme(HFMM_M2L_UNIFORM)): Lagrange
interpolation on the equispaced roots of Kernels/Uniform/FUnifRoots.hpp:
130-240 (P2M, M2M, L2L, L2P of FUnifInterpolator.hpp) and the M2L between the
equispaced nodes of target and source cells (FUnifM2LHandler.hpp computes the
same product by a zero-padded DFT), against the oracle's restatement.

Bars: FMM values within 1e-12 (relative L2) of the oracle's FUnifKernel FMM
for every pass subset and orders 5..8; the reference's unit test
(UTests/utestLagrange.cpp:44-212: order 6, DirectDouble.bfma, 4 levels) with
its gates; the reference's published errors of "Lagrange FMM ... ORDER 7"
(Data/ResultCompareKernels.txt) to the printed digits; a switch back to the
symmetric scheme restores the Chebyshev results bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _engine(h, width, center, order):
    from scalfmm_amd import _abi
    from scalfmm_amd.fmm import Engine
    e = Engine(h, width, center, order)
    e.set_m2l_scheme(_abi.M2L_UNIFORM)
    return e


@pytest.mark.parametrize("ops_name", ["all", "far", "near"])
def test_uniform_fmm_matches_oracle(oracle, fixture20k, ops_name):
    ops = {"all": oracle.ALL, "far": oracle.ALL & ~oracle.P2P, "near": oracle.P2P}[ops_name]
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    e = _engine(5, w, c, 5)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(ops)
    got = e.results()
    ref = oracle.fmm_global(oracle.ChebGlobal(5, uniform=True), x, y, z, q, c, w, 5, ops=ops)
    errs = [_rel(a, b) for a, b in zip(got, ref[:4])]
    print(f"  {ops_name}: rel L2 vs oracle {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-12
    e.close()


@pytest.mark.parametrize("order", [6, 7, 8])
def test_uniform_orders_match_oracle(oracle, order):
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(3000, 70 + order)
    c = (0.5, 0.5, 0.5)
    e = _engine(4, 1.0, c, order)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    ref = oracle.fmm_global(oracle.ChebGlobal(order, uniform=True), x, y, z, q, c, 1.0, 4)
    errs = [_rel(a, b) for a, b in zip(got, ref[:4])]
    print(f"  order {order}: rel L2 vs oracle {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-12
    e.close()


def test_uniform_levels_and_kept_cells(oracle, fixture20k):
    from scalfmm_amd import _abi
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    e = _engine(5, w, c, 5)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(oracle.ALL, (3, 5))
    ref = oracle.fmm_global(oracle.ChebGlobal(5, uniform=True), x, y, z, q, c, w, 5, levels=(3, 5))
    assert max(_rel(a, b) for a, b in zip(e.results(), ref[:4])) < 1e-12
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


def test_scheme_switch_restores_chebyshev(fixture20k):
    """Symmetric -> uniform -> symmetric: the Chebyshev tables come back and
    the symmetric results repeat bit for bit; uniform -> dense -> uniform
    likewise."""
    from scalfmm_amd import _abi
    from scalfmm_amd.fmm import Engine
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    e = Engine(5, w, c, 5)
    e.set_particles(d[:, 0], d[:, 1], d[:, 2], d[:, 3])
    e.build_tree()
    runs = {}
    for scheme in (_abi.M2L_SYMMETRIC, _abi.M2L_UNIFORM, _abi.M2L_DENSE, _abi.M2L_UNIFORM, _abi.M2L_SYMMETRIC):
        e.set_m2l_scheme(scheme)
        e.evaluate()
        r = [a.copy() for a in e.results()]
        if scheme in runs:
            for u, v in zip(r, runs[scheme]):
                np.testing.assert_array_equal(u, v)
        runs[scheme] = r
    assert _rel(runs[_abi.M2L_UNIFORM][0], runs[_abi.M2L_SYMMETRIC][0]) > 1e-9
    e.close()


def test_uniform_rejects_mixed_precision():
    from scalfmm_amd import FmmError, _abi
    from scalfmm_amd.fmm import Engine
    e = Engine(3, 1.0, (0.5, 0.5, 0.5), 5, precision=_abi.PRECISION_MIXED)
    with pytest.raises(FmmError, match="FP64 only"):
        e.set_m2l_scheme(_abi.M2L_UNIFORM)
    e.close()


def test_published_lagrange_errors(oracle, fixture20k):
    """ResultCompareKernels.txt, "Lagrange FMM ... ORDER 7": 20k fixture,
    height 5; the errors against the fixture's direct sums to the printed
    digits."""
    from scalfmm_amd import FUnifKernel
    d = fixture20k["data"]
    k = FUnifKernel(5, fixture20k["width"], fixture20k["center"], 7)
    e = k.engine
    e.set_particles(d[:, 0], d[:, 1], d[:, 2], d[:, 3])
    e.build_tree()
    e.evaluate()
    got = e.results()
    published = [2.11079e-08, 2.23423e-06, 2.36648e-06, 2.36806e-06]
    for j in range(4):
        err = oracle.accuracy(d[:, 4 + j], got[j])["rel_l2"]
        print(f"  component {j}: {err:.5e} (published {published[j]:.5e})", flush=True)
        assert abs(err - published[j]) <= 5e-6 * published[j]
    e.close()


def test_utest_lagrange_unif_kernel(oracle):
    """TestUnifKernel: FUnifKernel of order 6 over the classic octree of 4
    levels on DirectDouble.bfma, with the reference's gates."""
    from scalfmm_amd import FUnifKernel, octree
    from scalfmm_amd.fma import FFmaGenericLoader
    loader = FFmaGenericLoader(os.path.join(ROOT, "tests", "golden", "DirectDouble.bfma"))
    d = loader.data
    n = loader.getNumberOfParticles()
    tree = octree.FOctree(4, 2, loader.getBoxWidth(), loader.getCenterOfBox())
    tree.insert_many(d[:, 0], d[:, 1], d[:, 2], d[:, 3], np.arange(n))
    octree.FFmmAlgorithm(tree, FUnifKernel(4, loader.getBoxWidth(), loader.getCenterOfBox(), 6)).execute()
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
    ftot = np.linalg.norm(forces - d[:, 5:8]) / np.linalg.norm(d[:, 5:8])
    assert ftot < 9e-2
    energy = float(np.dot(pot, d[:, 3]))
    energy_d = float(np.dot(d[:, 4], d[:, 3]))
    assert abs(energy - energy_d) / abs(energy_d) < 9e-3
    ref = oracle.fmm_global(oracle.ChebGlobal(6, uniform=True), d[:, 0], d[:, 1], d[:, 2], d[:, 3],
                            loader.getCenterOfBox(), loader.getBoxWidth(), 4)
    assert _rel(pot, ref[0]) < 1e-12
    for k in range(3):
        assert _rel(forces[:, k], ref[1 + k]) < 1e-12


def test_scheme_at_creation(fixture20k):
    """A context created with the uniform scheme (hfmm_params.m2l_scheme) skips
    the symmetric compression: its results equal a symmetric context switched
    to the uniform scheme bit for bit; the 16 symmetric operators it compresses
    on demand equal the symmetric context's; switched to the symmetric scheme
    it gives that context's results bit for bit."""
    from scalfmm_amd import _abi
    from scalfmm_amd.fmm import Engine
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    engines = []
    for created in (_abi.M2L_SYMMETRIC, _abi.M2L_UNIFORM):
        e = Engine(5, w, c, 5, m2l_scheme=created)
        e.set_particles(d[:, 0], d[:, 1], d[:, 2], d[:, 3])
        e.build_tree()
        engines.append(e)
    sym, uni = engines
    sym.evaluate()
    ref_sym = [a.copy() for a in sym.results()]
    sym.set_m2l_scheme(_abi.M2L_UNIFORM)
    sym.evaluate()
    uni.evaluate()
    for a, b in zip(uni.results(), sym.results()):
        np.testing.assert_array_equal(a, b)
    for idx in (269, 276, 342):  # canonical positions (2,0,0), (2,1,0), (3,3,3)
        Ua, Va = sym.operator(idx)
        Ub, Vb = uni.operator(idx)
        np.testing.assert_array_equal(Ua, Ub)
        np.testing.assert_array_equal(Va, Vb)
    uni.set_m2l_scheme(_abi.M2L_SYMMETRIC)
    uni.evaluate()
    for a, b in zip(uni.results(), ref_sym):
        np.testing.assert_array_equal(a, b)
    for e in engines:
        e.close()


def test_scheme_at_creation_rejects_mixed():
    from scalfmm_amd import FmmError, _abi
    from scalfmm_amd.fmm import Engine
    with pytest.raises(FmmError):
        Engine(3, 1.0, (0.5, 0.5, 0.5), 5, precision=_abi.PRECISION_MIXED, m2l_scheme=_abi.M2L_UNIFORM)
