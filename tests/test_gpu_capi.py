"""The reference's public C API for its Chebyshev kernel (include/scalfmm_capi.h)
on the GPU, called through ctypes as a C program calls it.

Bars: with particles as one set (sequential / multi_thread), potentials and
forces within 1e-12 of the oracle's FMM at order 7 (the API's fixed order) and
within the reference's unit-test gates of the 5k fixture's direct results;
results accumulate over executions; npart getters address particles by
index; moved positions take effect at the next execution; with the
source_target algorithm the targets receive the sources' field only (FMM
error against a direct sum over the sources, order 7)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHEBYSHEV, LAGRANGE = 1, 2
SEQUENTIAL, MULTI_THREAD, PERIODIC, SOURCE_TARGET = 0, 1, 2, 3
SOURCE, TARGET, BOTH = 0, 1, 2


class Desc(C.Structure):
    _fields_ = [(f"f{i}", C.c_void_p) for i in range(5)]


def _lib():
    from scalfmm_amd import _abi
    L = _abi.lib()
    dp, ip, i, h = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_void_p
    L.scalfmm_init.restype = h
    L.scalfmm_init.argtypes = [i, i]
    L.scalfmm_build_tree.argtypes = [h, i, C.c_double, dp, Desc, Desc]
    for name in ("scalfmm_tree_insert_particles",):
        getattr(L, name).argtypes = [h, i, dp, dp, dp, i]
    L.scalfmm_tree_insert_particles_xyz.argtypes = [h, i, dp, i]
    for name in ("scalfmm_set_physical_values", "scalfmm_get_physical_values", "scalfmm_get_potentials",
                 "scalfmm_set_potentials", "scalfmm_get_forces_xyz", "scalfmm_set_forces_xyz",
                 "scalfmm_get_positions_xyz", "scalfmm_add_to_positions_xyz", "scalfmm_set_positions_xyz"):
        getattr(L, name).argtypes = [h, i, dp, i]
    for name in ("scalfmm_get_potentials_npart", "scalfmm_get_forces_xyz_npart", "scalfmm_get_physical_values_npart",
                 "scalfmm_set_potentials_npart"):
        getattr(L, name).argtypes = [h, i, ip, dp, i]
    L.scalfmm_get_forces.argtypes = [h, i, dp, dp, dp, i]
    for name in ("scalfmm_execute_fmm", "scalfmm_execute_fmm_far_field", "scalfmm_update_tree"):
        getattr(L, name).argtypes = [h]
    L.scalfmm_dealloc_handle.argtypes = [h, C.c_void_p]
    L.scalfmm_get_nb_timers.argtypes = [h]
    L.scalfmm_get_timers.argtypes = [h, dp]
    L.scalfmm_algorithm_config.argtypes = [h, i]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _handle(L, algo, height, width, center):
    h = L.scalfmm_init(CHEBYSHEV, algo)
    assert h
    c = np.array(center, np.float64)
    L.scalfmm_build_tree(h, height, width, _p(c), Desc(), Desc())
    return h


def _fixture():
    from scalfmm_amd.fma import FFmaGenericLoader
    return FFmaGenericLoader(os.path.join(ROOT, "tests", "golden", "DirectDouble.bfma"))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("algo", [SEQUENTIAL, MULTI_THREAD])
def test_capi_both(oracle, algo):
    L = _lib()
    ld = _fixture()
    d = ld.data
    n = len(d)
    h = _handle(L, algo, 4, ld.getBoxWidth(), ld.getCenterOfBox())
    xyz = np.ascontiguousarray(d[:, :3])
    q = np.ascontiguousarray(d[:, 3])
    L.scalfmm_tree_insert_particles_xyz(h, n, _p(xyz), BOTH)
    L.scalfmm_set_physical_values(h, n, _p(q), BOTH)
    L.scalfmm_execute_fmm(h)
    pot = np.zeros(n)
    f = np.zeros(3 * n)
    L.scalfmm_get_potentials(h, n, _p(pot), BOTH)
    L.scalfmm_get_forces_xyz(h, n, _p(f), BOTH)
    f = f.reshape(n, 3)
    ref = oracle.fmm(d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(), q.copy(), ld.getCenterOfBox(),
                     ld.getBoxWidth(), 4, 7)
    assert _rel(pot, ref[0]) <= 1e-12
    for k in range(3):
        assert _rel(f[:, k], ref[1 + k]) <= 1e-12
    assert oracle.accuracy(d[:, 4], pot)["rel_l2"] < 9e-3
    nt = L.scalfmm_get_nb_timers(h)
    t = np.zeros(nt)
    L.scalfmm_get_timers(h, _p(t))
    assert nt == 9 and t[2] > 0 and np.all(t >= 0)
    # accumulation, reset, npart
    L.scalfmm_execute_fmm(h)
    pot2 = np.zeros(n)
    L.scalfmm_get_potentials(h, n, _p(pot2), BOTH)
    assert np.array_equal(pot2, 2 * pot)
    zero = np.zeros(n)
    L.scalfmm_set_potentials(h, n, _p(zero), BOTH)
    idx = np.array([17, 3, 4999, 0], np.int32)
    sub = np.zeros(4)
    L.scalfmm_get_potentials_npart(h, 4, idx.ctypes.data_as(C.POINTER(C.c_int)), _p(sub), BOTH)
    assert np.all(sub == 0)
    fsub = np.zeros(12)
    L.scalfmm_get_forces_xyz_npart(h, 4, idx.ctypes.data_as(C.POINTER(C.c_int)), _p(fsub), BOTH)
    assert np.array_equal(fsub.reshape(4, 3), 2 * f[idx])
    L.scalfmm_dealloc_handle(h, None)


def test_capi_moved_positions(oracle):
    """add_to_positions then execute equals a fresh run on the moved particles."""
    from scalfmm_amd.distributions import uniform_cube
    L = _lib()
    x, y, z, q = uniform_cube(6000, 61)
    xyz = np.ascontiguousarray(np.stack([x, y, z], 1) * 0.9 + 0.05)
    h = _handle(L, MULTI_THREAD, 4, 1.0, (0.5, 0.5, 0.5))
    L.scalfmm_tree_insert_particles_xyz(h, len(q), _p(xyz), BOTH)
    L.scalfmm_set_physical_values(h, len(q), _p(q), BOTH)
    dxyz = np.ascontiguousarray(np.random.default_rng(2).uniform(-0.04, 0.04, xyz.shape))
    L.scalfmm_add_to_positions_xyz(h, len(q), _p(dxyz), BOTH)
    L.scalfmm_update_tree(h)
    L.scalfmm_execute_fmm(h)
    pot = np.zeros(len(q))
    L.scalfmm_get_potentials(h, len(q), _p(pot), BOTH)
    moved = xyz + dxyz
    got_xyz = np.zeros(3 * len(q))
    L.scalfmm_get_positions_xyz(h, len(q), _p(got_xyz), BOTH)
    assert np.array_equal(got_xyz.reshape(-1, 3), moved)
    ref = oracle.fmm(moved[:, 0].copy(), moved[:, 1].copy(), moved[:, 2].copy(), q, (0.5, 0.5, 0.5), 1.0, 4, 7)
    assert _rel(pot, ref[0]) <= 1e-12
    L.scalfmm_dealloc_handle(h, None)


def test_capi_source_target():
    """Targets receive the potential and force (q_t grad phi) of the sources only."""
    from scalfmm_amd.distributions import uniform_cube
    L = _lib()
    xs, ys, zs, qs = uniform_cube(4000, 62)
    xt, yt, zt, qt = uniform_cube(2500, 63)
    h = _handle(L, SOURCE_TARGET, 5, 1.0, (0.5, 0.5, 0.5))
    S = np.ascontiguousarray(np.stack([xs, ys, zs], 1))
    T = np.ascontiguousarray(np.stack([xt, yt, zt], 1))
    L.scalfmm_tree_insert_particles_xyz(h, len(qs), _p(S), SOURCE)
    L.scalfmm_tree_insert_particles_xyz(h, len(qt), _p(T), TARGET)
    L.scalfmm_set_physical_values(h, len(qs), _p(qs), SOURCE)
    L.scalfmm_set_physical_values(h, len(qt), _p(qt), TARGET)
    L.scalfmm_execute_fmm(h)
    pot = np.zeros(len(qt))
    f = np.zeros(3 * len(qt))
    L.scalfmm_get_potentials(h, len(qt), _p(pot), TARGET)
    L.scalfmm_get_forces_xyz(h, len(qt), _p(f), TARGET)
    D = T[:, None, :] - S[None, :, :]
    r = np.sqrt((D ** 2).sum(-1))
    phi = (qs[None, :] / r).sum(1)
    grad = -(qs[None, :, None] * D / r[:, :, None] ** 3).sum(1)
    assert _rel(pot, phi) < 1e-5
    assert _rel(f.reshape(-1, 3), qt[:, None] * grad) < 1e-4
    spot = np.ones(len(qs))
    L.scalfmm_get_potentials(h, len(qs), _p(spot), SOURCE)
    assert np.all(spot == 0)  # sources receive nothing
    L.scalfmm_dealloc_handle(h, None)


def test_capi_unsupported():
    L = _lib()
    assert not L.scalfmm_init(LAGRANGE, SEQUENTIAL)
    assert not L.scalfmm_init(CHEBYSHEV, PERIODIC)
