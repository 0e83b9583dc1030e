"""Several charge vectors in one evaluation (the reference's NVALS, FChebSymKernel.hpp:45 and
Kernels/P2P/FP2PMultiRhs.hpp): the near field of the leaves of at most 64 particles runs once for
all vectors, the pair geometry (separation, 1/r) once per pair, while the big leaves and the far
field run per vector.  Each vector's terms are summed exactly as in a single-vector evaluation, so
the bar is bit-equality with a context holding that vector alone, for every precision form, on a
Plummer set whose leaves take both paths, and for vector counts that exercise every chunk size
(8 = one chunk of 8; 7 = 4 + 2 + 1).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H, P = 5, 5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _precision(mode):
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED_HILO
    return {"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED, "hilo": PRECISION_MIXED_HILO}[mode]


def _particles(n, nrhs, seed):
    from scalfmm_amd.distributions import Drand48, plummer
    x, y, z, _ = plummer(n, seed)
    g = Drand48(seed + 100)
    q = 2.0 * g.draw(nrhs * n).reshape(nrhs, n) - 1.0
    return x, y, z, q


def _leaf_sizes(x, y, z):
    n = 1 << (H - 1)
    c = [np.clip((a * n).astype(np.int64), 0, n - 1) for a in (x, y, z)]
    _, cnt = np.unique((c[0] * n + c[1]) * n + c[2], return_counts=True)
    return cnt


def _alone(mode, x, y, z, qr):
    from scalfmm_amd import Engine
    e = Engine(H, 1.0, (0.5, 0.5, 0.5), P, 0.0, 0, _precision(mode))
    e.set_particles(x, y, z, qr)
    e.build_tree()
    e.evaluate()
    return e.results(), e.field()


@pytest.mark.parametrize("nrhs", [8, 7])
@pytest.mark.parametrize("mode", ["fp64", "mixed", "hilo"])
def test_vectors_together_equal_vectors_alone(mode, nrhs):
    from scalfmm_amd import Engine
    n = 30000
    x, y, z, q = _particles(n, nrhs, 11)
    sizes = _leaf_sizes(x, y, z)
    assert sizes.max() > 64 and sizes.min() <= 64  # both near-field paths run
    _say(f"{mode} nrhs={nrhs}: {n} Plummer particles, h{H} p{P}, leaves {len(sizes)} (largest {sizes.max()})")
    e = Engine(H, 1.0, (0.5, 0.5, 0.5), P, 0.0, 0, _precision(mode))
    e.set_particle_sets(np.array([0, n]), x, y, z, q)
    e.build_tree()
    e.evaluate()
    got, fld = e.results(), e.field()
    assert got[0].shape == (nrhs, n)
    for r in range(nrhs):
        ref, ref_fld = _alone(mode, x, y, z, q[r])
        for k in range(4):
            assert np.array_equal(got[k][r], ref[k]), (r, k)
        for k in range(3):
            assert np.array_equal(fld[k][r], ref_fld[k]), (r, k)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_near_field_only(mode):
    """The near field alone (P2P without the far field) with three vectors (chunks 2 + 1)."""
    from scalfmm_amd import Engine
    from scalfmm_amd._abi import P2P
    n, nrhs = 20000, 3
    x, y, z, q = _particles(n, nrhs, 12)
    e = Engine(H, 1.0, (0.5, 0.5, 0.5), P, 0.0, 0, _precision(mode))
    e.set_particle_sets(np.array([0, n]), x, y, z, q)
    e.build_tree()
    e.evaluate(P2P)
    got = e.results()
    for r in range(nrhs):
        a = Engine(H, 1.0, (0.5, 0.5, 0.5), P, 0.0, 0, _precision(mode))
        a.set_particles(x, y, z, q[r])
        a.build_tree()
        a.evaluate(P2P)
        ref = a.results()
        for k in range(4):
            assert np.array_equal(got[k][r], ref[k]), (r, k)


@pytest.mark.parametrize("scheme", ["global", "uniform"])
def test_other_m2l_schemes(scheme):
    """FChebKernel (one global SVD) and FUnifKernel (Lagrange basis) with three vectors: the near field
    of all vectors at once beside the scheme's far field, every vector bit-equal to a context holding it
    alone (order 5, FP64)."""
    from scalfmm_amd import Engine, _abi
    m2l = {"global": _abi.M2L_GLOBAL_SVD, "uniform": _abi.M2L_UNIFORM}[scheme]
    n, nrhs = 20000, 3
    x, y, z, q = _particles(n, nrhs, 13)

    def engine():
        e = Engine(H, 1.0, (0.5, 0.5, 0.5), P, 0.0, 0, _precision("fp64"))
        e.set_m2l_scheme(m2l)
        return e

    e = engine()
    e.set_particle_sets(np.array([0, n]), x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    _say(f"{scheme}: {n} Plummer particles, {nrhs} vectors")
    for r in range(nrhs):
        a = engine()
        a.set_particles(x, y, z, q[r])
        a.build_tree()
        a.evaluate()
        ref = a.results()
        for k in range(4):
            assert np.array_equal(got[k][r], ref[k]), (r, k)
