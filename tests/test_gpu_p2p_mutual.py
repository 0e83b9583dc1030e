"""GPU parity of the mutual near field of the mixed forms.

Each unordered pair of two neighbour leaves with more than 64 particles is
evaluated once and added to both particles (the reference's P2POuter over the
positions below 13 plus its algorithms' opposite-direction calls,
FChebSymKernel.hpp:468-478, FGroupSeqAlgorithm.hpp:343-394).  Checks:
  - against the FP64 oracle's near field on the same particles: relative L2
    <= 1e-6 (potential) and <= 3e-6 (forces), the mixed form's stated bounds;
  - against the ordered-pair mixed kernel (hfmm_set_p2p_mutual(0)): the same
    bounds (both sum FP32 pair terms, in different orders);
  - repeated evaluations give the same bits (the source-side sums are added in
    a fixed order, no atomics);
  - leaves of ragged sizes (tails of the 64-source chunks and the 256-target
    work items, big next to small neighbours);
  - two particles at the same position: dropped pair, finite results.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POT_TOL = 1e-6
FORCE_TOL = 3e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _engine(h, p, precision=None):
    from scalfmm_amd import PRECISION_MIXED, Engine
    return Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, PRECISION_MIXED if precision is None else precision)


def _run(x, y, z, q, h, p, ops, mutual=True):
    e = _engine(h, p)
    e.set_p2p_mutual(mutual)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(ops)
    out = e.results()
    e.close()
    return out


def _check(got, ref, what):
    errs = [_rel(a, b) for a, b in zip(got, ref[:4])]
    _say(f"  {what}: pot {errs[0]:.3e} fx {errs[1]:.3e} fy {errs[2]:.3e} fz {errs[3]:.3e}")
    assert errs[0] <= POT_TOL, errs
    assert max(errs[1:]) <= FORCE_TOL, errs


@pytest.fixture(scope="module")
def plummer60k():
    from scalfmm_amd.distributions import plummer
    return plummer(60_000, 21)


def test_mutual_near_field_vs_oracle_and_ordered(oracle, plummer60k):
    from scalfmm_amd import FFmmP2P
    x, y, z, q = plummer60k
    h, p = 5, 5
    _say(f"mutual P2P plummer n={len(x)} h={h}")
    got = _run(x, y, z, q, h, p, FFmmP2P)
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p, ops=FFmmP2P)
    _check(got, ref, "mutual vs FP64 oracle")
    ordered = _run(x, y, z, q, h, p, FFmmP2P, mutual=False)
    _check(ordered, ref, "ordered vs FP64 oracle")
    _check(got, ordered, "mutual vs ordered")


def test_mutual_full_fmm_and_determinism(oracle, plummer60k):
    from scalfmm_amd import FFmmNearAndFarFields
    x, y, z, q = plummer60k
    h, p = 5, 7
    e = _engine(h, p)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(FFmmNearAndFarFields)
    a = [v.copy() for v in e.results()]
    e.build_tree()
    e.evaluate(FFmmNearAndFarFields)
    b = e.results()
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    e.close()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p)
    _check(a, ref, "full FMM, mutual near field vs FP64 oracle")


def test_mutual_ragged_leaves(oracle):
    """Big leaves of 65..1100 particles (chunk and work-item tails) beside
    small ones, at height 3 (4 x 4 x 4 leaves): every big-big pair, big-small
    pair and small-small pair kind occurs."""
    from scalfmm_amd import FFmmP2P
    rng = np.random.default_rng(7)
    sizes = [65, 64, 127, 128, 129, 255, 256, 257, 300, 511, 1100, 1, 0, 70, 200, 3]
    pts = []
    L = 4
    for i in range(L ** 3):
        m = sizes[i % len(sizes)] if (i * 7) % 5 else sizes[(i * 3) % len(sizes)]
        cx, cy, cz = i // 16, (i // 4) % 4, i % 4
        u = rng.random((m, 3))
        pts.append((np.array([cx, cy, cz]) + u) / L)
    P = np.concatenate(pts)
    x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
    q = rng.uniform(-1, 1, len(x))
    _say(f"mutual P2P ragged leaves n={len(x)} h=3")
    got = _run(x, y, z, q, 3, 5, FFmmP2P)
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, 3, 5, ops=FFmmP2P)
    _check(got, ref, "ragged, mutual vs FP64 oracle")


def test_mutual_coincident_particles_are_finite():
    """Two particles at one position in a big leaf and two in a small leaf:
    the fixed-point coordinates cannot separate them, so the mixed form drops
    their pair (the FP64 form divides by zero); every other particle matches
    the FP64 form."""
    from scalfmm_amd import FFmmP2P, PRECISION_FP64
    rng = np.random.default_rng(11)
    n_big = 300
    # a big leaf: 300 particles inside the leaf [0.25, 0.375)^3 of a height-4 tree
    P = np.concatenate([rng.random((n_big, 3)) * 0.12 + 0.252, rng.random((600, 3))])
    P[1] = P[0]
    far = np.argmax(np.linalg.norm(P[n_big:] - 0.3, axis=1)) + n_big
    P = np.concatenate([P, P[far:far + 1]])  # a copy of a particle in a sparse region
    x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
    q = rng.uniform(-1, 1, len(x))
    got = _run(x, y, z, q, 4, 5, FFmmP2P)
    e = _engine(4, 5, PRECISION_FP64)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(FFmmP2P)
    ref = e.results()
    e.close()
    dup = np.array([0, 1, far, len(x) - 1])
    for v in got:
        assert np.all(np.isfinite(v))
    keep = np.setdiff1d(np.arange(len(x)), dup)
    _check([v[keep] for v in got], [v[keep] for v in ref], "coincident pairs dropped, others vs FP64 form")
