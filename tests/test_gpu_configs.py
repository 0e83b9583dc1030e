"""Full-size GPU tests of the two large workloads.

Mixed-precision workload: 10,000,000 particles of a Plummer sphere (scale 1,
truncated at radius 10, rescaled into the unit cube; drand48 seed 3), charges
U[-1,1], height 8, orders 5 and 8.  Checks:
  - the device tree against the oracle's (sorted keys, permutation, cell keys
    of every level, leaf ranges, P2P lists and the M2L lists of every level),
    bit for bit;
  - potentials and forces against the device FP64 direct sum on 4,000 sampled
    targets: order 8 mixed potential <= 1e-6 and forces <= the FP64 form's
    error + 3e-6; order 5 in the two-part (hi/lo) form within 1% of the FP64
    form's error; the mixed forms against the FP64 form on every particle.

Batched workload: 4,096 sets of floor(1000 * 64^u) particles (u = drand48 seed
4), each uniform in its own unit cube, 8 charge vectors U[-1,1], order 5,
heights by the default rule, all sets and vectors in one graph launch.
Checks: graph launch = eager run bit for bit; 16 sampled sets x 8 vectors
against FP64 direct sums (the order-5 truncation bound); 4 sampled sets
against single-set contexts bit for bit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def plummer10m():
    from scalfmm_amd.distributions import plummer
    _say("plummer 10M: generating (seed 3)")
    return plummer(10_000_000, 3)


@pytest.fixture(scope="module")
def direct10m(plummer10m):
    """FP64 direct sums (device) on 4,000 sampled targets."""
    from scalfmm_amd import Engine
    x, y, z, q = plummer10m
    t = np.sort(np.random.default_rng(5).choice(len(x), 4000, replace=False))
    e = Engine(8, 1.0, (0.5, 0.5, 0.5), 5)
    e.set_particles(x, y, z, q)
    _say("plummer 10M: FP64 direct sums on 4000 targets")
    ref = e.direct(t)
    e.close()
    return t, ref


def _evaluate(x, y, z, q, h, p, precision):
    from scalfmm_amd import Engine
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, precision)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    out = e.results()
    e.close()
    return out


def test_plummer10m_tree_bit_exact(oracle, plummer10m):
    from scalfmm_amd import PRECISION_MIXED, Engine
    x, y, z, q = plummer10m
    h = 8
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), 8, 0.0, 0, PRECISION_MIXED)
    e.set_particles(x, y, z, q)
    _say("plummer 10M h8: device tree")
    e.build_tree()
    keys, perm = e.sorted_keys()
    _say("plummer 10M h8: oracle tree")
    t = oracle.Tree(x, y, z, [0.5, 0.5, 0.5], 1.0, h)
    okeys, operm = t.sorted()
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)
    for level in range(2, h):
        assert np.array_equal(e.cell_keys(level), t.cell_keys(level)), level
    s, cnt = e.leaf_ranges()
    os_, oc = t.leaf_ranges()
    assert np.array_equal(s, os_) and np.array_equal(cnt, oc)
    _say(f"  {len(cnt)} leaves, largest {int(cnt.max())} particles")
    for a, b in zip(e.p2p_list(), t.p2p_list()):
        assert np.array_equal(a, b)
    for level in range(2, h):
        for a, b in zip(e.m2l_list(level), t.m2l_list(level)):
            assert np.array_equal(a, b), level
    e.close()


def test_plummer10m_order8_mixed(plummer10m, direct10m):
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED
    x, y, z, q = plummer10m
    t, ref = direct10m
    _say("plummer 10M h8 p8: FP64 form")
    f64 = _evaluate(x, y, z, q, 8, 8, PRECISION_FP64)
    _say("plummer 10M h8 p8: mixed form")
    mix = _evaluate(x, y, z, q, 8, 8, PRECISION_MIXED)
    e64 = [_rel(f64[k][t], ref[k]) for k in range(4)]
    emx = [_rel(mix[k][t], ref[k]) for k in range(4)]
    _say(f"  vs direct: FP64 pot {e64[0]:.3e} F {max(e64[1:]):.3e}; mixed pot {emx[0]:.3e} F {max(emx[1:]):.3e}")
    assert emx[0] <= 1e-6
    assert e64[0] <= 1e-6
    for k in range(1, 4):
        assert emx[k] <= e64[k] + 3e-6, (k, emx, e64)
    d = [_rel(mix[k], f64[k]) for k in range(4)]
    _say(f"  mixed vs FP64 on every particle: pot {d[0]:.3e} F {max(d[1:]):.3e}")
    assert d[0] <= 1e-6 and max(d[1:]) <= 3e-6, d


def test_plummer10m_order5_hilo(plummer10m, direct10m):
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED_HILO
    x, y, z, q = plummer10m
    t, ref = direct10m
    _say("plummer 10M h8 p5: FP64 form")
    f64 = _evaluate(x, y, z, q, 8, 5, PRECISION_FP64)
    _say("plummer 10M h8 p5: hi/lo form")
    hl = _evaluate(x, y, z, q, 8, 5, PRECISION_MIXED_HILO)
    e64 = [_rel(f64[k][t], ref[k]) for k in range(4)]
    ehl = [_rel(hl[k][t], ref[k]) for k in range(4)]
    _say(f"  vs direct: FP64 {['%.4e' % v for v in e64]}; hi/lo {['%.4e' % v for v in ehl]}")
    for k in range(4):
        assert abs(ehl[k] - e64[k]) <= 0.01 * e64[k], (k, ehl, e64)


@pytest.fixture(scope="module")
def batch4096():
    from scalfmm_amd.distributions import batch_sets
    _say("batch: generating 4096 sets x 8 charge vectors (seed 4)")
    return batch_sets(4096, 8, 4)


def _batch_run(sets, precision):
    from scalfmm_amd import BatchEngine
    sizes, x, y, z, q = sets
    b = BatchEngine(5, sizes, nrhs=8, precision=precision)
    b.set_particles(x, y, z, q)
    b.build()
    b.evaluate()  # eager run + capture
    eager = b.results(fields=("pot", "fx"))
    b.evaluate()  # graph launch
    b.synchronize()
    graph = b.results(fields=("pot", "fx"))
    for k in range(2):
        assert np.array_equal(graph[k], eager[k]), k
    variant = {t.set_height: t.m2l_variant()[0] for t in b.trees()}
    info = [(t.set_height, t.nsets, t.n) for t in b.trees()]
    offsets = b.offsets.copy()
    for tr in b.trees():
        tr.close()
    b.close()
    return eager, variant, info, offsets


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_batch4096_rhs8(batch4096, mode):
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED, Engine, default_height
    sizes, x, y, z, q = batch4096
    precision = PRECISION_MIXED if mode == "mixed" else PRECISION_FP64
    _say(f"batch 4096 x 8 RHS, {mode}: build, eager run, graph launch")
    got, variant, info, off = _batch_run(batch4096, precision)
    _say(f"  trees (height, sets, particles): {info}")
    assert sum(s for _, s, _ in info) == 4096 and sum(n for _, _, n in info) == int(sizes.sum())
    rng = np.random.default_rng(17)
    errs = []
    ref_engine = Engine(3, 1.0, (0.5, 0.5, 0.5), 5)
    for s in np.sort(rng.choice(4096, 16, replace=False)):
        lo, hi = int(off[s]), int(off[s + 1])
        t = np.sort(rng.choice(hi - lo, min(200, hi - lo), replace=False))
        for r in range(8):
            ref_engine.set_particles(x[lo:hi], y[lo:hi], z[lo:hi], q[r, lo:hi])
            d = ref_engine.direct(t)
            errs.append(max(_rel(got[0][r, lo + t], d[0]), 0.0))
            assert _rel(got[1][r, lo + t], d[1]) <= 3e-2
    ref_engine.close()
    _say(f"  pot rel L2 vs direct over 16 sets x 8 RHS: median {np.median(errs):.3e} max {max(errs):.3e}")
    assert max(errs) <= 3e-3  # the order-5 truncation error (about 1e-3)
    for s in np.sort(rng.choice(4096, 4, replace=False)):
        lo, hi = int(off[s]), int(off[s + 1])
        h = default_height(hi - lo)
        for r in range(8):
            e = Engine(h, 1.0, (0.5, 0.5, 0.5), 5, 0.0, 0, precision)
            if mode == "mixed":
                e.set_m2l_variant(variant[h])
            e.set_particles(x[lo:hi], y[lo:hi], z[lo:hi], q[r, lo:hi])
            e.build_tree()
            e.evaluate()
            one = e.results()
            e.close()
            assert np.array_equal(got[0][r, lo:hi], one[0]), (s, r)
            assert np.array_equal(got[1][r, lo:hi], one[1]), (s, r)
