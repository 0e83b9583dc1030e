"""GPU parity of the mixed-precision form (HFMM_PRECISION_MIXED) through the C-ABI.

The mixed form changes only the arithmetic of M2L (bfloat16 matrix cores on
three-part operand splits, FP32 accumulation) and P2P (FP32 pair terms,
compensated per-target sums); the tree, the lists and the other passes are the
FP64 ones.  Stated tolerances:
  - against the FP64 oracle on the same inputs: relative L2 <= 1e-6 for the
    potential and <= 3e-6 for each force component (24-bit operands: a few
    units of 2^-24 per interaction, accumulated over ~190 M2L terms);
  - against FP64 direct sums: the FP64 form's own error plus that bound, and
    <= 1e-6 for the potential at order 8 (the mixed form's accuracy target).
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


def _mixed(h, p, width=1.0, center=(0.5, 0.5, 0.5)):
    from scalfmm_amd import PRECISION_MIXED, Engine
    return Engine(h, width, center, p, 0.0, 0, PRECISION_MIXED)


def _check(got, ref, pot_tol=POT_TOL, force_tol=FORCE_TOL):
    errs = [_rel(a, b) for a, b in zip(got, ref[:4])]
    _say(f"  rel L2 vs FP64: pot {errs[0]:.3e} fx {errs[1]:.3e} fy {errs[2]:.3e} fz {errs[3]:.3e}")
    assert errs[0] <= pot_tol, errs
    assert max(errs[1:]) <= force_tol, errs
    return errs


def test_mfma_bf16_layout():
    e = _mixed(3, 5)
    D = e.selftest_mfma_bf16()
    i = np.arange(16)[:, None]
    k = np.arange(32)[None, :]
    A = (7 * i + 3 * k) % 11 - 5
    B = (5 * np.arange(32)[:, None] + 2 * np.arange(16)[None, :]) % 13 - 6
    assert np.array_equal(D, (A @ B).astype(np.float32))


@pytest.mark.parametrize("ops_name", ["p2p", "far", "all"])
def test_mixed_fixture_order7(oracle, fixture20k, ops_name):
    from scalfmm_amd import FFmmFarField, FFmmNearAndFarFields, FFmmP2P
    ops = {"p2p": FFmmP2P, "far": FFmmFarField, "all": FFmmNearAndFarFields}[ops_name]
    f = fixture20k
    d = f["data"]
    x, y, z, q = (d[:, k].copy() for k in range(4))
    _say(f"mixed fixture20k h5 p7 ops={ops}")
    e = _mixed(5, 7)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(ops)
    ref = oracle.fmm(x, y, z, q, f["center"], f["width"], 5, 7, ops=ops)
    _check(e.results(), ref)


def test_mixed_fixture_vs_direct(oracle, fixture20k):
    """The fixture's FP64 direct results: the mixed error stays within the
    FP64 form's error plus the mixed bound."""
    from test_oracle_fmm import PUBLISHED
    f = fixture20k
    d = f["data"]
    e = _mixed(5, 7)
    e.set_particles(*(d[:, k].copy() for k in range(4)))
    e.build_tree()
    e.evaluate()
    for name, col, got in zip(("pot", "fx", "fy", "fz"), range(4, 8), e.results()):
        acc = oracle.accuracy(d[:, col], got)
        bound = PUBLISHED[name][0] + (POT_TOL if name == "pot" else FORCE_TOL)
        _say(f"  {name}: rel L2 vs direct {acc['rel_l2']:.3e} (FP64 form {PUBLISHED[name][0]:.3e})")
        assert acc["rel_l2"] <= bound, (name, acc)


@pytest.mark.parametrize("order,h,n", [(5, 5, 20000), (6, 4, 8000), (8, 4, 6000), (8, 5, 40000)])
def test_mixed_uniform_orders(oracle, order, h, n):
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(n, 200 + order)
    _say(f"mixed uniform n={n} h={h} p={order}")
    e = _mixed(h, order)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, order)
    _check(e.results(), ref)


@pytest.fixture(scope="module")
def plummer100k(oracle):
    """100k Plummer particles, height 6: core leaves hold thousands of particles."""
    from scalfmm_amd.distributions import plummer
    n, h, p = 100_000, 6, 8
    x, y, z, q = plummer(n, 11)
    return {"n": n, "h": h, "p": p, "xyzq": (x, y, z, q)}


@pytest.mark.parametrize("ops_name", ["p2p", "far", "all"])
def test_mixed_plummer(oracle, plummer100k, ops_name):
    from scalfmm_amd import FFmmFarField, FFmmNearAndFarFields, FFmmP2P
    ops = {"p2p": FFmmP2P, "far": FFmmFarField, "all": FFmmNearAndFarFields}[ops_name]
    d = plummer100k
    x, y, z, q = d["xyzq"]
    _say(f"mixed plummer n={d['n']} h={d['h']} p={d['p']} ops={ops}")
    e = _mixed(d["h"], d["p"])
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(ops)
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, d["h"], d["p"], ops=ops)
    _check(got, ref)
    if ops_name == "all":
        t = np.sort(np.random.default_rng(3).choice(d["n"], 500, replace=False))
        dp = oracle.direct(x, y, z, q, targets=t)[0]
        err = _rel(got[0][t], dp)
        _say(f"  pot rel L2 vs direct on {len(t)} targets: {err:.3e} (FP64 form {_rel(ref[0][t], dp):.3e})")
        assert err <= 1e-6


def test_mixed_repeat_is_deterministic():
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(30000, 9)
    e = _mixed(5, 6)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    a = e.results()
    e.build_tree()
    e.evaluate()
    b = e.results()
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_mixed_matches_fp64_engine_lists():
    """Same tree and lists as the FP64 form (the precision mode changes arithmetic only)."""
    from scalfmm_amd import Engine
    from scalfmm_amd.distributions import plummer
    x, y, z, q = plummer(20000, 4)
    a = _mixed(5, 5)
    b = Engine(5, 1.0, (0.5, 0.5, 0.5), 5)
    for e in (a, b):
        e.set_particles(x, y, z, q)
        e.build_tree()
    assert np.array_equal(a.sorted_keys(), b.sorted_keys())
    for level in range(2, 5):
        for u, v in zip(a.m2l_list(level), b.m2l_list(level)):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("order,h,n,variant", [(5, 5, 20000, 0), (5, 5, 20000, 1), (5, 5, 20000, 2),
                                               (5, 5, 20000, 3), (6, 4, 8000, 0), (6, 4, 8000, 1),
                                               (8, 4, 6000, 3), (7, 5, 30000, 3)])
def test_mixed_m2l_variants(oracle, order, h, n, variant):
    """Every M2L tile shape of the mixed form against the FP64 oracle, and the
    wave-level shapes against the four-wave one (same products, other
    accumulation order: within the mixed tolerance)."""
    from scalfmm_amd import FFmmFarField
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(n, 300 + order)
    _say(f"mixed M2L variant {variant}: n={n} h={h} p={order}")
    got = {}
    for v in sorted({0, variant}):
        e = _mixed(h, order)
        e.set_m2l_variant(v)
        e.set_particles(x, y, z, q)
        e.build_tree()
        e.evaluate(FFmmFarField)
        assert e.m2l_variant()[0] == v
        got[v] = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, order, ops=FFmmFarField)
    _check(got[variant], ref)
    for a, b in zip(got[variant], got[0]):
        assert _rel(a, b) <= 1e-6


def test_mixed_m2l_autotune():
    """Variant -1 times the available shapes once and keeps the fastest."""
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(30000, 12)
    e = _mixed(5, 5)
    assert e.m2l_variant()[0] == -1
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    v, times = e.m2l_variant()
    _say(f"  autotune picked {v}; times {times}")
    assert v in (0, 1, 2, 3)
    assert all(t > 0 for t in times)
    assert times[v] == min(times)
    a = e.results()
    e.evaluate()
    assert all(np.array_equal(u, w) for u, w in zip(a, e.results()))
    with pytest.raises(Exception):
        _mixed(7, 7).set_m2l_variant(1)  # the wave-level shapes stop at order 6


# HFMM_PRECISION_MIXED_HILO: two-part (hi/lo) bfloat16 splits in the M2L, three
# partial products; each product carries about 2^-16 relative error, so the bars
# are those of 16-bit operands: the far field within 2e-5 (potential) and 6e-5
# (forces) of the FP64 oracle, the full FMM within 1e-5 / 3e-5.
HILO_FAR = (2e-5, 6e-5)
HILO_ALL = (1e-5, 3e-5)


@pytest.mark.parametrize("order,h,n,variant", [(5, 5, 20000, 0), (5, 5, 20000, 1), (5, 5, 20000, 2),
                                               (5, 5, 20000, 3), (6, 4, 8000, 1), (8, 4, 6000, 0),
                                               (8, 4, 6000, 3)])
def test_hilo_m2l_variants(oracle, order, h, n, variant):
    from scalfmm_amd import PRECISION_MIXED_HILO, Engine, FFmmFarField, FFmmNearAndFarFields
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(n, 300 + order)
    _say(f"hi/lo M2L variant {variant}: n={n} h={h} p={order}")
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), order, 0.0, 0, PRECISION_MIXED_HILO)
    e.set_m2l_variant(variant)
    e.set_particles(x, y, z, q)
    e.build_tree()
    for ops, tol in ((FFmmFarField, HILO_FAR), (FFmmNearAndFarFields, HILO_ALL)):
        e.evaluate(ops)
        ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, order, ops=ops)
        _check(e.results(), ref, *tol)
    e.close()


def test_hilo_plummer_order5_vs_direct(oracle, plummer100k):
    """Order 5 on clustered particles: the hi/lo form's error against a direct
    sum is the FP64 form's (the interpolation error, ~1e-4) to within 1%."""
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED_HILO, Engine
    x, y, z, q = plummer100k["xyzq"]
    t = np.sort(np.random.default_rng(8).choice(len(x), 2000, replace=False))
    errs = {}
    for name, prec in (("fp64", PRECISION_FP64), ("hilo", PRECISION_MIXED_HILO)):
        e = Engine(6, 1.0, (0.5, 0.5, 0.5), 5, 0.0, 0, prec)
        e.set_particles(x, y, z, q)
        e.build_tree()
        e.evaluate()
        pot = e.results()[0]
        d = e.direct(t)[0]
        errs[name] = _rel(pot[t], d)
        e.close()
    _say(f"  pot rel L2 vs direct: fp64 {errs['fp64']:.3e} hilo {errs['hilo']:.3e}")
    assert errs["hilo"] <= 1.01 * errs["fp64"]
