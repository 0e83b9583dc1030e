"""hfmm_direct_noalloc: the direct sum with its staging in the static per-device
array, called through the C-ABI, against hfmm_direct (per-call buffers) on the
same context: the same kernel, so the results must be equal bit for bit.  Two
calls in a row, the second with fewer targets, so a stale staging entry from
the first call would show; the maximum target count is enforced.
"""
import ctypes as C

import numpy as np
import pytest

from scalfmm_amd import _abi
from scalfmm_amd.distributions import plummer, uniform_cube

pytestmark = pytest.mark.gpu


def _engine(x, y, z, q, h=5, p=5):
    from scalfmm_amd import Engine
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p)
    e.set_particles(x, y, z, q)
    return e


def _err(lib, e):
    return lib.hfmm_last_error(e.h).decode()


def _call(lib, e, fn, targets):
    t = np.ascontiguousarray(targets, dtype=np.int64)
    out = [np.full(len(t), np.nan) for _ in range(4)]
    dp = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in out]
    rc = getattr(lib, fn)(e.h, len(t), t.ctypes.data_as(C.POINTER(C.c_int64)), *dp)
    return rc, out


def test_direct_noalloc_equals_direct_on_two_calls(oracle):
    lib = _abi.lib()
    x, y, z, q = plummer(200_000, 5)
    e = _engine(x, y, z, q)
    rng = np.random.default_rng(11)
    big = rng.choice(len(x), _abi.DIRECT_NOALLOC_MAX_TARGETS, replace=False)
    small = rng.choice(len(x), 777, replace=False)
    for targets in (big, small):
        print(f"direct_noalloc: n={len(x)} targets={len(targets)}", flush=True)
        rc0, ref = _call(lib, e, "hfmm_direct", targets)
        assert rc0 == 0, _err(lib, e)
        rc1, got = _call(lib, e, "hfmm_direct_noalloc", targets)
        assert rc1 == 0, _err(lib, e)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    # the oracle's direct sum on a few of the last call's targets
    xs, ys, zs, qs = (np.ascontiguousarray(a) for a in (x, y, z, q))
    pick = small[:16]
    want = oracle.direct(xs, ys, zs, qs, pick)
    for k in range(4):
        a, b = got[k][:16], want[k]
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), k


def test_direct_noalloc_rejects_too_many_targets_and_keeps_working():
    lib = _abi.lib()
    x, y, z, q = uniform_cube(3000, 2)
    e = _engine(x, y, z, q, h=4)
    n = _abi.DIRECT_NOALLOC_MAX_TARGETS + 1
    targets = np.arange(n, dtype=np.int64) % len(x)
    rc, _ = _call(lib, e, "hfmm_direct_noalloc", targets)
    assert rc == _abi.ERR_ARG
    assert "exceed" in _err(lib, e)
    rc, got = _call(lib, e, "hfmm_direct_noalloc", np.arange(len(x)))
    assert rc == 0
    rc, ref = _call(lib, e, "hfmm_direct", np.arange(len(x)))
    assert rc == 0
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_direct_noalloc_two_contexts_share_the_device_array():
    """Two contexts on one GPU, called one after the other, each get their own
    results (the second call rewrites the whole staging it reads)."""
    lib = _abi.lib()
    a = uniform_cube(4000, 3)
    b = plummer(5000, 4)
    ea, eb = _engine(*a, h=4), _engine(*b, h=4)
    ta, tb = np.arange(4000)[::3], np.arange(5000)[::7]
    _, ra = _call(lib, ea, "hfmm_direct", ta)
    _, rb = _call(lib, eb, "hfmm_direct", tb)
    for e, t, r in ((ea, ta, ra), (eb, tb, rb), (ea, ta, ra)):
        rc, got = _call(lib, e, "hfmm_direct_noalloc", t)
        assert rc == 0
        for u, v in zip(got, r):
            assert np.array_equal(u, v)
