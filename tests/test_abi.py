"""CPU checks of the C-ABI library: it loads without a GPU and exports every
symbol include/scalfmm_hip.h declares; host-side helpers agree with libc /
the reference conventions."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "scalfmm_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|int|const char\*)\s+(hfmm_\w+)\(", src, flags=re.M)))


def test_library_exports_header():
    from scalfmm_amd import _abi
    L = _abi.lib()
    names = _declared()
    assert names, "no declarations parsed"
    assert sorted(_abi.EXPORTS) == names
    for n in names:
        assert hasattr(L, n)
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported
    assert L.hfmm_version().startswith(b"scalfmm_amd")


def test_null_context_errors():
    from scalfmm_amd import _abi
    L = _abi.lib()
    assert L.hfmm_build_tree(None) == _abi.ERR_ARG
    assert L.hfmm_evaluate(None, 63) == _abi.ERR_ARG
    assert L.hfmm_create(None, None) == _abi.ERR_ARG
    assert L.hfmm_last_error(None) == b"null context"


def test_create_rejects_bad_params():
    from scalfmm_amd import _abi
    L = _abi.lib()
    for order, height, width in ((4, 5, 1.0), (9, 5, 1.0), (7, 2, 1.0), (7, 21, 1.0), (7, 5, 0.0), (7, 5, -1.0)):
        p = _abi.Params(tree_height=height, order=order, box_width=width, epsilon=0.0, precision=0, device=0)
        h = ctypes.c_void_p()
        assert L.hfmm_create(ctypes.byref(p), ctypes.byref(h)) == _abi.ERR_ARG


def test_batch_height_rule_and_argument_checks():
    """Default set height: the smallest h in [3, 20] with n <= 32 * 8^(h-1);
    hfmm_batch_create rejects bad set tables before touching a device."""
    from scalfmm_amd import _abi
    L = _abi.lib()
    for n, h in ((1, 3), (1000, 3), (2048, 3), (2049, 4), (16384, 4), (16385, 5), (64000, 5), (131072, 5),
                 (131073, 6), (10 ** 7, 8)):
        assert L.hfmm_batch_default_height(n) == h, n
    p = _abi.Params(tree_height=0, order=5, box_width=1.0, epsilon=0.0, precision=0, device=0)
    h = ctypes.c_void_p()
    i64 = ctypes.c_int64

    def create(off, heights=None, nrhs=1):
        o = (i64 * len(off))(*off)
        hh = (ctypes.c_int32 * len(heights))(*heights) if heights else None
        return L.hfmm_batch_create(ctypes.byref(p), len(off) - 1, o, hh, nrhs, ctypes.byref(h))

    assert create([1, 5]) == _abi.ERR_ARG           # must start at 0
    assert create([0, 5, 3]) == _abi.ERR_ARG        # decreasing
    assert create([0, 0, 0]) == _abi.ERR_ARG        # no particles
    assert create([0, 5], heights=[2]) == _abi.ERR_ARG
    assert create([0, 5], heights=[21]) == _abi.ERR_ARG
    assert create([0, 5], nrhs=0) == _abi.ERR_ARG
    assert create([0, 5], nrhs=_abi.MAX_RHS + 1) == _abi.ERR_ARG
    assert L.hfmm_batch_build(None) == _abi.ERR_ARG
    assert L.hfmm_batch_last_error(None) == b"null batch"


def test_drand48_matches_libc():
    from scalfmm_amd.distributions import Drand48
    libc = ctypes.CDLL(None)
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    libc.srand48(12345)
    ref = np.array([libc.drand48() for _ in range(70000)])
    g = Drand48(12345)
    got = np.concatenate([g.draw(1000), g.draw(69000)])
    assert np.array_equal(ref, got)


def test_plummer_matches_sequential_rejection():
    """The vectorised Plummer sampler consumes draws like the sequential loop
    of Src/Files/FGenerateDistribution.hpp:319-331."""
    from scalfmm_amd.distributions import Drand48, plummer
    n, R = 500, 3.0
    g = Drand48(9)
    uv = g.draw(2 * n).reshape(n, 2)
    radii = []
    while len(radii) < n:
        u = g.draw(1)[0] ** (2.0 / 3.0)
        r = (u / (1.0 - u)) ** 0.5
        if r <= R:
            radii.append(r)
    q = 2.0 * g.draw(n) - 1.0
    theta = 2 * np.pi * uv[:, 0]
    phi = np.arccos(2 * uv[:, 1] - 1)
    xs = 0.5 + np.cos(theta) * np.sin(phi) * np.array(radii) / (2 * R)
    x, y, z, qq = plummer(n, 9, R)
    assert np.allclose(x, xs, rtol=0, atol=1e-15)
    assert np.array_equal(qq, q)
