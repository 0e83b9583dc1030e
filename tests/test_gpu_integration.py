"""The integration shapes compiled and run on the GPU.

  - tests/adapter/adapter_demo.cpp: a C++ program with tree types of its own
    (the reference's particle- and cell-block byte layouts, a classic octree
    of indexed leaves) driving include/scalfmm_hip_algorithm.hpp:
    FGroupHipAlgorithm over the group tree (the 20k fixture, height 5, 250
    cells per block, order 6), split executions on a second tree, and
    hip_execute over the classic octree (order 7).  Bars: results within
    1e-12 of the oracle's FMM; the split executions equal one execution bit
    for bit; the level-3 cell blocks hold the oracle's expansions (M 1e-13,
    L 1e-12) under the oracle's cell keys;
  - tests/capi/capi_demo.c: the reference's C API (include/scalfmm_capi.h)
    from C, 10,000 particles of its own drand48 stream, order 7, height 5:
    potentials and forces within 1e-12 of the oracle's FMM.
Both are compiled here with g++ / gcc against the in-tree library.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _build(compiler, std, src, exe):
    from scalfmm_amd import _abi
    lib = _abi.LIB_PATH
    r = subprocess.run([compiler, std, "-O2", "-I", os.path.join(ROOT, "include"), src, "-o", str(exe),
                        "-L", os.path.dirname(lib), "-lscalfmm_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _run(cmd):
    print("running", " ".join(map(str, cmd)), flush=True)
    sys.stdout.flush()
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr


def test_cpp_adapter_program(oracle, tmp_path):
    exe = tmp_path / "adapter_demo"
    _build("g++", "-std=c++17", os.path.join(ROOT, "tests", "adapter", "adapter_demo.cpp"), exe)
    out = tmp_path / "adapter.bin"
    fixture = os.path.join(ROOT, "tests", "golden", "unitCubeRef20kDouble.bfma")
    _run([exe, fixture, out])
    raw = open(out, "rb").read()
    n, nc = np.frombuffer(raw[:16], np.int64)
    vals = np.frombuffer(raw[16:], np.uint8)
    off = 0

    def take(count, dtype=np.float64):
        nonlocal off
        a = vals[off:off + count * 8].view(dtype)
        off += count * 8
        return a
    full, split, classic = (take(4 * n).reshape(4, n) for _ in range(3))
    keys = take(nc, np.int64)
    M, L = take(nc * 216).reshape(nc, 216), take(nc * 216).reshape(nc, 216)
    f = oracle.read_bfma(fixture)
    d = f["data"]
    x, y, z, q = (d[:, k].copy() for k in range(4))
    ref6 = oracle.fmm(x, y, z, q, f["center"], f["width"], 5, 6)
    ref7 = oracle.fmm(x, y, z, q, f["center"], f["width"], 5, 7)
    for k in range(4):
        assert _rel(full[k], ref6[k]) <= 1e-12, k
        assert _rel(classic[k], ref7[k]) <= 1e-12, k
        assert np.array_equal(split[k], full[k]), k
    run = oracle.Run(oracle.Cheb(6), x, y, z, q, f["center"], f["width"], 5)  # one full execution
    assert np.array_equal(keys, run.tree.cell_keys(3))
    OM, OL = run.expansions(3)
    assert _rel(M, OM) <= 1e-13 and _rel(L, OL) <= 1e-12


def test_c_api_program(oracle, tmp_path):
    exe = tmp_path / "capi_demo"
    _build("gcc", "-std=c99", os.path.join(ROOT, "tests", "capi", "capi_demo.c"), exe)
    out = tmp_path / "capi.bin"
    _run([exe, 10000, out])
    raw = np.fromfile(out, np.uint8)
    n = int(raw[:8].view(np.int64)[0])
    v = raw[8:].view(np.float64)
    xyz = v[:3 * n].reshape(n, 3)
    q = v[3 * n:4 * n]
    pot = v[4 * n:5 * n]
    frc = v[5 * n:8 * n].reshape(n, 3)
    ref = oracle.fmm(xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), q.copy(), (0.5, 0.5, 0.5), 1.0, 5, 7)
    assert _rel(pot, ref[0]) <= 1e-12
    for k in range(3):
        assert _rel(frc[:, k], ref[1 + k]) <= 1e-12
