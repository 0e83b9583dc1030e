"""The C API header is valid C and the library exports every function it
declares: a C program written against it compiles and links (no GPU)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "scalfmm_capi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:void|int|scalfmm_handle)\s+(scalfmm_\w+)\(", src, flags=re.M)))


def test_capi_exports_and_links(tmp_path):
    from scalfmm_amd import _abi
    lib = _abi.LIB_PATH
    names = _declared()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported, sorted(set(names) - exported)
    exe = tmp_path / "capi_demo"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "capi", "capi_demo.c"), "-o", str(exe),
                        "-L", os.path.dirname(lib), "-lscalfmm_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_adapter_header_compiles(tmp_path):
    """include/scalfmm_hip_algorithm.hpp with the demo's own tree types: a C++
    program compiles and links against it (no GPU; run by
    tests/test_gpu_integration.py)."""
    from scalfmm_amd import _abi
    lib = _abi.LIB_PATH
    exe = tmp_path / "adapter_demo"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "adapter", "adapter_demo.cpp"), "-o", str(exe),
                        "-L", os.path.dirname(lib), "-lscalfmm_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
