"""The reference's particle and operator files through the C-ABI (no GPU):
FFmaGenericLoader / FFmaGenericWriter formats and the operator cache file name."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_bfma_matches_fixture_reader(oracle):
    from scalfmm_amd.fma import FFmaGenericLoader
    for name in ("DirectDouble.bfma", "unitCubeRef20kDouble.bfma"):
        path = os.path.join(GOLDEN, name)
        ld = FFmaGenericLoader(path)
        ref = oracle.read_bfma(path)
        assert np.array_equal(ld.data, ref["data"])
        assert ld.getBoxWidth() == ref["width"]
        assert np.array_equal(ld.center, ref["center"])
        assert ld.getNbRecordPerline() == 8


def test_text_fma_known_answers():
    """Data/test20k.fma: x, y, z are the C library's rand()/RAND_MAX from the
    default seed, printed to 6 digits; q = 0.01; header half width 0.5."""
    from scalfmm_amd.fma import FFmaGenericLoader
    ld = FFmaGenericLoader(os.path.join(GOLDEN, "test20k.fma"))
    assert ld.getNumberOfParticles() == 20000 and ld.getNbRecordPerline() == 4
    assert ld.getBoxWidth() == 1.0 and ld.getCenterOfBox() == (0.5, 0.5, 0.5)
    libc = ctypes.CDLL(None)
    libc.srand(1)
    r = np.array([libc.rand() for _ in range(60000)], dtype=np.float64) / 2147483647.0
    assert np.abs(ld.data[:, :3].ravel() - r).max() < 5.1e-7
    assert np.all(ld.data[:, 3] == 0.01)


@pytest.mark.parametrize("ext", [".fma", ".bfma"])
def test_round_trip(tmp_path, ext):
    from scalfmm_amd.fma import FFmaGenericLoader, FFmaGenericWriter
    rng = np.random.default_rng(3)
    data = rng.normal(size=(257, 8)) * 10.0 ** rng.integers(-5, 5, size=(257, 8))
    path = str(tmp_path / ("p" + ext))
    w = FFmaGenericWriter(path)
    w.writeHeader((0.25, -1.5, 3.0), 2.5, 257, 8, 8)
    w.writeArrayOfReal(data, 8, 257)
    ld = FFmaGenericLoader(path)
    assert ld.getBoxWidth() == 2.5 and ld.getCenterOfBox() == (0.25, -1.5, 3.0)
    if ext == ".bfma":
        assert np.array_equal(ld.data, data)
    else:  # 10 significant digits, as the reference's writer prints
        assert np.max(np.abs(ld.data - data) / np.abs(data)) < 6e-10
        first = open(path).read().split("\n")[:2]
        assert first[0].split() == ["8", "8"] and first[1].split()[:2] == ["257", "1.25"]


def test_bad_files(tmp_path):
    from scalfmm_amd import _abi
    from scalfmm_amd.fma import FFmaGenericLoader, FmaError
    L = _abi.lib()
    n, nv = ctypes.c_int64(), ctypes.c_int32()
    assert L.hfmm_fma_read(b"particles.txt", ctypes.byref(n), ctypes.byref(nv), None, None, None) == _abi.ERR_ARG
    assert L.hfmm_fma_read(str(tmp_path / "none.bfma").encode(), ctypes.byref(n), ctypes.byref(nv), None, None,
                           None) == _abi.ERR_IO
    single = tmp_path / "single.bfma"  # element size 4: not FP64 (the reference exits)
    single.write_bytes(np.array([4, 4], np.uint32).tobytes() + np.array([1], np.int64).tobytes() +
                       np.zeros(4, np.float64).tobytes() + np.zeros(4, np.float32).tobytes())
    with pytest.raises(FmaError):
        FFmaGenericLoader(str(single))
    short = tmp_path / "short.fma"
    short.write_text("8 4\n3 0.5 0.5 0.5 0.5\n0.1 0.2 0.3 1.0\n")
    with pytest.raises(FmaError):
        FFmaGenericLoader(str(short))


def test_operator_file_name():
    """sym2l_d_o<ORDER>_e<epsilon>.bin with the stream formatting of epsilon
    (FChebSymM2LHandler.hpp:485-487)."""
    from scalfmm_amd import _abi
    L = _abi.lib()
    buf = ctypes.create_string_buffer(64)
    for order, eps, name in ((8, 0.0, "sym2l_d_o8_e1e-08.bin"), (5, 0.0, "sym2l_d_o5_e1e-05.bin"),
                             (7, 1e-7, "sym2l_d_o7_e1e-07.bin"), (6, 2.5e-4, "sym2l_d_o6_e0.00025.bin")):
        assert L.hfmm_operator_file_name(order, eps, buf, 64) == 0
        assert buf.value.decode() == name
    assert L.hfmm_operator_file_name(8, 0.0, buf, 5) == _abi.ERR_ARG
