"""The reference's particle files: FFmaGenericLoader / FFmaGenericWriter
(Src/Files/FFmaGenericLoader.hpp:34-850) over hfmm_fma_read / hfmm_fma_write.

    loader = FFmaGenericLoader("particles.bfma")
    loader.getNumberOfParticles(), loader.getBoxWidth(), loader.getCenterOfBox()
    data = loader.data          # (n, values per particle): x, y, z, q[, pot, fx, fy, fz]

    writer = FFmaGenericWriter("out.fma")
    writer.writeHeader(center, width, n, 8, 8)
    writer.writeArrayOfReal(data, 8, n)

".bfma" in the name selects the binary format, else ".fma" the text format;
the header stores half the box width (the loader doubles it back).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class FmaError(IOError):
    pass


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class FFmaGenericLoader:
    def __init__(self, filename: str):
        L = _abi.lib()
        n, nv = C.c_int64(), C.c_int32()
        w = C.c_double()
        ctr = np.zeros(3)
        rc = L.hfmm_fma_read(filename.encode(), C.byref(n), C.byref(nv), C.byref(w), _dp(ctr), None)
        if rc:
            raise FmaError(f"cannot read {filename} (status {rc})")
        self.data = np.empty((n.value, nv.value))
        rc = L.hfmm_fma_read(filename.encode(), C.byref(n), C.byref(nv), C.byref(w), _dp(ctr), _dp(self.data))
        if rc:
            raise FmaError(f"cannot read the particles of {filename} (status {rc})")
        self.filename = filename
        self.box_width = w.value
        self.center = ctr

    def getNumberOfParticles(self):
        return self.data.shape[0]

    def getBoxWidth(self):
        return self.box_width

    def getCenterOfBox(self):
        return tuple(self.center)

    def getNbRecordPerline(self):
        return self.data.shape[1]

    def positions(self):
        return tuple(self.data[:, d].copy() for d in range(3))

    def physical_values(self):
        return self.data[:, 3].copy()


class FFmaGenericWriter:
    def __init__(self, filename: str):
        if ".bfma" not in filename and ".fma" not in filename:
            raise FmaError("only .fma or .bfma files are written")
        self.filename = filename
        self._header = None

    def writeHeader(self, center, width, n, data_type=8, nb_data_per_record=4):
        if data_type != 8:
            raise FmaError("only FP64 records are written")
        self._header = (tuple(float(c) for c in center), float(width), int(n), int(nb_data_per_record))

    def writeArrayOfReal(self, data, nb_data=None, n=None):
        if self._header is None:
            raise FmaError("writeHeader first")
        center, width, hn, nv = self._header
        a = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)
        nv = int(nb_data or nv)
        n = int(hn if n is None else n)
        if a.size != n * nv:
            raise FmaError(f"expected {n} x {nv} values, got {a.size}")
        ctr = np.array(center, dtype=np.float64)
        rc = _abi.lib().hfmm_fma_write(self.filename.encode(), n, nv, width, _dp(ctr), _dp(a))
        if rc:
            raise FmaError(f"cannot write {self.filename} (status {rc})")
