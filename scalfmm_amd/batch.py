"""Batches of independent particle sets evaluated by one HIP graph launch.

    b = BatchEngine(order, set_sizes, nrhs=8, box_width=1.0, box_center=(.5, .5, .5))
    b.set_particles(x, y, z, q)    # all sets concatenated; q of shape (nrhs, n)
    b.build()
    b.evaluate()                   # first call: passes run and are captured
    b.evaluate()                   # later calls: one graph launch
    pot, fx, fy, fz = b.results()  # (nrhs, n) each, input order

Each set lives in its own copy of the box; its tree height is the smallest h
in [3, 20] with at most 32 particles per leaf on average, unless `heights`
gives one per set.  Sets of equal height share one device tree.  The
reference evaluates one FGroupTree per set and up to NVALS right-hand sides
per run (FChebSymKernel.hpp:45-488); it has no batch entry point.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .fmm import NEAR_AND_FAR, FmmError, _dp


def default_height(n: int) -> int:
    return int(_abi.lib().hfmm_batch_default_height(int(n)))


class BatchEngine:
    def __init__(self, order, set_sizes, nrhs=1, box_width=1.0, box_center=(0.5, 0.5, 0.5), heights=None,
                 epsilon=0.0, device=0, precision=_abi.PRECISION_FP64):
        self._L = _abi.lib()
        sizes = np.asarray(set_sizes, dtype=np.int64)
        self.offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64))
        self.nsets = len(sizes)
        self.n = int(self.offsets[-1])
        self.nrhs = int(nrhs)
        self.order = int(order)
        p = _abi.Params()
        p.tree_height = 0
        p.order = int(order)
        p.box_width = float(box_width)
        for d in range(3):
            p.box_center[d] = float(box_center[d])
        p.epsilon = float(epsilon or 0.0)
        p.precision = int(precision)
        p.device = int(device)
        hp = None
        if heights is not None:
            hh = np.ascontiguousarray(heights, dtype=np.int32)
            if len(hh) != self.nsets:
                raise ValueError("one height per set")
            hp = hh.ctypes.data_as(C.POINTER(C.c_int32))
        self.heights = (np.asarray(heights, dtype=np.int32) if heights is not None
                        else np.array([default_height(s) for s in sizes], dtype=np.int32))
        h = C.c_void_p()
        rc = self._L.hfmm_batch_create(C.byref(p), self.nsets, self.offsets.ctypes.data_as(C.POINTER(C.c_int64)),
                                       hp, self.nrhs, C.byref(h))
        if rc:
            raise FmmError(rc, "hfmm_batch_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.hfmm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise FmmError(rc, self._L.hfmm_batch_last_error(self.h).decode())

    def set_particles(self, x, y, z, q):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        qq = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(self.nrhs, -1))
        if any(len(a) != self.n for a in arrs) or qq.shape[1] != self.n:
            raise ValueError(f"expected {self.n} particles and ({self.nrhs}, {self.n}) charges")
        self._check(self._L.hfmm_batch_set_particles(self.h, *(_dp(a) for a in arrs), _dp(qq)))

    def build(self):
        self._check(self._L.hfmm_batch_build(self.h))

    def evaluate(self, ops=NEAR_AND_FAR):
        self._check(self._L.hfmm_batch_evaluate(self.h, int(ops)))

    def synchronize(self):
        self._check(self._L.hfmm_batch_synchronize(self.h))

    def results(self, fields=("pot", "fx", "fy", "fz")):
        """(nrhs, n) arrays in input order for the requested fields (None for the others)."""
        names = ("pot", "fx", "fy", "fz")
        out = [np.empty((self.nrhs, self.n)) if k in fields else None for k in names]
        self._check(self._L.hfmm_batch_get_results(self.h, *(_dp(a) if a is not None else None for a in out)))
        return tuple(out)

    def trees(self):
        """Views (Engine, not owning) of the per-height trees, increasing
        height: tree inspection, M2L counts and kernel profiling."""
        from .fmm import Engine
        out = []
        for i in range(self.info()["trees"]):
            h, hh, ns = C.c_void_p(), C.c_int32(), C.c_int32()
            self._check(self._L.hfmm_batch_tree(self.h, i, C.byref(h), C.byref(hh), C.byref(ns)))
            e = Engine._borrowed(h, hh.value, self.order, self)
            e.set_height = hh.value
            e.nsets = ns.value
            out.append(e)
        return out

    def info(self):
        nt, gr = C.c_int32(), C.c_int32()
        self._check(self._L.hfmm_batch_info(self.h, C.byref(nt), C.byref(gr)))
        return {"trees": nt.value, "graph_ready": bool(gr.value)}
