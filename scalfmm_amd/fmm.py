"""Host-side mirror of the reference's Chebyshev 1/r path over the HIP engine.

Names follow the reference (paths relative to its source tree):

  FChebSymKernel(tree_height, box_width, box_center, order, epsilon)
      Src/Kernels/Chebyshev/FChebSymKernel.hpp:110-140 (ORDER is a template
      parameter there, an argument here; epsilon defaults to 10^-order)
  FGroupTree(tree_height, box_width, box_center, x, y, z, q)
      Src/GroupTree/Core/FGroupTree.hpp:178-360 (built on the device)
  FGroupTaskAlgorithm(tree, kernel).execute(ops)
      Src/GroupTree/Core/FGroupTaskAlgorithm.hpp:82-111 and
      Src/Core/FCoreCommon.hpp:73-78, with the FFmmOperations bits

Errors the reference turns into aborts (a particle outside the box,
FCoordinateComputer.hpp:18) or exceptions (FChebSymM2LHandler.hpp:156-160)
raise FmmError here.  Results come back in the input order; force = q grad phi.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import FAR_FIELD, L2L, L2P, M2L, M2M, NEAR_AND_FAR, NEAR_FIELD, P2M, P2P  # noqa: F401

FFmmP2P, FFmmP2M, FFmmM2M, FFmmM2L, FFmmL2L, FFmmL2P = P2P, P2M, M2M, M2L, L2L, L2P
FFmmNearField, FFmmFarField, FFmmNearAndFarFields = NEAR_FIELD, FAR_FIELD, NEAR_AND_FAR


class FmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _ip32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Engine:
    """One device context: operators, particles, tree and expansions."""

    def __init__(self, tree_height, box_width, box_center, order, epsilon=0.0, device=0,
                 precision=_abi.PRECISION_FP64, m2l_scheme=_abi.M2L_SYMMETRIC):
        self._L = _abi.lib()
        p = _abi.Params()
        p.tree_height = int(tree_height)
        p.order = int(order)
        p.box_width = float(box_width)
        for d in range(3):
            p.box_center[d] = float(box_center[d])
        p.epsilon = float(epsilon or 0.0)
        p.precision = int(precision)
        p.device = int(device)
        p.m2l_scheme = int(m2l_scheme)
        h = C.c_void_p()
        rc = self._L.hfmm_create(C.byref(p), C.byref(h))
        if rc:
            raise FmmError(rc, "hfmm_create failed")
        self.h = h
        self.params = p
        self.n = 0
        self.nrhs = 1
        self.nsets = 1
        self.tree_stamp = 0  # counts tree builds (the autograd wrapper checks it)
        self.height = int(tree_height)
        self.order = int(order)
        self.nnodes = self.order ** 3

    @classmethod
    def _borrowed(cls, handle, tree_height, order, owner):
        """A view of a context owned by `owner` (a batch): inspection and
        profiling only; closing it does not destroy the context."""
        e = cls.__new__(cls)
        e._L = _abi.lib()
        e.h = handle
        e._owner = owner
        e.n = 0
        e.nrhs = e.nsets = 1
        e.tree_stamp = 0
        e.height = int(tree_height)
        e.order = int(order)
        e.nnodes = e.order ** 3
        ti = e.tree_info()
        e.n = int(ti.n_particles)
        e.height = int(ti.height)
        return e

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self.h = None
            return
        if getattr(self, "h", None):
            self._L.hfmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise FmmError(rc, self._L.hfmm_last_error(self.h).decode())

    # ---- data ---------------------------------------------------------------
    def set_particles(self, x, y, z, q):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q)]
        n = len(arrs[0])
        if any(len(a) != n for a in arrs):
            raise ValueError("x, y, z, q must have the same length")
        self._check(self._L.hfmm_set_particles(self.h, n, *(_dp(a) for a in arrs)))
        self.n = n
        self.nrhs = self.nsets = 1

    def set_particle_sets(self, offsets, x, y, z, q):
        """Independent particle sets (set s = particles offsets[s]..offsets[s+1]-1,
        each in its own copy of the box) with q of shape (nrhs, n) or (n,)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        n = len(arrs[0])
        qq = np.ascontiguousarray(np.atleast_2d(np.asarray(q, dtype=np.float64)))
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        if any(len(a) != n for a in arrs) or qq.shape[1] != n:
            raise ValueError("x, y, z and every charge vector must have the same length")
        self._check(self._L.hfmm_set_particle_sets(self.h, len(off) - 1, _ip64(off), qq.shape[0], n,
                                                   *(_dp(a) for a in arrs), _dp(qq)))
        self.n = n
        self.nrhs = qq.shape[0]
        self.nsets = len(off) - 1

    def set_particles_device(self, n, x_ptr, y_ptr, z_ptr, q_ptr):
        self._check(self._L.hfmm_set_particles_device(self.h, int(n), x_ptr, y_ptr, z_ptr, q_ptr))
        self.n = int(n)

    def set_charges(self, q):
        """New charges on the built tree (input order, (nrhs, n) or (n,)); the tree is kept."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.size != self.n * self.nrhs:
            raise ValueError("one charge per particle and charge vector")
        self._check(self._L.hfmm_set_charges(self.h, _dp(q)))

    def build_tree(self):
        self.tree_stamp += 1
        self._check(self._L.hfmm_build_tree(self.h))

    def evaluate(self, ops=NEAR_AND_FAR, levels=None):
        """Runs the selected passes; `levels` = (upper, lower) restricts the
        working levels as FAbstractAlgorithm::execute(ops, upper, lower)."""
        if levels is None:
            self._check(self._L.hfmm_evaluate(self.h, int(ops)))
        else:
            self._check(self._L.hfmm_evaluate_levels(self.h, int(ops), int(levels[0]), int(levels[1])))

    def set_cell_mode(self, mode):
        """_abi.CELLS_FRESH (each evaluation starts from zero expansions) or
        _abi.CELLS_KEPT (the passes add into the expansions of earlier
        evaluations, as the reference's tree cells keep them)."""
        self._check(self._L.hfmm_set_cell_mode(self.h, int(mode)))

    def set_m2l_scheme(self, scheme):
        """_abi.M2L_SYMMETRIC (FChebSymKernel), _abi.M2L_GLOBAL_SVD (FChebKernel:
        builds the global operators on first use), _abi.M2L_DENSE
        (FChebDenseKernel) or _abi.M2L_UNIFORM (FUnifKernel: the Lagrange basis
        on equispaced roots)."""
        self._check(self._L.hfmm_set_m2l_scheme(self.h, int(scheme)))

    def global_operators(self):
        """FChebKernel's operators: U, B (nnodes x k), C (343 x k x k with
        C[pos][i, j]), and the singular values of K_col and K_row."""
        k = C.c_int32()
        self._check(self._L.hfmm_get_global_operators(self.h, C.byref(k), None, None, None, None, None))
        k = k.value
        nn = self.nnodes
        U, B = np.empty(nn * k), np.empty(nn * k)
        Cm = np.empty(343 * k * k)
        S1, S2 = np.empty(nn), np.empty(nn)
        self._check(self._L.hfmm_get_global_operators(self.h, None, _dp(U), _dp(B), _dp(Cm), _dp(S1), _dp(S2)))
        return U.reshape(k, nn).T, B.reshape(k, nn).T, Cm.reshape(343, k, k).transpose(0, 2, 1), S1, S2

    def set_stream_overlap(self, on=True):
        """Run the P2P on a second stream beside P2M..L2L (default off)."""
        self._check(self._L.hfmm_set_stream_overlap(self.h, 1 if on else 0))

    def set_p2p_mutual(self, on=True):
        """Mixed forms: each unordered pair of two big neighbour leaves once (default on)."""
        self._check(self._L.hfmm_set_p2p_mutual(self.h, 1 if on else 0))

    def reset_cells(self):
        """tree.forEachCell(resetToInitialState) for the kept cells."""
        self._check(self._L.hfmm_reset_cells(self.h))

    def evaluate_particle_blocks(self, blocks, attributes, nvals, ops=NEAR_AND_FAR, upper=2, lower=None):
        """hfmm_evaluate_particle_blocks: `blocks` are the raw particle-block
        buffers of a group tree, `attributes` their attribute buffers (numpy
        arrays; the results are added to the latter)."""
        nb = len(blocks)
        if len(attributes) != nb:
            raise ValueError("one attributes buffer per block")
        for a in attributes:
            if a.dtype != np.float64 or not a.flags.c_contiguous or not a.flags.writeable:
                raise ValueError("attribute buffers must be writeable contiguous float64 arrays")
        bp = (C.c_void_p * max(nb, 1))(*[b.ctypes.data for b in blocks])
        ap = (C.POINTER(C.c_double) * max(nb, 1))(*[_dp(a) for a in attributes])
        lower = self.height if lower is None else int(lower)
        rc = self._L.hfmm_evaluate_particle_blocks(self.h, nb, bp, ap, int(nvals), int(ops), int(upper), lower)
        self._check(rc)
        n = int(sum(int(b[56:64].view(np.int64)[0]) for b in blocks))
        if n:
            self.n, self.nrhs, self.nsets = n, int(nvals), 1
            self.tree_stamp += 1

    def _cell_args(self, blocks, multipoles, locals_):
        nb = len(blocks)
        for arrs in (multipoles, locals_):
            if arrs is None:
                continue
            if len(arrs) != nb:
                raise ValueError("one expansion buffer per cell block")
            for a in arrs:
                if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous):
                    raise ValueError("expansion buffers must be contiguous float64 arrays")
        bp = (C.c_void_p * max(nb, 1))(*[b.ctypes.data for b in blocks])

        def ptrs(arrs):
            if arrs is None:
                return None
            return (C.POINTER(C.c_double) * max(nb, 1))(*[_dp(a) if a is not None else None for a in arrs])
        return nb, bp, ptrs(multipoles), ptrs(locals_)

    def get_cell_blocks(self, level, blocks, multipoles, locals_, nvals=1):
        """hfmm_get_cell_blocks: the expansions of `level` into the cell
        groups' multipole / local buffers (FChebCellPOD layout)."""
        nb, bp, mp, lp = self._cell_args(blocks, multipoles, locals_)
        self._check(self._L.hfmm_get_cell_blocks(self.h, int(level), nb, bp, mp, lp, int(nvals)))

    def set_cell_blocks(self, level, blocks, multipoles, locals_, nvals=1):
        """hfmm_set_cell_blocks: the cell groups' expansions into the kept cells."""
        nb, bp, mp, lp = self._cell_args(blocks, multipoles, locals_)
        self._check(self._L.hfmm_set_cell_blocks(self.h, int(level), nb, bp, mp, lp, int(nvals)))

    def synchronize(self):
        self._check(self._L.hfmm_synchronize(self.h))

    def _shape(self):
        return (self.n,) if self.nrhs == 1 else (self.nrhs, self.n)

    def results(self):
        """pot, fx, fy, fz in input order; (nrhs, n) arrays with several charge vectors."""
        out = [np.empty(self._shape()) for _ in range(4)]
        self._check(self._L.hfmm_get_results(self.h, *(_dp(a) for a in out)))
        return tuple(out)

    # ---- inspection -----------------------------------------------------------
    def field(self):
        """grad(phi) at every particle (input order): the force without the target charge."""
        out = [np.empty(self._shape()) for _ in range(3)]
        self._check(self._L.hfmm_get_field(self.h, *(_dp(a) for a in out)))
        return tuple(out)

    def tree_info(self):
        ti = _abi.TreeInfo()
        self._check(self._L.hfmm_get_tree_info(self.h, C.byref(ti)))
        return ti

    def sorted_keys(self):
        k = np.empty(self.n, np.int64)
        p = np.empty(self.n, np.int64)
        self._check(self._L.hfmm_get_sorted_keys(self.h, _ip64(k), _ip64(p)))
        return k, p

    def cell_keys(self, level):
        ti = self.tree_info()
        k = np.empty(max(ti.n_cells[level], 1), np.int64)
        self._check(self._L.hfmm_get_cell_keys(self.h, level, _ip64(k)))
        return k[: ti.n_cells[level]]

    def leaf_ranges(self):
        nl = self.tree_info().n_cells[self.height - 1]
        s = np.empty(nl, np.int64)
        c = np.empty(nl, np.int64)
        self._check(self._L.hfmm_get_leaf_ranges(self.h, _ip64(s), _ip64(c)))
        return s, c

    def p2p_list(self):
        ti = self.tree_info()
        nl = ti.n_cells[self.height - 1]
        off = np.empty(nl + 1, np.int64)
        nb = np.empty(max(ti.p2p_entries, 1), np.int64)
        ps = np.empty(max(ti.p2p_entries, 1), np.int32)
        self._check(self._L.hfmm_get_p2p_list(self.h, _ip64(off), _ip64(nb), _ip32(ps)))
        return off, nb[: ti.p2p_entries], ps[: ti.p2p_entries]

    def m2l_list(self, level):
        ti = self.tree_info()
        nc = ti.n_cells[level]
        m = ti.m2l_entries[level]
        off = np.empty(nc + 1, np.int64)
        sr = np.empty(max(m, 1), np.int64)
        ps = np.empty(max(m, 1), np.int32)
        self._check(self._L.hfmm_get_m2l_list(self.h, level, _ip64(off), _ip64(sr), _ip32(ps)))
        return off, sr[:m], ps[:m]

    def set_m2l_variant(self, v):
        """Mixed M2L tile shape (0, 1, 2, 3) or -1 to autotune at the next M2L."""
        self._check(self._L.hfmm_set_m2l_variant(self.h, int(v)))

    def m2l_variant(self):
        """(variant in use or -1, times per variant measured by the last autotune)."""
        v = C.c_int32()
        t = (C.c_float * 4)()
        self._check(self._L.hfmm_get_m2l_variant(self.h, C.byref(v), t))
        return v.value, [float(x) for x in t]

    def m2l_counts(self, level):
        """Interactions of a level per relative position (343 counts)."""
        out = np.zeros(343, np.int64)
        self._check(self._L.hfmm_get_m2l_counts(self.h, int(level), _ip64(out)))
        return out

    def save_operators(self, path):
        """The compressed M2L operators in the reference's cache file format."""
        self._check(self._L.hfmm_save_operators(self.h, str(path).encode()))

    def load_operators(self, path):
        """Replaces the compressed M2L operators by those of a cache file."""
        self._check(self._L.hfmm_load_operators(self.h, str(path).encode()))

    def operator(self, idx):
        r = C.c_int32()
        self._check(self._L.hfmm_get_operator(self.h, idx, C.byref(r), None))
        k = r.value
        if k == 0:
            return None, None
        K = np.empty(2 * k * self.nnodes)
        self._check(self._L.hfmm_get_operator(self.h, idx, C.byref(r), _dp(K)))
        U = K[: k * self.nnodes].reshape(k, self.nnodes).T
        V = K[k * self.nnodes:].reshape(k, self.nnodes).T
        return U, V

    def canonical_ranks(self):
        out = {}
        for i in range(2, 4):
            for j in range(0, i + 1):
                for k in range(0, j + 1):
                    r = C.c_int32()
                    self._check(self._L.hfmm_get_operator(self.h, (i + 3) * 49 + (j + 3) * 7 + (k + 3),
                                                          C.byref(r), None))
                    out[(i, j, k)] = r.value
        return out

    def expansions(self, level):
        nc = self.tree_info().n_cells[level]
        M = np.empty(max(nc, 1) * self.nnodes)
        Lx = np.empty(max(nc, 1) * self.nnodes)
        self._check(self._L.hfmm_get_expansions(self.h, level, _dp(M), _dp(Lx)))
        return M[: nc * self.nnodes].reshape(nc, -1), Lx[: nc * self.nnodes].reshape(nc, -1)

    # ---- timing ---------------------------------------------------------------
    def set_profiling(self, on=True):
        self._check(self._L.hfmm_set_profiling(self.h, 1 if on else 0))

    def set_profiling_filter(self, names=()):
        """Time only the named kernels (an empty list: every kernel)."""
        self._check(self._L.hfmm_set_profiling_filter(self.h, ",".join(names).encode()))

    def kernel_times(self):
        cnt = C.c_int32()
        buf = (_abi.KernelTime * 64)()
        self._check(self._L.hfmm_get_kernel_times(self.h, buf, 64, C.byref(cnt)))
        return {buf[i].name.decode(): (buf[i].total_ms, buf[i].launches) for i in range(min(cnt.value, 64))}

    def reset_kernel_times(self):
        self._check(self._L.hfmm_reset_kernel_times(self.h))

    def selftest_mfma_f64(self):
        out = np.empty(256)
        self._check(self._L.hfmm_selftest_mfma_f64(self.h, _dp(out)))
        return out.reshape(16, 16)

    def direct(self, targets):
        """FP64 direct sum on the device for the given input-order targets."""
        t = np.ascontiguousarray(targets, dtype=np.int64)
        out = [np.empty(len(t)) for _ in range(4)]
        self._check(self._L.hfmm_direct(self.h, len(t), _ip64(t), *(_dp(a) for a in out)))
        return tuple(out)

    def selftest_mfma_bf16(self):
        out = np.empty(256, dtype=np.float32)
        self._check(self._L.hfmm_selftest_mfma_bf16(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out.reshape(16, 16)


class FChebSymKernel:
    """Kernel parameters of FChebSymKernel<double, ..., FInterpMatrixKernelR, ORDER>."""

    def __init__(self, tree_height, box_width, box_center, order, epsilon=None, device=0,
                 precision=_abi.PRECISION_FP64, m2l_scheme=_abi.M2L_SYMMETRIC):
        self.tree_height = int(tree_height)
        self.box_width = float(box_width)
        self.box_center = tuple(float(c) for c in box_center)
        self.order = int(order)
        self.epsilon = float(epsilon) if epsilon else 0.0
        self.device = device
        self.precision = int(precision)
        self.engine = Engine(tree_height, box_width, box_center, order, self.epsilon, device, self.precision,
                             m2l_scheme)


class FChebKernel(FChebSymKernel):
    """Kernel parameters of FChebKernel<double, ..., FInterpMatrixKernelR, ORDER>
    (Src/Kernels/Chebyshev/FChebKernel.hpp:59-84): the same interpolation, the
    M2L through one global SVD of the 316 far-field operators (FP64)."""

    def __init__(self, tree_height, box_width, box_center, order, epsilon=None, device=0):
        super().__init__(tree_height, box_width, box_center, order, epsilon, device, _abi.PRECISION_FP64,
                         _abi.M2L_GLOBAL_SVD)


class FChebDenseKernel(FChebSymKernel):
    """Kernel parameters of FChebDenseKernel<double, ..., FInterpMatrixKernelR, ORDER>
    (Src/Kernels/Chebyshev/FChebDenseKernel.hpp): the 316 M2L operators
    applied uncompressed (FP64)."""

    def __init__(self, tree_height, box_width, box_center, order, device=0):
        super().__init__(tree_height, box_width, box_center, order, None, device, _abi.PRECISION_FP64,
                         _abi.M2L_DENSE)


class FUnifKernel(FChebSymKernel):
    """Kernel parameters of FUnifKernel<double, ..., FInterpMatrixKernelR, ORDER>
    (Src/Kernels/Uniform/FUnifKernel.hpp:56-69): Lagrange interpolation on
    equispaced roots, the M2L applying the kernel between the equispaced
    nodes of target and source cells (FP64)."""

    def __init__(self, tree_height, box_width, box_center, order, device=0):
        super().__init__(tree_height, box_width, box_center, order, None, device, _abi.PRECISION_FP64,
                         _abi.M2L_UNIFORM)


class FGroupTree:
    """Particles of one simulation box; the octree itself lives on the device."""

    def __init__(self, tree_height, box_width, box_center, x, y, z, q):
        self.tree_height = int(tree_height)
        self.box_width = float(box_width)
        self.box_center = tuple(float(c) for c in box_center)
        self.x, self.y, self.z, self.q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
        self.potentials = self.forces_x = self.forces_y = self.forces_z = None


class FGroupTaskAlgorithm:
    """execute(ops) builds the tree on the device and runs the selected passes."""

    def __init__(self, tree: FGroupTree, kernel: FChebSymKernel):
        if (tree.tree_height, tree.box_width, tree.box_center) != (kernel.tree_height, kernel.box_width,
                                                                   kernel.box_center):
            raise ValueError("tree and kernel disagree on height / box")
        self.tree = tree
        self.kernel = kernel

    def execute(self, operations=FFmmNearAndFarFields, upper_level=None, lower_level=None):
        """execute(ops) or execute(ops, upperWorkingLevel, lowerWorkingLevel)
        (Src/Core/FCoreCommon.hpp:65-94)."""
        e = self.kernel.engine
        t = self.tree
        e.set_particles(t.x, t.y, t.z, t.q)
        e.build_tree()
        levels = None
        if upper_level is not None or lower_level is not None:
            levels = (2 if upper_level is None else upper_level,
                      t.tree_height if lower_level is None else lower_level)
        e.evaluate(operations, levels)
        t.potentials, t.forces_x, t.forces_y, t.forces_z = e.results()
        return t
