"""ctypes bindings of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(scalfmm_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

P2P, P2M, M2M, M2L, L2L, L2P = 1, 2, 4, 8, 16, 32
ALL = 63

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("t_tree", "t_p2m", "t_m2m", "t_m2l", "t_l2l", "t_l2p", "t_p2p", "t_total")] + \
               [("n_m2l", C.c_int64), ("n_p2p_pairs", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    L.orc_morton_encode.restype = C.c_int64
    L.orc_morton_encode.argtypes = [C.c_int, C.c_int, C.c_int]
    L.orc_morton_decode.argtypes = [C.c_int64, _i32p]
    L.orc_particle_keys.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int, vp, _i64p]
    L.orc_neighbor_keys.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i64p, _i32p]
    L.orc_interaction_keys.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i64p, _i32p]
    L.orc_cheb_roots.argtypes = [C.c_int, _f64p]
    L.orc_cheb_T.restype = C.c_double
    L.orc_cheb_T.argtypes = [C.c_uint, C.c_double]
    L.orc_cheb_create.argtypes = [C.c_int, C.c_double, C.POINTER(vp)]
    L.orc_cheb_destroy.argtypes = [vp]
    L.orc_cheb_rank.argtypes = [vp, C.c_int]
    L.orc_cheb_get_K.argtypes = [vp, C.c_int, _f64p]
    L.orc_cheb_get_symmetries.argtypes = [vp, _u32p, _u32p]
    L.orc_cheb_get_child_interp.argtypes = [vp, _f64p]
    L.orc_cheb_apply_m2l.argtypes = [vp, C.c_int, C.c_double, _f64p, _f64p]
    L.orc_cheb_dense_block.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f64p]
    L.orc_p2m.argtypes = [vp, _f64p, C.c_double, C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_l2p.argtypes = [vp, _f64p, C.c_double, _f64p, C.c_int64, _f64p, _f64p, _f64p, _f64p,
                          _f64p, _f64p, _f64p, _f64p]
    L.orc_m2m.argtypes = [vp, C.c_int, _f64p, _f64p]
    L.orc_l2l.argtypes = [vp, C.c_int, _f64p, _f64p]
    L.orc_tree_build.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int, C.POINTER(vp)]
    L.orc_tree_destroy.argtypes = [vp]
    L.orc_tree_sorted.argtypes = [vp, _i64p, _i64p]
    L.orc_tree_num_cells.restype = C.c_int64
    L.orc_tree_num_cells.argtypes = [vp, C.c_int]
    L.orc_tree_cell_keys.argtypes = [vp, C.c_int, _i64p]
    L.orc_tree_leaf_ranges.argtypes = [vp, _i64p, _i64p]
    L.orc_tree_p2p_list_size.restype = C.c_int64
    L.orc_tree_p2p_list_size.argtypes = [vp]
    L.orc_tree_p2p_list.argtypes = [vp, _i64p, _i64p, _i32p]
    L.orc_tree_m2l_list_size.restype = C.c_int64
    L.orc_tree_m2l_list_size.argtypes = [vp, C.c_int]
    L.orc_tree_m2l_list.argtypes = [vp, C.c_int, _i64p, _i64p, _i32p]
    L.orc_fmm.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int, C.c_int,
                          C.c_double, C.c_uint, C.c_int, _f64p, _f64p, _f64p, _f64p, C.POINTER(Stats)]
    L.orc_fmm_with.argtypes = [vp, C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                               C.c_uint, C.c_int, _f64p, _f64p, _f64p, _f64p, C.POINTER(Stats)]
    L.orc_fmm_levels.argtypes = [vp, C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                 C.c_uint, C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, C.POINTER(Stats)]
    L.orc_run_create.argtypes = [vp, C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                 C.c_uint, C.POINTER(vp)]
    L.orc_run_destroy.argtypes = [vp]
    L.orc_run_tree.restype = vp
    L.orc_run_tree.argtypes = [vp]
    L.orc_run_expansions.argtypes = [vp, C.c_int, _f64p, _f64p]
    L.orc_chebg_create.argtypes = [C.c_int, C.c_double, C.POINTER(vp)]
    L.orc_chebg_destroy.argtypes = [vp]
    L.orc_chebd_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.orc_unif_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.orc_chebg_rank.argtypes = [vp]
    L.orc_chebg_get.argtypes = [vp, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_chebg_set.argtypes = [vp, C.c_int, _f64p, _f64p, _f64p]
    L.orc_fmm_global.argtypes = [vp, C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                 C.c_uint, C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, C.POINTER(Stats)]
    L.orc_run_execute.argtypes = [vp, C.c_uint, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p]
    L.orc_run_reset_cells.argtypes = [vp]
    L.orc_direct.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p, C.c_int64, vp, C.c_int,
                             _f64p, _f64p, _f64p, _f64p]
    L.orc_qr.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p]
    L.orc_svd.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
    L.orc_svd_rank.argtypes = [_f64p, C.c_int, C.c_double]
    _lib = L
    return L


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _center(center):
    return _f(np.asarray(center, dtype=np.float64).reshape(3))


# ---- indexing ---------------------------------------------------------------
def morton(x, y, z):
    return int(lib().orc_morton_encode(int(x), int(y), int(z)))


def demorton(key):
    out = np.zeros(3, np.int32)
    lib().orc_morton_decode(int(key), out)
    return tuple(int(v) for v in out)


def particle_keys(x, y, z, center, width, height):
    n = len(x)
    keys = np.zeros(max(n, 1), np.int64)
    st = lib().orc_particle_keys(n, _f(x), _f(y), _f(z), _center(center), float(width), int(height), None, keys)
    return st, keys[:n]


def neighbor_keys(x, y, z, height):
    k = np.zeros(26, np.int64); p = np.zeros(26, np.int32)
    m = lib().orc_neighbor_keys(x, y, z, height, k, p)
    return k[:m], p[:m]


def interaction_keys(x, y, z, level):
    k = np.zeros(216, np.int64); p = np.zeros(216, np.int32)
    m = lib().orc_interaction_keys(x, y, z, level, k, p)
    return k[:m], p[:m]


def roots(order):
    r = np.zeros(order)
    if lib().orc_cheb_roots(order, r):
        raise ValueError(order)
    return r


# ---- Chebyshev operators ------------------------------------------------------
class Cheb:
    def __init__(self, order, epsilon=0.0):
        self.order = order
        self.nnodes = order ** 3
        h = C.c_void_p()
        st = lib().orc_cheb_create(order, float(epsilon), C.byref(h))
        if st:
            raise RuntimeError(f"orc_cheb_create failed: {st}")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_cheb_destroy(self.h)
            self.h = None

    def rank(self, idx):
        return int(lib().orc_cheb_rank(self.h, idx))

    def canonical_ranks(self):
        out = {}
        for i in range(2, 4):
            for j in range(0, i + 1):
                for k in range(0, j + 1):
                    idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3)
                    out[(i, j, k)] = self.rank(idx)
        return out

    def K(self, idx):
        r = self.rank(idx)
        out = np.zeros(2 * r * self.nnodes)
        lib().orc_cheb_get_K(self.h, idx, out)
        U = out[: r * self.nnodes].reshape(r, self.nnodes).T
        V = out[r * self.nnodes:].reshape(r, self.nnodes).T
        return U, V

    def symmetries(self):
        pi = np.zeros(343, np.uint32)
        pv = np.zeros(343 * self.nnodes, np.uint32)
        lib().orc_cheb_get_symmetries(self.h, pi, pv)
        return pi, pv.reshape(343, self.nnodes)

    def child_interp(self):
        out = np.zeros(8 * 3 * self.order ** 2)
        lib().orc_cheb_get_child_interp(self.h, out)
        return out.reshape(8, 3, self.order, self.order)

    def apply_m2l(self, idx, scale, M, L=None):
        L = np.zeros(self.nnodes) if L is None else _f(L).copy()
        st = lib().orc_cheb_apply_m2l(self.h, idx, float(scale), _f(M), L)
        if st:
            raise ValueError(idx)
        return L

    def p2m(self, center, width, x, y, z, q):
        M = np.zeros(self.nnodes)
        lib().orc_p2m(self.h, _center(center), float(width), len(x), _f(x), _f(y), _f(z), _f(q), M)
        return M

    def l2p(self, center, width, L, x, y, z, q):
        n = len(x)
        pot, fx, fy, fz = (np.zeros(n) for _ in range(4))
        lib().orc_l2p(self.h, _center(center), float(width), _f(L), n, _f(x), _f(y), _f(z), _f(q), pot, fx, fy, fz)
        return pot, fx, fy, fz

    def m2m(self, child, Mc, Mp=None):
        Mp = np.zeros(self.nnodes) if Mp is None else _f(Mp).copy()
        lib().orc_m2m(self.h, child, _f(Mc), Mp)
        return Mp

    def l2l(self, child, Lp, Lc=None):
        Lc = np.zeros(self.nnodes) if Lc is None else _f(Lc).copy()
        lib().orc_l2l(self.h, child, _f(Lp), Lc)
        return Lc


def dense_block(order, i, j, k):
    nn = order ** 3
    out = np.zeros(nn * nn)
    lib().orc_cheb_dense_block(order, i, j, k, out)
    return out.reshape(nn, nn)  # [m (source node), n (target node)]


# ---- tree ----------------------------------------------------------------------
class Tree:
    def __init__(self, x, y, z, center, width, height, handle=None, owner=True):
        self.height = height
        self.n = len(x)
        if handle is None:
            h = C.c_void_p()
            st = lib().orc_tree_build(self.n, _f(x), _f(y), _f(z), _center(center), float(width), height,
                                      C.byref(h))
            if st:
                raise RuntimeError(f"orc_tree_build failed: {st}")
            handle = h
        self.h = handle
        self.owner = owner

    def __del__(self):
        if getattr(self, "owner", False) and getattr(self, "h", None) and _lib is not None:
            _lib.orc_tree_destroy(self.h)
            self.h = None

    def sorted(self):
        k = np.zeros(max(self.n, 1), np.int64); p = np.zeros(max(self.n, 1), np.int64)
        lib().orc_tree_sorted(self.h, k, p)
        return k[: self.n], p[: self.n]

    def num_cells(self, level):
        return int(lib().orc_tree_num_cells(self.h, level))

    def cell_keys(self, level):
        out = np.zeros(max(self.num_cells(level), 1), np.int64)
        lib().orc_tree_cell_keys(self.h, level, out)
        return out[: self.num_cells(level)]

    def leaf_ranges(self):
        nl = self.num_cells(self.height - 1)
        s = np.zeros(max(nl, 1), np.int64); c = np.zeros(max(nl, 1), np.int64)
        lib().orc_tree_leaf_ranges(self.h, s, c)
        return s[:nl], c[:nl]

    def p2p_list(self):
        nl = self.num_cells(self.height - 1)
        m = int(lib().orc_tree_p2p_list_size(self.h))
        off = np.zeros(nl + 1, np.int64); nb = np.zeros(max(m, 1), np.int64); ps = np.zeros(max(m, 1), np.int32)
        lib().orc_tree_p2p_list(self.h, off, nb, ps)
        return off, nb[:m], ps[:m]

    def m2l_list(self, level):
        nc = self.num_cells(level)
        m = int(lib().orc_tree_m2l_list_size(self.h, level))
        off = np.zeros(nc + 1, np.int64); sr = np.zeros(max(m, 1), np.int64); ps = np.zeros(max(m, 1), np.int32)
        lib().orc_tree_m2l_list(self.h, level, off, sr, ps)
        return off, sr[:m], ps[:m]


# ---- evaluation ------------------------------------------------------------------
def fmm(x, y, z, q, center, width, height, order, epsilon=0.0, ops=ALL, nthreads=0, cheb=None, levels=None):
    """FMM over the tree of `height`; `levels` = (upper, lower) restricts the
    working levels as FAbstractAlgorithm::execute(ops, upper, lower) does."""
    n = len(x)
    pot, fx, fy, fz = (np.zeros(max(n, 1)) for _ in range(4))
    st = Stats()
    if levels is not None:
        own = cheb is None
        cheb = cheb or Cheb(order, epsilon)
        rc = lib().orc_fmm_levels(cheb.h, n, _f(x), _f(y), _f(z), _f(q), _center(center), float(width), height,
                                  ops, int(levels[0]), int(levels[1]), nthreads, pot, fx, fy, fz, C.byref(st))
        if own:
            del cheb
    elif cheb is not None:
        rc = lib().orc_fmm_with(cheb.h, n, _f(x), _f(y), _f(z), _f(q), _center(center), float(width), height,
                                ops, nthreads, pot, fx, fy, fz, C.byref(st))
    else:
        rc = lib().orc_fmm(n, _f(x), _f(y), _f(z), _f(q), _center(center), float(width), height, order,
                           float(epsilon), ops, nthreads, pot, fx, fy, fz, C.byref(st))
    if rc:
        raise RuntimeError(f"orc_fmm failed: {rc}")
    return pot[:n], fx[:n], fy[:n], fz[:n], st


class Run:
    """Level-by-level run kept alive for inspecting expansions."""

    def __init__(self, cheb, x, y, z, q, center, width, height, ops=ALL):
        h = C.c_void_p()
        rc = lib().orc_run_create(cheb.h, len(x), _f(x), _f(y), _f(z), _f(q), _center(center), float(width),
                                  height, ops, C.byref(h))
        if rc:
            raise RuntimeError(f"orc_run_create failed: {rc}")
        self.h = h
        self.cheb = cheb
        self.n = len(x)
        self.height = int(height)
        self.tree = Tree(x, y, z, center, width, height, handle=lib().orc_run_tree(h), owner=False)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_run_destroy(self.h)
            self.h = None

    def execute(self, ops=ALL, upper=2, lower=None):
        """Another execute on this run's tree: passes add into the cells and
        particle results of earlier ones; returns the accumulated results."""
        n = self.n
        out = [np.zeros(max(n, 1)) for _ in range(4)]
        lower = self.height if lower is None else lower
        rc = lib().orc_run_execute(self.h, ops, int(upper), int(lower), *out)
        if rc:
            raise RuntimeError(f"orc_run_execute failed: {rc}")
        return tuple(a[:n] for a in out)

    def reset_cells(self):
        lib().orc_run_reset_cells(self.h)

    def expansions(self, level):
        nc = self.tree.num_cells(level)
        M = np.zeros(max(nc, 1) * self.cheb.nnodes); L = np.zeros(max(nc, 1) * self.cheb.nnodes)
        lib().orc_run_expansions(self.h, level, M, L)
        return M[: nc * self.cheb.nnodes].reshape(nc, -1), L[: nc * self.cheb.nnodes].reshape(nc, -1)


class ChebGlobal:
    """FChebKernel's operators: one global SVD of the 316 far-field operators."""

    def __init__(self, order, epsilon=0.0, dense=False, uniform=False):
        self.order = order
        self.nnodes = order ** 3
        h = C.c_void_p()
        if uniform:  # FUnifKernel: Lagrange on equispaced roots, the kernel applied exactly
            st = lib().orc_unif_create(order, C.byref(h))
        elif dense:  # FChebDenseKernel: the operators uncompressed, identity bases
            st = lib().orc_chebd_create(order, C.byref(h))
        else:
            st = lib().orc_chebg_create(order, float(epsilon), C.byref(h))
        if st:
            raise RuntimeError(f"orc_chebg_create failed: {st}")
        self.h = h
        self.rank = int(lib().orc_chebg_rank(h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_chebg_destroy(self.h)
            self.h = None

    def set_operators(self, U, B, Cm):
        """Replaces U, B (nnodes x k) and C (343 x k x k, C[pos][i, j])."""
        k = U.shape[1]
        Uf = np.ascontiguousarray(U.T).reshape(-1)
        Bf = np.ascontiguousarray(B.T).reshape(-1)
        Cf = np.ascontiguousarray(np.asarray(Cm).transpose(0, 2, 1)).reshape(-1)
        st = lib().orc_chebg_set(self.h, int(k), Uf, Bf, Cf)
        if st:
            raise RuntimeError(f"orc_chebg_set failed: {st}")
        self.rank = int(k)

    def operators(self):
        """U, B (nnodes x k), C (343 x k x k with C[pos][i, j]), S1, S2."""
        nn, k = self.nnodes, self.rank
        U, B = np.zeros(nn * k), np.zeros(nn * k)
        Cm = np.zeros(343 * k * k)
        S1, S2 = np.zeros(nn), np.zeros(nn)
        lib().orc_chebg_get(self.h, U, B, Cm, S1, S2)
        return (U.reshape(k, nn).T, B.reshape(k, nn).T, Cm.reshape(343, k, k).transpose(0, 2, 1), S1, S2)


def fmm_global(g, x, y, z, q, center, width, height, ops=ALL, levels=None, nthreads=0):
    """The FMM with FChebKernel's compressed M2L over operators `g` (ChebGlobal)."""
    n = len(x)
    pot, fx, fy, fz = (np.zeros(max(n, 1)) for _ in range(4))
    st = Stats()
    up, lo = levels if levels is not None else (2, height)
    rc = lib().orc_fmm_global(g.h, n, _f(x), _f(y), _f(z), _f(q), _center(center), float(width), height, ops,
                              int(up), int(lo), nthreads, pot, fx, fy, fz, C.byref(st))
    if rc:
        raise RuntimeError(f"orc_fmm_global failed: {rc}")
    return pot[:n], fx[:n], fy[:n], fz[:n], st


def direct(x, y, z, q, targets=None, nthreads=0):
    n = len(x)
    if targets is None:
        m = n
        tp = None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        m = len(targets)
        tp = targets.ctypes.data_as(C.c_void_p)
    pot, fx, fy, fz = (np.zeros(max(m, 1)) for _ in range(4))
    lib().orc_direct(n, _f(x), _f(y), _f(z), _f(q), m, tp, nthreads, pot, fx, fy, fz)
    return pot[:m], fx[:m], fy[:m], fz[:m]


def qr(A):
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    Q = np.zeros(m * n); R = np.zeros(n * n)
    rc = lib().orc_qr(m, n, A.ravel(order="F").copy(), Q, R)
    if rc:
        raise RuntimeError(rc)
    return Q.reshape(n, m).T, R.reshape(n, n).T


def svd(A):
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    U = np.zeros(n * n); S = np.zeros(n); VT = np.zeros(n * n)
    rc = lib().orc_svd(n, A.ravel(order="F").copy(), U, S, VT)
    if rc:
        raise RuntimeError(rc)
    return U.reshape(n, n).T, S, VT.reshape(n, n).T


def svd_rank(S, eps):
    return int(lib().orc_svd_rank(_f(S), len(S), float(eps)))


# ---- error metrics: FMath::FAccurater (Utils/FMath.hpp:519-601) -------------------------
def accuracy(good, bad):
    good = np.asarray(good, np.float64); bad = np.asarray(bad, np.float64)
    d = bad - good
    l2diff = float(np.dot(d, d)); l2dot = float(np.dot(good, good))
    return {"rel_l2": (l2diff / l2dot) ** 0.5 if l2dot > 0 else float("nan"),
            "rms": (l2diff / len(good)) ** 0.5 if len(good) else float("nan"),
            "rel_inf": float(np.max(np.abs(d)) / np.max(np.abs(good))) if len(good) else float("nan")}


# ---- fixtures: FFmaGenericLoader binary format (Files/FFmaGenericLoader.hpp:480-505) ------
def read_bfma(path):
    with open(path, "rb") as f:
        b = f.read()
    sz, nv = np.frombuffer(b[:8], dtype=np.uint32)
    n = int(np.frombuffer(b[8:16], dtype=np.int64)[0])
    assert sz == 8
    hdr = np.frombuffer(b[16:48], dtype=np.float64)
    data = np.frombuffer(b[48:48 + 8 * n * int(nv)], dtype=np.float64).reshape(n, int(nv))
    return {"half_width": float(hdr[0]), "width": 2.0 * float(hdr[0]), "center": hdr[1:4].copy(),
            "data": data.copy()}
