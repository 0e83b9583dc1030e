"""ctypes declarations of the C-ABI in include/scalfmm_hip.h.

The product path has no CPU fallback: when the HIP library is missing,
`lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libscalfmm_hip.so")

OK, ERR_ARG, ERR_OUTSIDE_BOX, ERR_NUMERIC, ERR_NOMEM, ERR_DEVICE, ERR_STATE, ERR_IO = range(8)

# FFmmOperations (Src/Core/FCoreCommon.hpp:17-29)
P2P, P2M, M2M, M2L, L2L, L2P = 1, 2, 4, 8, 16, 32
NEAR_FIELD = P2P
FAR_FIELD = P2M | M2M | M2L | L2L | L2P
NEAR_AND_FAR = NEAR_FIELD | FAR_FIELD

M2L_SYMMETRIC = 0
M2L_GLOBAL_SVD = 1
M2L_DENSE = 2
M2L_UNIFORM = 3

CELLS_FRESH = 0
CELLS_KEPT = 1

PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_MIXED_HILO = 2

# every symbol the header declares
EXPORTS = (
    "hfmm_create", "hfmm_destroy", "hfmm_last_error", "hfmm_set_particles", "hfmm_set_particles_device",
    "hfmm_build_tree", "hfmm_evaluate", "hfmm_synchronize", "hfmm_get_results", "hfmm_get_results_device",
    "hfmm_get_tree_info", "hfmm_get_sorted_keys", "hfmm_get_cell_keys", "hfmm_get_leaf_ranges",
    "hfmm_get_p2p_list", "hfmm_get_m2l_list", "hfmm_get_operator", "hfmm_get_expansions",
    "hfmm_set_profiling", "hfmm_get_kernel_times", "hfmm_reset_kernel_times", "hfmm_selftest_mfma_f64",
    "hfmm_selftest_mfma_bf16", "hfmm_direct", "hfmm_direct_noalloc",
    "hfmm_get_field", "hfmm_get_field_device", "hfmm_set_charges", "hfmm_set_charges_device",
    "hfmm_version", "hfmm_set_particle_sets",
    "hfmm_batch_default_height", "hfmm_batch_create", "hfmm_batch_destroy", "hfmm_batch_last_error",
    "hfmm_batch_set_particles", "hfmm_batch_build", "hfmm_batch_evaluate", "hfmm_batch_synchronize",
    "hfmm_batch_get_results", "hfmm_batch_info", "hfmm_batch_tree", "hfmm_get_m2l_counts",
    "hfmm_set_m2l_variant", "hfmm_get_m2l_variant", "hfmm_evaluate_levels",
    "hfmm_fma_read", "hfmm_fma_write", "hfmm_save_operators", "hfmm_load_operators", "hfmm_operator_file_name",
    "hfmm_evaluate_particle_blocks", "hfmm_set_cell_mode", "hfmm_reset_cells", "hfmm_set_stream_overlap",
    "hfmm_set_m2l_scheme", "hfmm_get_global_operators", "hfmm_set_p2p_mutual", "hfmm_set_profiling_filter",
    "hfmm_get_cell_blocks", "hfmm_set_cell_blocks",
)
MAX_RHS = 64
DIRECT_NOALLOC_MAX_TARGETS = 65536  # HFMM_DIRECT_NOALLOC_MAX_TARGETS
_NON_STATUS = ("hfmm_last_error", "hfmm_version", "hfmm_batch_last_error", "hfmm_batch_default_height")


class Params(C.Structure):
    _fields_ = [("tree_height", C.c_int32), ("order", C.c_int32), ("box_width", C.c_double),
                ("box_center", C.c_double * 3), ("epsilon", C.c_double), ("precision", C.c_int32),
                ("device", C.c_int32), ("m2l_scheme", C.c_int32)]


class TreeInfo(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("height", C.c_int32), ("order", C.c_int32),
                ("n_cells", C.c_int64 * 21), ("p2p_entries", C.c_int64), ("m2l_entries", C.c_int64 * 21)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("total_ms", C.c_double), ("launches", C.c_int64)]


_lib = None


class HipLibraryMissing(RuntimeError):
    pass


def lib(path: str | None = None):
    """Loads the HIP engine (no fallback): `path`, else $SCALFMM_HIP_LIB, else
    the in-tree build."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("SCALFMM_HIP_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise HipLibraryMissing(f"HIP engine not built: {p} is missing (run __graft_entry__.build())")
    L = C.CDLL(p)
    vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    dp = C.POINTER(C.c_double)
    L.hfmm_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.hfmm_destroy.argtypes = [vp]
    L.hfmm_last_error.restype = C.c_char_p
    L.hfmm_last_error.argtypes = [vp]
    L.hfmm_set_particles.argtypes = [vp, i64, dp, dp, dp, dp]
    L.hfmm_set_particles_device.argtypes = [vp, i64, vp, vp, vp, vp]
    L.hfmm_build_tree.argtypes = [vp]
    L.hfmm_evaluate.argtypes = [vp, C.c_uint32]
    L.hfmm_synchronize.argtypes = [vp]
    L.hfmm_get_results.argtypes = [vp, dp, dp, dp, dp]
    L.hfmm_get_results_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    L.hfmm_get_tree_info.argtypes = [vp, C.POINTER(TreeInfo)]
    L.hfmm_get_sorted_keys.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.hfmm_get_cell_keys.argtypes = [vp, i32, C.POINTER(i64)]
    L.hfmm_get_leaf_ranges.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.hfmm_get_p2p_list.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]
    L.hfmm_get_m2l_list.argtypes = [vp, i32, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]
    L.hfmm_get_operator.argtypes = [vp, i32, C.POINTER(i32), dp]
    L.hfmm_get_expansions.argtypes = [vp, i32, dp, dp]
    L.hfmm_set_profiling.argtypes = [vp, i32]
    L.hfmm_get_kernel_times.argtypes = [vp, C.POINTER(KernelTime), i32, C.POINTER(i32)]
    L.hfmm_reset_kernel_times.argtypes = [vp]
    L.hfmm_selftest_mfma_f64.argtypes = [vp, dp]
    L.hfmm_selftest_mfma_bf16.argtypes = [vp, C.POINTER(C.c_float)]
    L.hfmm_direct.argtypes = [vp, C.c_int64, C.POINTER(C.c_int64), dp, dp, dp, dp]
    L.hfmm_direct_noalloc.argtypes = [vp, C.c_int64, C.POINTER(C.c_int64), dp, dp, dp, dp]
    L.hfmm_get_field.argtypes = [vp, dp, dp, dp]
    L.hfmm_set_charges.argtypes = [vp, dp]
    L.hfmm_set_charges_device.argtypes = [vp, vp]
    L.hfmm_get_field_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    L.hfmm_version.restype = C.c_char_p
    L.hfmm_set_particle_sets.argtypes = [vp, i32, C.POINTER(i64), i32, i64, dp, dp, dp, dp]
    L.hfmm_batch_default_height.restype = i32
    L.hfmm_batch_default_height.argtypes = [i64]
    L.hfmm_batch_create.argtypes = [C.POINTER(Params), i32, C.POINTER(i64), C.POINTER(i32), i32, C.POINTER(vp)]
    L.hfmm_batch_destroy.argtypes = [vp]
    L.hfmm_batch_last_error.restype = C.c_char_p
    L.hfmm_batch_last_error.argtypes = [vp]
    L.hfmm_batch_set_particles.argtypes = [vp, dp, dp, dp, dp]
    L.hfmm_batch_build.argtypes = [vp]
    L.hfmm_batch_evaluate.argtypes = [vp, C.c_uint32]
    L.hfmm_batch_synchronize.argtypes = [vp]
    L.hfmm_batch_get_results.argtypes = [vp, dp, dp, dp, dp]
    L.hfmm_batch_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    L.hfmm_batch_tree.argtypes = [vp, i32, C.POINTER(vp), C.POINTER(i32), C.POINTER(i32)]
    L.hfmm_get_m2l_counts.argtypes = [vp, i32, C.POINTER(i64)]
    L.hfmm_set_m2l_variant.argtypes = [vp, i32]
    L.hfmm_evaluate_levels.argtypes = [vp, C.c_uint32, i32, i32]
    L.hfmm_fma_read.argtypes = [C.c_char_p, C.POINTER(i64), C.POINTER(i32), dp, dp, dp]
    L.hfmm_fma_write.argtypes = [C.c_char_p, i64, i32, d, dp, dp]
    L.hfmm_save_operators.argtypes = [vp, C.c_char_p]
    L.hfmm_load_operators.argtypes = [vp, C.c_char_p]
    L.hfmm_operator_file_name.argtypes = [i32, d, C.c_char_p, i32]
    L.hfmm_get_m2l_variant.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_float)]
    L.hfmm_set_cell_mode.argtypes = [vp, i32]
    L.hfmm_set_stream_overlap.argtypes = [vp, i32]
    L.hfmm_set_p2p_mutual.argtypes = [vp, i32]
    L.hfmm_set_profiling_filter.argtypes = [vp, C.c_char_p]
    L.hfmm_set_m2l_scheme.argtypes = [vp, i32]
    L.hfmm_get_global_operators.argtypes = [vp, C.POINTER(i32), dp, dp, dp, dp, dp]
    L.hfmm_reset_cells.argtypes = [vp]
    L.hfmm_evaluate_particle_blocks.argtypes = [vp, i32, C.POINTER(vp), C.POINTER(dp), i32, C.c_uint32, i32, i32]
    L.hfmm_get_cell_blocks.argtypes = [vp, i32, i32, C.POINTER(vp), C.POINTER(dp), C.POINTER(dp), i32]
    L.hfmm_set_cell_blocks.argtypes = [vp, i32, i32, C.POINTER(vp), C.POINTER(dp), C.POINTER(dp), i32]
    for name in EXPORTS:
        if name not in _NON_STATUS:
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
