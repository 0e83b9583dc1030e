// fmm_common.hpp -- device helpers shared by the FP64 and mixed-precision
// pass kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hfmm_internal.hpp"

namespace hfmm {

constexpr int kM2LWaves = 4;  // waves per M2L workgroup

__device__ __forceinline__ uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return static_cast<uint32_t>(x);
}

struct Box {
    double corner[3];
    double leafw;
    uint32_t cmask;  // leaves per axis of one set, minus 1: masks the set bits off a leaf coordinate
};

// Leaf centre (getLeafCellCenter, FAbstractChebKernel.hpp:53-58) and the
// affine map bounds a = c - w/2, b = c + w/2 (FInterpMapping.hpp:30-37), in
// the box of the leaf's own set.
__device__ __forceinline__ void leaf_bounds(const Box& bx, uint64_t key, double a[3], double b[3]) {
    const int co[3] = {static_cast<int>(compact3(key >> 2) & bx.cmask), static_cast<int>(compact3(key >> 1) & bx.cmask),
                       static_cast<int>(compact3(key) & bx.cmask)};
    for (int d = 0; d < 3; ++d) {
        const double c = bx.corner[d] + (static_cast<double>(co[d]) + .5) * bx.leafw;
        a[d] = c - bx.leafw / 2.;
        b[d] = c + bx.leafw / 2.;
    }
}

// Workgroups of the persistent big-leaf P2P kernels (one wave each): the
// item count is known only on the device, the capacity bounds it.
inline unsigned p2p_tiled_grid(const Context* c) {
    const int64_t g = c->p2p_items_cap < 4096 ? c->p2p_items_cap : 4096;
    return static_cast<unsigned>(g > 0 ? g : 1);
}

// Raises a kernel's dynamic LDS limit to `bytes` when a launch needs more than
// the default 64 KiB and more than set before on this device: the call is
// made once per kernel and device, by an eager run, never while a stream is
// being captured into a graph.
inline int raise_lds_limit(Context* c, const void* kernel, size_t bytes, int* set_per_device) {
    if (bytes <= 64 * 1024) return HFMM_OK;
    if (c->device < 0 || c->device >= 64) return fail(c, HFMM_ERR_ARG, "device ordinal above 63");
    if (static_cast<int>(bytes) <= set_per_device[c->device]) return HFMM_OK;
    HFMM_HIP(c, hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    set_per_device[c->device] = static_cast<int>(bytes);
    return HFMM_OK;
}

// The box of the context (of each set) with its leaf width.
Box make_box(const Context* c);

struct M2LLevels {
    int nlev;
    int tile_begin[kMaxLevels + 1];  // prefix of (upper bounds of) tiles per level
    int64_t cell_off[kMaxLevels];    // first cell of the level in the cell arrays
    int pcount_off[kMaxLevels];      // parity counts of the level
    double scale[kMaxLevels];
};

// Tile bookkeeping of the all-level M2L launch: levels top..h-1, each a segment
// of 8 x ceil(ncell / tile_targets) workgroups (parity = workgroup % 8).
M2LLevels make_levels(const Context* c, int tile_targets, int& tiles);

}  // namespace hfmm
