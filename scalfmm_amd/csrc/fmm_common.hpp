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
};

// Leaf centre (getLeafCellCenter, FAbstractChebKernel.hpp:53-58) and the
// affine map bounds a = c - w/2, b = c + w/2 (FInterpMapping.hpp:30-37).
__device__ __forceinline__ void leaf_bounds(const Box& bx, uint64_t key, double a[3], double b[3]) {
    const int co[3] = {static_cast<int>(compact3(key >> 2)), static_cast<int>(compact3(key >> 1)),
                       static_cast<int>(compact3(key))};
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

// The box of the context with its leaf width (level height-1).
Box make_box(const Context* c);

struct M2LLevels {
    int nlev;
    int tile_begin[kMaxLevels + 1];  // prefix of (upper bounds of) tiles per level
    int64_t cell_off[kMaxLevels];    // first cell of the level in the cell arrays
    int pcount_off[kMaxLevels];      // parity counts of the level
    double scale[kMaxLevels];
};

// Tile bookkeeping of the all-level M2L launch: levels 2..h-1, each a segment
// of 8 x ceil(ncell / tile_targets) workgroups (parity = workgroup % 8).
M2LLevels make_levels(const Context* c, int tile_targets, int& tiles);

}  // namespace hfmm
