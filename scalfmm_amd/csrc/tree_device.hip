// tree_device.hip -- octree construction on the device (FP64 keys, int32
// indices): leaf coordinates and Morton keys, a stable LSD radix sort,
// run-length encoded leaves, parent levels, P2P neighbour slots and M2L
// interaction slots in the reference's enumeration order.
//
// Follows (reference paths relative to its source tree):
//   coordinates   Src/Containers/FCoordinateComputer.hpp:15-24, 42-54
//   Morton keys   Src/Containers/FTreeCoordinate.hpp:91-142
//   sort/leaves   Src/GroupTree/Core/FGroupTree.hpp:192-360
//   P2P list      Src/Containers/FTreeCoordinate.hpp:189-214
//   M2L list      Src/Containers/FTreeCoordinate.hpp:219-264
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // elements per scan / sort tile

// ---------------------------------------------------------------------------
// Morton helpers: bit 3k <- z_k, 3k+1 <- y_k, 3k+2 <- x_k.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
__device__ __forceinline__ uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return static_cast<uint32_t>(x);
}
__device__ __forceinline__ uint64_t morton3(int x, int y, int z) {
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}
__device__ __forceinline__ void demorton3(uint64_t k, int& x, int& y, int& z) {
    z = compact3(k);
    y = compact3(k >> 1);
    x = compact3(k >> 2);
}

// Index of a key among the sorted keys of one level (-1 when absent): a
// dense Morton -> index table when the level is small enough, otherwise a
// lower-bound search.
struct KeyIndex {
    const uint64_t* keys;
    int n;
    const int32_t* dense;  // 8^level entries or null
    __device__ __forceinline__ int find(uint64_t key) const {
        if (dense) return dense[key];
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (keys[mid] < key) lo = mid + 1; else hi = mid;
        }
        return (lo < n && keys[lo] == key) ? lo : -1;
    }
};

__global__ __launch_bounds__(256) void k_dense_index(int n, const uint64_t* __restrict__ keys,
                                                    int32_t* __restrict__ dense) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dense[keys[i]] = i;
}

// ---------------------------------------------------------------------------
// Exclusive scan (three passes: tile sums, scan of tile sums, tile scans).
template <typename T>
__device__ T block_exclusive_scan(T v, T* sh, T& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
        const T add = t >= d ? sh[t - d] : T(0);
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[kBlock - 1];
    const T incl = sh[t];
    __syncthreads();
    return incl - v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T* __restrict__ in, int64_t n, T* __restrict__ part) {
    __shared__ T sh[kBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    T s = 0;
    for (int i = 0; i < kItems; ++i)
        if (base + i < n) s += in[base + i];
    T total;
    block_exclusive_scan<T>(s, sh, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_top(T* __restrict__ part, int64_t nparts, T* __restrict__ total_out) {
    __shared__ T sh[kBlock];
    T carry = 0;
    for (int64_t c0 = 0; c0 < nparts; c0 += kTile) {
        const int64_t base = c0 + threadIdx.x * kItems;
        T v[kItems];
        T s = 0;
        for (int i = 0; i < kItems; ++i) {
            v[i] = base + i < nparts ? part[base + i] : T(0);
            s += v[i];
        }
        T tot;
        T ex = block_exclusive_scan<T>(s, sh, tot) + carry;
        for (int i = 0; i < kItems; ++i) {
            if (base + i < nparts) part[base + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_apply(const T* in, T* out, int64_t n, const T* __restrict__ part) {
    __shared__ T sh[kBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    T v[kItems];
    T s = 0;
    for (int i = 0; i < kItems; ++i) {
        v[i] = base + i < n ? in[base + i] : T(0);
        s += v[i];
    }
    T tot;
    T ex = block_exclusive_scan<T>(s, sh, tot) + part[blockIdx.x];
    for (int i = 0; i < kItems; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

// out = exclusive scan of in (may alias); *total_dev = sum.
template <typename T>
int exclusive_scan(Context* c, const T* in, T* out, int64_t n, T* total_dev, const char* tag) {
    if (n <= 0) {
        HFMM_HIP(c, hipMemsetAsync(total_dev, 0, sizeof(T), c->stream));
        return HFMM_OK;
    }
    const int64_t tiles = (n + kTile - 1) / kTile;
    HFMM_TRY(ensure(c, c->d_scan_part, sizeof(T) * static_cast<size_t>(tiles + 1)));
    T* part = c->d_scan_part.as<T>();
    prof_begin(c, tag);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, c->stream, in, n, part);
    hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(kBlock), 0, c->stream, part, tiles, total_dev);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, c->stream, in, out, n,
                       part);
    prof_end(c);
    return check_launch(c, tag);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (uint64 key, int32 value), 8-bit digits.
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                      int32_t* __restrict__ hist, int64_t ntiles) {
    __shared__ int32_t s[256];
    s[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
    for (int r = 0; r < kItems; ++r) {
        const int64_t e = base + r * kBlock + threadIdx.x;
        if (e < n) atomicAdd(&s[(keys[e] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * ntiles + blockIdx.x] = s[threadIdx.x];
}

// Scatter one tile: elements are visited in rounds of 256 consecutive
// elements; within a wave the rank among equal digits comes from ballots,
// across the four waves from per-digit counters, so the order of equal keys
// is preserved.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                         const int32_t* __restrict__ vin, int64_t n, int shift,
                                                         const int32_t* __restrict__ offs, int64_t ntiles,
                                                         uint64_t* __restrict__ kout, int32_t* __restrict__ vout) {
    __shared__ int32_t s_base[256];
    __shared__ int32_t s_wave[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    s_base[tid] = offs[static_cast<int64_t>(tid) * ntiles + blockIdx.x];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kItems; ++r) {
        const int64_t e = base + r * kBlock + tid;
        const bool valid = e < n;
        const uint64_t key = valid ? kin[e] : 0ull;
        const int digit = valid ? static_cast<int>((key >> shift) & 255u) : 0;
        for (int w = 0; w < 4; ++w) s_wave[w][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) s_wave[wave][digit] = __popcll(peers);
        __syncthreads();
        {
            int run = s_base[tid];
            for (int w = 0; w < 4; ++w) {
                const int cnt = s_wave[w][tid];
                s_wave[w][tid] = run;
                run += cnt;
            }
            s_base[tid] = run;
        }
        __syncthreads();
        if (valid) {
            const int dst = s_wave[wave][digit] + rank;
            kout[dst] = key;
            vout[dst] = vin[e];
        }
        __syncthreads();
    }
}

// Sorts (keys, vals) in place by the low `nbits` bits of the keys, using the
// context's second key/value buffers; the result ends in keys/vals.
int radix_sort(Context* c, uint64_t* keys, int32_t* vals, uint64_t* keys2, int32_t* vals2, int64_t n, int nbits) {
    if (n <= 1) return HFMM_OK;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    HFMM_TRY(ensure(c, c->d_hist, sizeof(int32_t) * static_cast<size_t>(256 * ntiles + 1)));
    HFMM_TRY(ensure(c, c->d_count, sizeof(int64_t) * 24));
    int32_t* hist = c->d_hist.as<int32_t>();
    const int passes = (nbits + 7) / 8;
    uint64_t *ka = keys, *kb = keys2;
    int32_t *va = vals, *vb = vals2;
    for (int p = 0; p < passes; ++p) {
        prof_begin(c, "radix_hist");
        hipLaunchKernelGGL(k_radix_hist, dim3(static_cast<unsigned>(ntiles)), dim3(kBlock), 0, c->stream, ka, n, 8 * p,
                           hist, ntiles);
        prof_end(c);
        HFMM_TRY(check_launch(c, "radix_hist"));
        HFMM_TRY(exclusive_scan<int32_t>(c, hist, hist, 256 * ntiles,
                                         reinterpret_cast<int32_t*>(c->d_count.as<int64_t>() + 15), "radix_scan"));
        prof_begin(c, "radix_scatter");
        hipLaunchKernelGGL(k_radix_scatter, dim3(static_cast<unsigned>(ntiles)), dim3(kBlock), 0, c->stream, ka, va, n,
                           8 * p, hist, ntiles, kb, vb);
        prof_end(c);
        HFMM_TRY(check_launch(c, "radix_scatter"));
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        HFMM_HIP(c, hipMemcpyAsync(keys, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, c->stream));
        HFMM_HIP(c, hipMemcpyAsync(vals, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    return HFMM_OK;
}

// ---------------------------------------------------------------------------
// Leaf coordinates and keys (FCoordinateComputer::GetTreeCoordinate):
// c = rel == width ? 2^(h-1)-1 : int(rel / leafwidth), rel in [0, width].
// With particle sets the key is (set << 3(h-1)) | set-local key; the set of a
// particle is found by bisection of the set offsets.
__global__ __launch_bounds__(kBlock) void k_morton(int64_t n, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ z,
                                                  double cx, double cy, double cz, double width, double leafw,
                                                  int lbits, const int64_t* __restrict__ setoff, int nsets,
                                                  uint64_t* __restrict__ keys, int32_t* __restrict__ vals,
                                                  int32_t* __restrict__ err) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int lim = 1 << lbits;
    const double rel[3] = {x[i] - cx, y[i] - cy, z[i] - cz};
    int co[3];
    bool bad = false;
    for (int d = 0; d < 3; ++d) {
        if (!(rel[d] >= 0 && rel[d] <= width)) { bad = true; co[d] = 0; continue; }
        co[d] = rel[d] == width ? lim - 1 : static_cast<int>(rel[d] / leafw);
        if (co[d] >= lim) { bad = true; co[d] = lim - 1; }
    }
    if (bad) atomicOr(err, 1);
    uint64_t set = 0;
    if (setoff) {  // last set whose first particle is <= i
        int lo = 0, hi = nsets - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (setoff[mid] <= i) lo = mid; else hi = mid - 1;
        }
        set = static_cast<uint64_t>(lo);
    }
    keys[i] = (set << (3 * lbits)) | morton3(co[0], co[1], co[2]);
    vals[i] = static_cast<int32_t>(i);
}

// The inverse of the sort permutation: input index -> sorted index.
__global__ __launch_bounds__(kBlock) void k_invert_perm(int64_t n, const int32_t* __restrict__ perm,
                                                       int32_t* __restrict__ iperm) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    iperm[perm[i]] = static_cast<int32_t>(i);
}

__global__ __launch_bounds__(kBlock) void k_gather_particles(int64_t n, const int32_t* __restrict__ idx,
                                                            const double* __restrict__ x, const double* __restrict__ y,
                                                            const double* __restrict__ z, const double* __restrict__ q,
                                                            double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t o = idx[i];
    out[i] = x[o];
    out[n + i] = y[o];
    out[2 * n + i] = z[o];
    out[3 * n + i] = q[o];
}

// New charges on an existing tree: sorted q[i] = q_in[perm[i]].
__global__ __launch_bounds__(kBlock) void k_gather_charges(int64_t n, const int32_t* __restrict__ perm,
                                                          const double* __restrict__ q, double* __restrict__ sq) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sq[i] = q[perm[i]];
}

// The sorted charges of vectors r0 .. r0 + R - 1 as [n][R] (T = float: the
// mixed forms' P2P charges, rounded as k_local_coords rounds them).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_charge_chunk(int64_t n, const int32_t* __restrict__ perm,
                                                               const double* __restrict__ q, int r0, int R,
                                                               T* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = perm[i];
    for (int r = 0; r < R; ++r) out[i * R + r] = static_cast<T>(q[static_cast<int64_t>(r0 + r) * n + o]);
}

// flags[i] = 1 where (keys[i] >> shift) starts a new group.
__global__ __launch_bounds__(kBlock) void k_head_flags(int64_t n, const uint64_t* __restrict__ keys, int shift,
                                                      int32_t* __restrict__ flags) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1 : 0;
}

// Writes one entry per group: key >> shift and the index of its first element;
// group_of[i] (optional) = group of element i.
__global__ __launch_bounds__(kBlock) void k_write_groups(int64_t n, const uint64_t* __restrict__ keys, int shift,
                                                        const int32_t* __restrict__ flags,
                                                        const int32_t* __restrict__ pos, uint64_t* __restrict__ gkey,
                                                        int32_t* __restrict__ gstart, int32_t* __restrict__ group_of) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t g = pos[i] + flags[i] - 1;
    if (group_of) group_of[i] = g;
    if (flags[i]) {
        gkey[g] = keys[i] >> shift;
        gstart[g] = static_cast<int32_t>(i);
    }
}

__global__ void k_set_int(int32_t* p, int32_t v) { *p = v; }

// P2P neighbour slots of every leaf: 26 slots in ascending position order,
// slot -> leaf index or -1 (outside the box or empty).  lim = leaves per axis
// of one set: the range test is on the set-local coordinates (the low bits),
// so no list leaves the set of its cell.
__global__ __launch_bounds__(kBlock) void k_p2p_slots(KeyIndex ix, int lim, int32_t* __restrict__ slots) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nleaf = ix.n;
    if (i >= nleaf) return;
    int x, y, z;
    demorton3(ix.keys[i], x, y, z);
    const int lx = x & (lim - 1), ly = y & (lim - 1), lz = z & (lim - 1);
    int s = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (!dx && !dy && !dz) continue;
                const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
                int f = -1;
                if (nx >= 0 && nx < lim && ny >= 0 && ny < lim && nz >= 0 && nz < lim)
                    f = ix.find(morton3(x + dx, y + dy, z + dz));
                slots[static_cast<int64_t>(i) * kNbrSlots + s] = f;
                ++s;
            }
}

// M2L slots of every cell of a level: 189 slots in the reference enumeration
// order (parent neighbour x,y,z from -1 to 1, then cousin 0..7, keeping the
// non-adjacent ones); slot -> cell index in the level or -1.  plim = parents
// per axis of one set (the range test is set-local, as for the P2P slots).
__global__ __launch_bounds__(kBlock) void k_m2l_slots(KeyIndex ix, int plim, int32_t* __restrict__ slots) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ix.n) return;
    int x, y, z;
    demorton3(ix.keys[i], x, y, z);
    const int px = x >> 1, py = y >> 1, pz = z >> 1;
    const int lx = px & (plim - 1), ly = py & (plim - 1), lz = pz & (plim - 1);
    const int bx = x & 1, by = y & 1, bz = z & 1;
    int s = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (!dx && !dy && !dz) continue;
                const int ox = lx + dx, oy = ly + dy, oz = lz + dz;
                const bool inside = ox >= 0 && ox < plim && oy >= 0 && oy < plim && oz >= 0 && oz < plim;
                const uint64_t pk = inside ? morton3(px + dx, py + dy, pz + dz) : 0ull;
                for (int cz = 0; cz < 8; ++cz) {
                    const int xd = 2 * dx + ((cz >> 2) & 1) - bx;
                    const int yd = 2 * dy + ((cz >> 1) & 1) - by;
                    const int zd = 2 * dz + (cz & 1) - bz;
                    if (abs(xd) <= 1 && abs(yd) <= 1 && abs(zd) <= 1) continue;
                    slots[static_cast<int64_t>(i) * kFarSlots + s] =
                        inside ? ix.find((pk << 3) | static_cast<uint64_t>(cz)) : -1;
                    ++s;
                }
            }
}

__global__ __launch_bounds__(kBlock) void k_parity_keys(int ncell, const uint64_t* __restrict__ ckeys,
                                                       uint64_t* __restrict__ pk, int32_t* __restrict__ idx,
                                                       int32_t* __restrict__ pcount) {
    __shared__ int32_t s[8];
    if (threadIdx.x < 8) s[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncell) {
        const int par = static_cast<int>(ckeys[i] & 7u);
        pk[i] = static_cast<uint64_t>(par);
        idx[i] = i;
        atomicAdd(&s[par], 1);
    }
    __syncthreads();
    if (threadIdx.x < 8 && s[threadIdx.x]) atomicAdd(&pcount[threadIdx.x], s[threadIdx.x]);
}

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Occupied M2L slots of one level counted per (parity, slot): counts[8][189].
__global__ __launch_bounds__(kBlock) void k_m2l_slot_counts(int ncell, const uint64_t* __restrict__ keys,
                                                           const int32_t* __restrict__ slots,
                                                           unsigned long long* __restrict__ counts) {
    __shared__ unsigned int bins[8 * kFarSlots];
    for (int j = threadIdx.x; j < 8 * kFarSlots; j += blockDim.x) bins[j] = 0u;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncell) {
        const int par = static_cast<int>(keys[i] & 7u);
        for (int j = 0; j < kFarSlots; ++j)
            if (slots[static_cast<int64_t>(i) * kFarSlots + j] >= 0) atomicAdd(&bins[par * kFarSlots + j], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 8 * kFarSlots; j += blockDim.x)
        if (bins[j]) atomicAdd(&counts[j], static_cast<unsigned long long>(bins[j]));
}

// Run-length groups of sorted keys >> shift: writes group keys and starts,
// returns the count on the host.
// P2P work items of the big leaves: ceil(count / kP2PItem) items of
// consecutive targets per leaf with more than kP2PBig particles.
template <int ITEM>
__global__ __launch_bounds__(kBlock) void k_p2p_item_counts(int nleaf, const int32_t* __restrict__ lstart,
                                                           int32_t* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int m = lstart[i + 1] - lstart[i];
    cnt[i] = m > kP2PBig ? (m + ITEM - 1) / ITEM : 0;
}

template <int ITEM>
__global__ __launch_bounds__(kBlock) void k_p2p_items(int nleaf, const int32_t* __restrict__ lstart,
                                                     const int32_t* __restrict__ off, int2* __restrict__ items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int m = lstart[i + 1] - lstart[i];
    if (m <= kP2PBig) return;
    const int k = (m + ITEM - 1) / ITEM;
    for (int q = 0; q < k; ++q) items[off[i] + q] = make_int2(i, lstart[i] + q * ITEM);
}

// Sort key of a big-leaf P2P work item: the items are taken costliest first,
// so the grid ends on small items instead of waiting for a large one taken
// late.  Cost = targets x sources the item visits: its own leaf and every
// neighbour (ordered pairs), or for the mutual kernel its own leaf, its small
// neighbours and its big forward neighbours (the backward ones are skipped).
// Keys ascend with falling cost; the padding past the item count sorts last.
template <int ITEM, bool MUTUAL>
__global__ __launch_bounds__(kBlock) void k_p2p_item_cost(int64_t cap, const int32_t* __restrict__ nitems,
                                                         const int2* __restrict__ items,
                                                         const int32_t* __restrict__ lstart,
                                                         const int32_t* __restrict__ slots, uint64_t* __restrict__ keys,
                                                         int32_t* __restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    uint64_t key = 0xFFFFFFFFull;
    if (i < *nitems) {
        const int2 it = items[i];
        const int leaf = it.x;
        int64_t src = lstart[leaf + 1] - lstart[leaf];
        for (int s = 0; s < kNbrSlots; ++s) {
            const int nb = slots[static_cast<int64_t>(leaf) * kNbrSlots + s];
            if (nb < 0) continue;
            const int sz = lstart[nb + 1] - lstart[nb];
            if (!MUTUAL || sz <= kP2PBig || s < kNbrForward) src += sz;
        }
        const int64_t tc = min(static_cast<int64_t>(ITEM), static_cast<int64_t>(lstart[leaf + 1] - it.y));
        const uint64_t cost = min(static_cast<uint64_t>(tc * src), 0xFFFFFFFEull);
        key = 0xFFFFFFFEull - cost;
    }
    keys[i] = key;
    vals[i] = static_cast<int32_t>(i);
}

__global__ __launch_bounds__(kBlock) void k_gather_items(int64_t cap, const int32_t* __restrict__ nitems,
                                                        const int32_t* __restrict__ idx, const int2* __restrict__ in,
                                                        int2* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cap || i >= *nitems) return;
    out[i] = in[idx[i]];
}

// Sorts the work items (count on the device, at most cap) costliest first, a
// stable sort: equal costs keep the leaf order.  A work item's results (and
// the mutual kernel's slots) follow from its leaf and first target, not from
// its place in the list, so the order changes no sum.
template <int ITEM, bool MUTUAL>
int sort_p2p_items(Context* c, int2* items, const int32_t* nitems, int64_t cap) {
    if (cap <= 1) return HFMM_OK;
    HFMM_TRY(ensure(c, c->d_itemkeys, sizeof(uint64_t) * 2 * static_cast<size_t>(cap)));
    HFMM_TRY(ensure(c, c->d_itemidx, sizeof(int32_t) * 2 * static_cast<size_t>(cap)));
    HFMM_TRY(ensure(c, c->d_items_tmp, sizeof(int2) * static_cast<size_t>(cap)));
    uint64_t* ik = c->d_itemkeys.as<uint64_t>();
    int32_t* iv = c->d_itemidx.as<int32_t>();
    prof_begin(c, "p2p_items");
    hipLaunchKernelGGL((k_p2p_item_cost<ITEM, MUTUAL>), dim3(blocks_for(cap)), dim3(kBlock), 0, c->stream, cap, nitems,
                       items, c->d_leafstart.as<int32_t>(), c->d_p2p.as<int32_t>(), ik, iv);
    prof_end(c);
    HFMM_TRY(check_launch(c, "p2p_items"));
    HFMM_TRY(radix_sort(c, ik, iv, ik + cap, iv + cap, cap, 32));
    HFMM_HIP(c, hipMemcpyAsync(c->d_items_tmp.ptr, items, sizeof(int2) * static_cast<size_t>(cap),
                               hipMemcpyDeviceToDevice, c->stream));
    prof_begin(c, "p2p_items");
    hipLaunchKernelGGL(k_gather_items, dim3(blocks_for(cap)), dim3(kBlock), 0, c->stream, cap, nitems, iv,
                       c->d_items_tmp.as<int2>(), items);
    prof_end(c);
    return check_launch(c, "p2p_items");
}

// Slot tables of the mutual mixed P2P.  A work item of big leaf A (tile t of
// kP2PMutualItem targets) evaluates A against its big forward neighbours
// (slots s < 13) once for both sides; the source-side sums for the particles
// of neighbour B go to B's slot csoff[B][12 - s] + t, A being B's backward
// neighbour in slot 25 - s.  Per big leaf B: the 13 offsets (one per backward
// slot 13..25, -1 when that neighbour is absent or small) and units[B] =
// (number of slots) x (particles of B) float4 entries.
__global__ __launch_bounds__(kBlock) void k_colslot_counts(int nleaf, const int32_t* __restrict__ lstart,
                                                          const int32_t* __restrict__ slots,
                                                          int32_t* __restrict__ csoff, int64_t* __restrict__ units) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int m = lstart[i + 1] - lstart[i];
    int run = 0;
    for (int b = 0; b < kNbrForward; ++b) {
        int o = -1;
        const int a = slots[static_cast<int64_t>(i) * kNbrSlots + kNbrForward + b];
        if (m > kP2PBig && a >= 0) {
            const int na = lstart[a + 1] - lstart[a];
            if (na > kP2PBig) {
                o = run;
                run += (na + kP2PMutualItem - 1) / kP2PMutualItem;
            }
        }
        csoff[static_cast<int64_t>(i) * kNbrForward + b] = o;
    }
    units[i] = static_cast<int64_t>(run) * m;
}

int make_groups(Context* c, const uint64_t* keys, int64_t n, int shift, uint64_t* gkey, int32_t* gstart,
                int32_t* group_of, int64_t& count, const char* tag) {
    HFMM_TRY(ensure(c, c->d_flags, sizeof(int32_t) * static_cast<size_t>(n + 1)));
    HFMM_TRY(ensure(c, c->d_scan, sizeof(int32_t) * static_cast<size_t>(n + 1)));
    int32_t* flags = c->d_flags.as<int32_t>();
    int32_t* pos = c->d_scan.as<int32_t>();
    int32_t* total = reinterpret_cast<int32_t*>(c->d_count.as<int64_t>() + 14);
    prof_begin(c, tag);
    hipLaunchKernelGGL(k_head_flags, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, keys, shift, flags);
    prof_end(c);
    HFMM_TRY(check_launch(c, tag));
    HFMM_TRY(exclusive_scan<int32_t>(c, flags, pos, n, total, "group_scan"));
    prof_begin(c, tag);
    hipLaunchKernelGGL(k_write_groups, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, keys, shift, flags, pos,
                       gkey, gstart, group_of);
    prof_end(c);
    HFMM_TRY(check_launch(c, tag));
    int32_t h_total = 0;
    HFMM_HIP(c, hipMemcpyAsync(&h_total, total, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    count = h_total;
    return HFMM_OK;
}

}  // namespace

int m2l_slot_counts(Context* c, int level, unsigned long long* counts) {
    const int64_t nc = c->ncell[level];
    HFMM_TRY(zero_device(c, counts, sizeof(unsigned long long) * 8 * kFarSlots));
    if (nc == 0) return HFMM_OK;
    hipLaunchKernelGGL(k_m2l_slot_counts, dim3(blocks_for(nc)), dim3(kBlock), 0, c->stream, static_cast<int>(nc),
                       c->d_cellkey.as<uint64_t>() + c->lvl_off[level],
                       c->d_m2l.as<int32_t>() + static_cast<size_t>(c->m2l_off[level]) * kFarSlots, counts);
    return check_launch(c, "m2l_slot_counts");
}

int update_charges_device(Context* c, int rhs) {
    prof_begin(c, "gather_charges");
    hipLaunchKernelGGL(k_gather_charges, dim3(blocks_for(c->n)), dim3(kBlock), 0, c->stream, c->n,
                       c->d_perm.as<int32_t>(), c->q + static_cast<size_t>(rhs) * c->n,
                       c->d_sorted.as<double>() + 3 * c->n);
    prof_end(c);
    return check_launch(c, "gather_charges");
}

int gather_all_charges(Context* c, bool f32) {
    const int64_t npad = (c->n + 7) & ~int64_t(7);  // chunk starts 64-byte aligned
    HFMM_TRY(ensure(c, c->d_qall, (f32 ? sizeof(float) : sizeof(double)) * static_cast<size_t>(npad) * c->nrhs));
    RhsChunk ch[HFMM_MAX_RHS];
    const int nch = rhs_chunks(c->nrhs, ch);
    prof_begin(c, "gather_charges");
    for (int k = 0; k < nch; ++k) {
        if (f32)
            hipLaunchKernelGGL(k_gather_charge_chunk<float>, dim3(blocks_for(c->n)), dim3(kBlock), 0, c->stream, c->n,
                               c->d_perm.as<int32_t>(), c->q, ch[k].r0, ch[k].R,
                               c->d_qall.as<float>() + npad * ch[k].r0);
        else
            hipLaunchKernelGGL(k_gather_charge_chunk<double>, dim3(blocks_for(c->n)), dim3(kBlock), 0, c->stream,
                               c->n, c->d_perm.as<int32_t>(), c->q, ch[k].r0, ch[k].R,
                               c->d_qall.as<double>() + npad * ch[k].r0);
    }
    prof_end(c);
    return check_launch(c, "gather_charges");
}

int build_tree_device(Context* c) {
    const int64_t n = c->n;
    const int h = c->height();  // of the virtual tree
    const int L = h - 1;
    const int Lloc = c->P.tree_height - 1;  // leaf level inside one set
    const int top = c->top();
    if (n <= 0) return fail(c, HFMM_ERR_ARG, "no particles");
    if (n > (int64_t(1) << 30)) return fail(c, HFMM_ERR_ARG, "more than 2^30 particles");
    HFMM_TRY(ensure(c, c->d_keys, sizeof(uint64_t) * n));
    HFMM_TRY(ensure(c, c->d_keys2, sizeof(uint64_t) * n));
    HFMM_TRY(ensure(c, c->d_vals, sizeof(int32_t) * n));
    HFMM_TRY(ensure(c, c->d_vals2, sizeof(int32_t) * n));
    HFMM_TRY(ensure(c, c->d_count, sizeof(int64_t) * 24));
    uint64_t* keys = c->d_keys.as<uint64_t>();
    int32_t* vals = c->d_vals.as<int32_t>();
    int32_t* err = reinterpret_cast<int32_t*>(c->d_count.as<int64_t>() + 13);
    HFMM_HIP(c, hipMemsetAsync(err, 0, sizeof(int32_t), c->stream));

    // 1. keys (boxCorner = center + (-(width/2)), FGroupTree.hpp:183)
    const double half = c->P.box_width / 2;
    const double leafw = c->P.box_width / static_cast<double>(1 << Lloc);
    prof_begin(c, "morton");
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, c->x, c->y, c->z,
                       c->P.box_center[0] + (-half), c->P.box_center[1] + (-half), c->P.box_center[2] + (-half),
                       c->P.box_width, leafw, Lloc, c->nsets > 1 ? c->d_setoff.as<int64_t>() : nullptr, c->nsets,
                       keys, vals, err);
    prof_end(c);
    HFMM_TRY(check_launch(c, "morton"));
    int32_t h_err = 0;
    HFMM_HIP(c, hipMemcpyAsync(&h_err, err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    if (h_err) return fail(c, HFMM_ERR_OUTSIDE_BOX, "a particle lies outside the simulation box");

    // 2. stable sort by key (ties keep the input order)
    HFMM_TRY(radix_sort(c, keys, vals, c->d_keys2.as<uint64_t>(), c->d_vals2.as<int32_t>(), n, 3 * L));

    // 3. sorted particle SoA
    // kSortedPad doubles after the four SoA arrays: the big-leaf P2P reads whole
    // blocks of sources past a leaf's end (and masks them)
    HFMM_TRY(ensure(c, c->d_sorted, sizeof(double) * (4 * n + kSortedPad)));
    HFMM_TRY(ensure(c, c->d_perm, sizeof(int32_t) * n));
    HFMM_HIP(c, hipMemcpyAsync(c->d_perm.ptr, vals, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, c->stream));
    // the results are written back in input order through the inverse
    HFMM_TRY(ensure(c, c->d_iperm, sizeof(int32_t) * n));
    hipLaunchKernelGGL(k_invert_perm, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, vals,
                       c->d_iperm.as<int32_t>());
    HFMM_TRY(check_launch(c, "invert_perm"));
    prof_begin(c, "gather");
    hipLaunchKernelGGL(k_gather_particles, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, vals, c->x, c->y, c->z,
                       c->q, c->d_sorted.as<double>());
    prof_end(c);
    HFMM_TRY(check_launch(c, "gather"));

    // 4. leaves, then the parent levels down to 2; the cell keys of every level
    // are first produced into a per-level scratch table of capacity n.
    const int64_t cap = n;
    HFMM_TRY(ensure(c, c->d_cellkey_tmp, sizeof(uint64_t) * static_cast<size_t>(cap) * h));
    HFMM_TRY(ensure(c, c->d_leafstart, sizeof(int32_t) * (n + 1)));
    HFMM_TRY(ensure(c, c->d_pleaf, sizeof(int32_t) * n));
    // child-start table: per level ncell+1 entries, at most cap+1 each
    HFMM_TRY(ensure(c, c->d_childstart, sizeof(int32_t) * static_cast<size_t>(cap + 1) * h));
    uint64_t* tkeys = c->d_cellkey_tmp.as<uint64_t>();
    int32_t* cstart = c->d_childstart.as<int32_t>();
    for (int l = 0; l < kMaxLevels; ++l) c->ncell[l] = 0;
    int64_t nl = 0;
    HFMM_TRY(make_groups(c, keys, n, 0, tkeys + static_cast<size_t>(L) * cap, c->d_leafstart.as<int32_t>(),
                         c->d_pleaf.as<int32_t>(), nl, "leaves"));
    c->ncell[L] = nl;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->d_leafstart.as<int32_t>() + nl,
                       static_cast<int32_t>(n));
    for (int l = L - 1; l >= top; --l) {
        int64_t np = 0;
        int32_t* cs = cstart + static_cast<size_t>(l) * (cap + 1);
        HFMM_TRY(make_groups(c, tkeys + static_cast<size_t>(l + 1) * cap, c->ncell[l + 1], 3,
                             tkeys + static_cast<size_t>(l) * cap, cs, nullptr, np, "levels"));
        c->ncell[l] = np;
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, cs + np, static_cast<int32_t>(c->ncell[l + 1]));
        c->cs_off[l] = static_cast<int64_t>(l) * (cap + 1);
    }
    // pack the cell keys of levels top..h-1 contiguously
    int64_t off = 0;
    for (int l = 0; l <= L; ++l) {
        c->lvl_off[l] = off;
        if (l >= top) off += c->ncell[l];
    }
    c->lvl_off[L + 1] = off;
    c->total_cells = off;
    HFMM_TRY(ensure(c, c->d_cellkey, sizeof(uint64_t) * static_cast<size_t>(off)));
    for (int l = top; l <= L; ++l)
        HFMM_HIP(c, hipMemcpyAsync(c->d_cellkey.as<uint64_t>() + c->lvl_off[l], tkeys + static_cast<size_t>(l) * cap,
                                   sizeof(uint64_t) * c->ncell[l], hipMemcpyDeviceToDevice, c->stream));

    // dense Morton -> cell tables for the levels with at most 2^24 cells
    int64_t dense_total = 0;
    int64_t dense_off[kMaxLevels];
    for (int l = 0; l <= L; ++l) {
        dense_off[l] = -1;
        if (l >= top && 3 * l <= 24) {
            dense_off[l] = dense_total;
            dense_total += int64_t(1) << (3 * l);
        }
    }
    HFMM_TRY(ensure(c, c->d_dense, sizeof(int32_t) * static_cast<size_t>(dense_total + 1)));
    HFMM_HIP(c, hipMemsetAsync(c->d_dense.ptr, 0xFF, sizeof(int32_t) * static_cast<size_t>(dense_total), c->stream));
    KeyIndex ix[kMaxLevels];
    for (int l = top; l <= L; ++l) {
        ix[l].keys = c->d_cellkey.as<uint64_t>() + c->lvl_off[l];
        ix[l].n = static_cast<int>(c->ncell[l]);
        ix[l].dense = dense_off[l] >= 0 ? c->d_dense.as<int32_t>() + dense_off[l] : nullptr;
        if (ix[l].dense) {
            prof_begin(c, "dense_index");
            hipLaunchKernelGGL(k_dense_index, dim3(blocks_for(c->ncell[l])), dim3(kBlock), 0, c->stream, ix[l].n,
                               ix[l].keys, c->d_dense.as<int32_t>() + dense_off[l]);
            prof_end(c);
            HFMM_TRY(check_launch(c, "dense_index"));
        }
    }

    // 5. P2P slots of the leaves
    HFMM_TRY(ensure(c, c->d_p2p, sizeof(int32_t) * static_cast<size_t>(nl) * kNbrSlots));
    prof_begin(c, "p2p_lists");
    hipLaunchKernelGGL(k_p2p_slots, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, ix[L], 1 << Lloc,
                       c->d_p2p.as<int32_t>());
    prof_end(c);
    HFMM_TRY(check_launch(c, "p2p_lists"));
    {
        // items of the big leaves; their count stays on the device (slot 12 of d_count)
        c->p2p_items_cap = n / kP2PItem + n / (kP2PBig + 1) + 1;
        HFMM_TRY(ensure(c, c->d_p2pitems, sizeof(int2) * static_cast<size_t>(c->p2p_items_cap)));
        HFMM_TRY(ensure(c, c->d_flags, sizeof(int32_t) * static_cast<size_t>(nl + 1)));
        HFMM_TRY(ensure(c, c->d_scan, sizeof(int32_t) * static_cast<size_t>(nl + 1)));
        prof_begin(c, "p2p_items");
        hipLaunchKernelGGL(k_p2p_item_counts<kP2PItem>, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, static_cast<int>(nl),
                           c->d_leafstart.as<int32_t>(), c->d_flags.as<int32_t>());
        prof_end(c);
        HFMM_TRY(check_launch(c, "p2p_items"));
        HFMM_TRY(exclusive_scan<int32_t>(c, c->d_flags.as<int32_t>(), c->d_scan.as<int32_t>(), nl, p2p_item_count(c),
                                         "p2p_items"));
        prof_begin(c, "p2p_items");
        hipLaunchKernelGGL(k_p2p_items<kP2PItem>, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, static_cast<int>(nl),
                           c->d_leafstart.as<int32_t>(), c->d_scan.as<int32_t>(), c->d_p2pitems.as<int2>());
        prof_end(c);
        HFMM_TRY(check_launch(c, "p2p_items"));
        // the ordered big-leaf kernels take the items at a fixed stride: costliest
        // first spreads the large ones over the waves
        HFMM_TRY((sort_p2p_items<kP2PItem, false>(c, c->d_p2pitems.as<int2>(), p2p_item_count(c), c->p2p_items_cap)));
    }

    // 6. M2L slots and parity-sorted target order of levels top..h-1
    HFMM_TRY(ensure(c, c->d_m2l, sizeof(int32_t) * static_cast<size_t>(off) * kFarSlots));
    HFMM_TRY(ensure(c, c->d_order, sizeof(int32_t) * static_cast<size_t>(off)));
    HFMM_TRY(ensure(c, c->d_pcount, sizeof(int32_t) * 8 * kMaxLevels));
    HFMM_HIP(c, hipMemsetAsync(c->d_pcount.ptr, 0, sizeof(int32_t) * 8 * kMaxLevels, c->stream));
    for (int l = top; l <= L; ++l) {
        const int64_t nc = c->ncell[l];
        c->m2l_off[l] = c->lvl_off[l];
        const uint64_t* ck = c->d_cellkey.as<uint64_t>() + c->lvl_off[l];
        prof_begin(c, "m2l_lists");
        hipLaunchKernelGGL(k_m2l_slots, dim3(blocks_for(nc)), dim3(kBlock), 0, c->stream, ix[l],
                           1 << (l - 1 - c->set_levels),
                           c->d_m2l.as<int32_t>() + static_cast<size_t>(c->m2l_off[l]) * kFarSlots);
        prof_end(c);
        HFMM_TRY(check_launch(c, "m2l_lists"));
        // parity order: stable sort of (key & 7) carrying the cell index
        int32_t* ord = c->d_order.as<int32_t>() + c->lvl_off[l];
        prof_begin(c, "parity");
        // (the per-level key scratch is free once the keys are packed)
        hipLaunchKernelGGL(k_parity_keys, dim3(blocks_for(nc)), dim3(kBlock), 0, c->stream, static_cast<int>(nc), ck,
                           tkeys, ord, c->d_pcount.as<int32_t>() + 8 * l);
        prof_end(c);
        HFMM_TRY(check_launch(c, "parity"));
        HFMM_TRY(radix_sort(c, tkeys, ord, tkeys + cap, c->d_vals2.as<int32_t>(), nc, 3));
    }
    c->tree_built = true;
    c->evaluated = false;
    c->p2p_mutual = false;
    if (is_mixed(c)) HFMM_TRY(build_p2p_mutual_tables(c));
    return HFMM_OK;
}

int build_p2p_mutual_tables(Context* c) {
    c->p2p_mutual = false;
    const int64_t nl = c->ncell[c->height() - 1];
    const int32_t* lstart = c->d_leafstart.as<int32_t>();
    c->p2p_items_m_cap = c->n / kP2PMutualItem + c->n / (kP2PBig + 1) + 1;
    HFMM_TRY(ensure(c, c->d_p2pitems_m, sizeof(int2) * static_cast<size_t>(c->p2p_items_m_cap)));
    HFMM_TRY(ensure(c, c->d_flags, sizeof(int32_t) * static_cast<size_t>(nl + 1)));
    HFMM_TRY(ensure(c, c->d_scan, sizeof(int32_t) * static_cast<size_t>(nl + 1)));
    prof_begin(c, "p2p_items");
    hipLaunchKernelGGL(k_p2p_item_counts<kP2PMutualItem>, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream,
                       static_cast<int>(nl), lstart, c->d_flags.as<int32_t>());
    prof_end(c);
    HFMM_TRY(check_launch(c, "p2p_items"));
    HFMM_TRY(exclusive_scan<int32_t>(c, c->d_flags.as<int32_t>(), c->d_scan.as<int32_t>(), nl,
                                     p2p_mutual_item_count(c), "p2p_items"));
    prof_begin(c, "p2p_items");
    hipLaunchKernelGGL(k_p2p_items<kP2PMutualItem>, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream,
                       static_cast<int>(nl), lstart, c->d_scan.as<int32_t>(), c->d_p2pitems_m.as<int2>());
    prof_end(c);
    HFMM_TRY(check_launch(c, "p2p_items"));
    HFMM_TRY((sort_p2p_items<kP2PMutualItem, true>(c, c->d_p2pitems_m.as<int2>(), p2p_mutual_item_count(c),
                                                    c->p2p_items_m_cap)));
    HFMM_TRY(ensure(c, c->d_csoff, sizeof(int32_t) * static_cast<size_t>(nl) * kNbrForward));
    HFMM_TRY(ensure(c, c->d_csbase, sizeof(int64_t) * static_cast<size_t>(nl + 1)));
    int64_t* base = c->d_csbase.as<int64_t>();
    prof_begin(c, "p2p_items");
    hipLaunchKernelGGL(k_colslot_counts, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, static_cast<int>(nl), lstart,
                       c->d_p2p.as<int32_t>(), c->d_csoff.as<int32_t>(), base);
    prof_end(c);
    HFMM_TRY(check_launch(c, "p2p_items"));
    HFMM_TRY(exclusive_scan<int64_t>(c, base, base, nl, p2p_mutual_units(c), "p2p_items"));
    int64_t units = 0;
    HFMM_HIP(c, hipMemcpyAsync(&units, p2p_mutual_units(c), sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HFMM_HIP(c, hipStreamSynchronize(c->stream));
    // base[nl] = the total, so a leaf's slot count is (base[i+1] - base[i]) / n_i
    HFMM_HIP(c, hipMemcpyAsync(base + nl, p2p_mutual_units(c), sizeof(int64_t), hipMemcpyDeviceToDevice, c->stream));
    const size_t bytes = sizeof(float) * 4 * static_cast<size_t>(units > 0 ? units : 1);
    if (bytes > c->d_cpart.bytes) {
        size_t free_b = 0, total_b = 0;
        HFMM_HIP(c, hipMemGetInfo(&free_b, &total_b));
        if (bytes > (free_b + c->d_cpart.bytes) / 4) return HFMM_OK;  // ordered-pair kernel instead
    }
    HFMM_TRY(ensure(c, c->d_cpart, bytes));
    c->cpart_units = units;
    c->p2p_mutual = true;
    return HFMM_OK;
}

}  // namespace hfmm
