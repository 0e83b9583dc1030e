// p2p_mixed.hip -- the near field (P2P) of the mixed-precision form
// (HFMM_PRECISION_MIXED and _HILO): FP32 pair terms on leaf-local fixed-point
// coordinates (exact integer separations, see k_local_coords), 1/r by rsqrt,
// per-target sums compensated (Kahan) once per block of 16 pair terms, results
// written in FP64.  Leaves of at most kP2PBig particles run one thread per
// target (k_p2p_mixed); bigger leaves run in waves over work items, either
// every ordered pair (k_p2p_mixed_big) or every unordered pair of two big
// neighbour leaves once (k_p2p_mixed_mutual + k_p2p_combine).
//
// Reference operators: Src/Kernels/P2P/FP2P.hpp:103-269,
// Src/Kernels/Chebyshev/FChebSymKernel.hpp:468-478.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fmm_common.hpp"
#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

inline unsigned blocks_of(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Leaf-local fixed-point coordinates of the sorted particles: the offset from
// the leaf centre in units of 2^-30 leaf widths, clamped to |value| < 2^29
// (one unit off at most, on the leaf faces), charge as FP32 bits in .w.  A
// separation (t - c) - s, c = the source leaf's centre offset (0 or +-2^30
// units), is exact in int32 (|.| <= 2^31 - 2) and rounds once when converted
// to FP32, relative to the separation itself; the coordinate resolution is
// 2^-31 leaf widths.
constexpr double kUnitsPerLeaf = 1073741824.;  // 2^30
constexpr int kHalfLeaf = 1 << 29;

__global__ __launch_bounds__(256) void k_local_coords(int64_t n, const int32_t* __restrict__ pleaf,
                                                     const uint64_t* __restrict__ lkeys, Box bx,
                                                     const double* __restrict__ sp, int4* __restrict__ loc) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[3], b[3];
    leaf_bounds(bx, lkeys[pleaf[i]], a, b);
    int v[3];
    for (int d = 0; d < 3; ++d) {
        const double x = (sp[d * n + i] - (a[d] + b[d]) / 2.) / bx.leafw;  // in [-1/2, 1/2]
        const double u = rint(x * kUnitsPerLeaf);
        v[d] = static_cast<int>(fmin(fmax(u, -(kHalfLeaf - 1.)), kHalfLeaf - 1.));
    }
    loc[i] = make_int4(v[0], v[1], v[2], __float_as_int(static_cast<float>(sp[3 * n + i])));
}

// Compensated (Kahan) add of a block partial into (sum, comp).
__device__ __forceinline__ void kahan_add(float& sum, float& comp, float part) {
    const float y = part - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// One source against one target; t = the target's coordinates minus the
// source leaf's centre offset.  With MASK the pair is dropped when `skip`
// holds (sources past a leaf's end) or when the two positions coincide: r2 = 0
// is the pair of a target with itself, or two particles less than 2^-31 leaf
// widths apart, which the fixed-point coordinates cannot separate.  Without
// MASK the source lies in another leaf and r2 >= 4 units^2 (k_local_coords
// keeps every position at least one unit inside its leaf's faces).
template <bool MASK>
__device__ __forceinline__ void pair_f32(const int (&t)[3], int4 s, bool skip, float (&part)[4]) {
    const float dx = static_cast<float>(t[0] - s.x);
    const float dy = static_cast<float>(t[1] - s.y);
    const float dz = static_cast<float>(t[2] - s.z);
    const float r2 = dx * dx + dy * dy + dz * dz;
    float inv = __builtin_amdgcn_rsqf(r2);
    if (MASK) inv = (skip || r2 == 0.f) ? 0.f : inv;
    const float qi = __int_as_float(s.w) * inv;
    const float qi3 = qi * (inv * inv);
    part[0] += qi;
    part[1] -= qi3 * dx;
    part[2] -= qi3 * dy;
    part[3] -= qi3 * dz;
}

// Centre offset of neighbour slot s (-1 = the leaf itself) in coordinate
// units: slot s is pos = s (s < 13) or s + 1 of the 3x3x3 block,
// pos = 9 (dx+1) + 3 (dy+1) + (dz+1).
__device__ __forceinline__ void slot_offset(int s, int (&c)[3]) {
    const int pos = s < 0 ? 13 : (s < 13 ? s : s + 1);
    c[0] = (pos / 9 - 1) * (2 * kHalfLeaf);
    c[1] = ((pos / 3) % 3 - 1) * (2 * kHalfLeaf);
    c[2] = (pos % 3 - 1) * (2 * kHalfLeaf);
}

constexpr int kKahanBlock = 16;  // pair terms summed plainly before a compensated add

// Scales of the unit-coordinate sums: potential / (leaf width 2^-30), field /
// (leaf width 2^-30)^2 (the charge is applied when the results are scattered).
__device__ __forceinline__ void write_p2p(int64_t n, int64_t i, const float (&sum)[4], const float (&comp)[4],
                                          double inv_u, double* __restrict__ res) {
    const double qf = inv_u * inv_u;
    res[i] = (static_cast<double>(sum[0]) - static_cast<double>(comp[0])) * inv_u;
    res[n + i] = (static_cast<double>(sum[1]) - static_cast<double>(comp[1])) * qf;
    res[2 * n + i] = (static_cast<double>(sum[2]) - static_cast<double>(comp[2])) * qf;
    res[3 * n + i] = (static_cast<double>(sum[3]) - static_cast<double>(comp[3])) * qf;
}

// Mixed P2P of the leaves with at most kP2PBig particles: one thread per
// target (sorted order), every pair of its 27 leaves (the pair with itself
// has r2 = 0 and is dropped).
__global__ __launch_bounds__(256) void k_p2p_mixed(int64_t n, const int32_t* __restrict__ pleaf,
                                                  const int32_t* __restrict__ lstart, const int32_t* __restrict__ slots,
                                                  const int4* __restrict__ loc, double inv_u, double* __restrict__ res) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = pleaf[i];
    if (lstart[li + 1] - lstart[li] > kP2PBig) return;  // k_p2p_mixed_big / k_p2p_mixed_mutual
    const int4 t = loc[i];
    float sum[4] = {0.f, 0.f, 0.f, 0.f}, comp[4] = {0.f, 0.f, 0.f, 0.f};
    for (int s = -1; s < kNbrSlots; ++s) {
        const int nb = s < 0 ? li : slots[static_cast<int64_t>(li) * kNbrSlots + s];
        if (nb < 0) continue;
        int c[3];
        slot_offset(s, c);
        const int tc[3] = {t.x - c[0], t.y - c[1], t.z - c[2]};
        const int j0 = lstart[nb], j1 = lstart[nb + 1];
        for (int jb = j0; jb < j1; jb += kKahanBlock) {
            float part[4] = {0.f, 0.f, 0.f, 0.f};
            const int je = min(j1, jb + kKahanBlock);
            for (int j = jb; j < je; ++j) pair_f32<true>(tc, loc[j], false, part);
#pragma unroll
            for (int k = 0; k < 4; ++k) kahan_add(sum[k], comp[k], part[k]);
        }
    }
    write_p2p(n, i, sum, comp, inv_u, res);
}

// k_p2p_mixed for R charge vectors at once: the separation, 1/r and 1/r^2 of
// a pair once, then the terms of each vector, each summed and compensated as
// in k_p2p_mixed (the same bits per vector).  qc: the chunk's sorted charges
// (FP32) as [n][R]; res: the results of its first vector, the next vectors 4n
// further each.
template <int R>
__global__ __launch_bounds__(256) void k_p2p_mixed_rhs(int64_t n, const int32_t* __restrict__ pleaf,
                                                      const int32_t* __restrict__ lstart,
                                                      const int32_t* __restrict__ slots, const int4* __restrict__ loc,
                                                      const float* __restrict__ qc, double inv_u,
                                                      double* __restrict__ res) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = pleaf[i];
    if (lstart[li + 1] - lstart[li] > kP2PBig) return;  // the big-leaf kernels, per vector
    const int4 t = loc[i];
    float sum[R][4] = {}, comp[R][4] = {};
    for (int s = -1; s < kNbrSlots; ++s) {
        const int nb = s < 0 ? li : slots[static_cast<int64_t>(li) * kNbrSlots + s];
        if (nb < 0) continue;
        int c[3];
        slot_offset(s, c);
        const int tc[3] = {t.x - c[0], t.y - c[1], t.z - c[2]};
        const int j0 = lstart[nb], j1 = lstart[nb + 1];
        for (int jb = j0; jb < j1; jb += kKahanBlock) {
            float part[R][4] = {};
            const int je = min(j1, jb + kKahanBlock);
            for (int j = jb; j < je; ++j) {
                const int4 sj = loc[j];
                const float dx = static_cast<float>(tc[0] - sj.x);
                const float dy = static_cast<float>(tc[1] - sj.y);
                const float dz = static_cast<float>(tc[2] - sj.z);
                const float r2 = dx * dx + dy * dy + dz * dz;
                float inv = __builtin_amdgcn_rsqf(r2);
                inv = r2 == 0.f ? 0.f : inv;  // the target itself (or an unresolvable pair)
                const float inv2 = inv * inv;
                const float* q = qc + static_cast<int64_t>(j) * R;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float qi = q[r] * inv;
                    const float qi3 = qi * inv2;
                    part[r][0] += qi;
                    part[r][1] -= qi3 * dx;
                    part[r][2] -= qi3 * dy;
                    part[r][3] -= qi3 * dz;
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < 4; ++k) kahan_add(sum[r][k], comp[r][k], part[r][k]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) write_p2p(n, i, sum[r], comp[r], inv_u, res + static_cast<int64_t>(r) * 4 * n);
}

// Big-leaf sources against TPL targets per lane, in blocks of 16; every
// source address is wave-uniform, so the sources arrive through scalar loads
// into SGPRs and feed the pair arithmetic as scalar operands (no LDS, no
// barriers).  A block running past the leaf's end drops the extra sources (the
// particle array has kKahanBlock padding entries, so the reads stay in
// bounds).  SELF: the target's own leaf (the pairs with r2 = 0 are dropped).
template <int TPL, bool SELF>
__device__ __forceinline__ void p2p_leaf_sources(const int4* __restrict__ loc, int j0, int j1, const int (&tc)[TPL][3],
                                                 float (&sum)[TPL][4], float (&comp)[TPL][4]) {
    int jb = j0;
    for (; jb + kKahanBlock <= j1; jb += kKahanBlock) {
        float part[TPL][4] = {};
#pragma unroll
        for (int jj = 0; jj < kKahanBlock; ++jj) {
            const int4 s = loc[jb + jj];
#pragma unroll
            for (int u = 0; u < TPL; ++u) pair_f32<SELF>(tc[u], s, false, part[u]);
        }
#pragma unroll
        for (int u = 0; u < TPL; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) kahan_add(sum[u][k], comp[u][k], part[u][k]);
    }
    if (jb < j1) {
        float part[TPL][4] = {};
#pragma unroll
        for (int jj = 0; jj < kKahanBlock; ++jj) {
            const int4 s = loc[jb + jj];
            const bool past = jb + jj >= j1;
#pragma unroll
            for (int u = 0; u < TPL; ++u) pair_f32<true>(tc[u], s, past, part[u]);
        }
#pragma unroll
        for (int u = 0; u < TPL; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) kahan_add(sum[u][k], comp[u][k], part[u][k]);
    }
}

// Targets of one work item: TPL per lane from `first`; a lane past the item's
// end holds a copy of the item's first target with charge 0 (never written).
template <int TPL>
__device__ __forceinline__ void load_targets(const int4* __restrict__ loc, int first, int tend, int lane, int4 (&t)[TPL],
                                             float (&qt)[TPL], int (&ti)[TPL]) {
#pragma unroll
    for (int u = 0; u < TPL; ++u) {
        ti[u] = first + 64 * u + lane;
        const bool ok = ti[u] < tend;
        t[u] = loc[ok ? ti[u] : first];
        qt[u] = ok ? __int_as_float(t[u].w) : 0.f;
    }
}

template <int TPL>
__device__ __forceinline__ void offset_targets(const int4 (&t)[TPL], int s, int (&tc)[TPL][3]) {
    int c[3];
    slot_offset(s, c);
#pragma unroll
    for (int u = 0; u < TPL; ++u) {
        tc[u][0] = t[u].x - c[0];
        tc[u][1] = t[u].y - c[1];
        tc[u][2] = t[u].z - c[2];
    }
}

// Ordered-pair mixed P2P of the big leaves.  Each wave takes work items of up
// to kP2PItem = 64 TPL consecutive targets of one leaf (TPL per lane) and
// walks the sources of the 27 neighbour leaves (p2p_leaf_sources).
template <int TPL>
__global__ __launch_bounds__(256) void k_p2p_mixed_big(const int2* __restrict__ items,
                                                      const int32_t* __restrict__ nitems,
                                                      const int32_t* __restrict__ lstart,
                                                      const int32_t* __restrict__ slots, const int4* __restrict__ loc,
                                                      int64_t n, double inv_u, double* __restrict__ res) {
    static_assert(64 * TPL == kP2PItem, "work item = one wave x TPL targets");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int total = *nitems;
    for (int it = blockIdx.x * 4 + wave; it < total; it += gridDim.x * 4) {
        const int2 item = items[it];
        const int leaf = __builtin_amdgcn_readfirstlane(item.x);
        const int first = __builtin_amdgcn_readfirstlane(item.y);
        const int tend = min(lstart[leaf + 1], first + kP2PItem);
        int ti[TPL];
        int4 t[TPL];
        float qt[TPL];
        float sum[TPL][4] = {}, comp[TPL][4] = {};
        load_targets<TPL>(loc, first, tend, lane, t, qt, ti);
        for (int s = -1; s < kNbrSlots; ++s) {
            const int nb = __builtin_amdgcn_readfirstlane(s < 0 ? leaf : slots[static_cast<int64_t>(leaf) * kNbrSlots + s]);
            if (nb < 0) continue;
            int tc[TPL][3];
            offset_targets<TPL>(t, s, tc);
            const int j0 = __builtin_amdgcn_readfirstlane(lstart[nb]);
            const int j1 = __builtin_amdgcn_readfirstlane(lstart[nb + 1]);
            if (s < 0)
                p2p_leaf_sources<TPL, true>(loc, j0, j1, tc, sum, comp);
            else
                p2p_leaf_sources<TPL, false>(loc, j0, j1, tc, sum, comp);
        }
#pragma unroll
        for (int u = 0; u < TPL; ++u)
            if (ti[u] < tend) write_p2p(n, ti[u], sum[u], comp[u], inv_u, res);
    }
}

// ---------------------------------------------------------------------------
// Mutual mixed P2P of the big leaves: every unordered pair of two big
// neighbour leaves is evaluated once and contributes to both particles, as
// the reference's P2POuter does for the positions below 13 and its algorithms
// for the opposite direction (FChebSymKernel.hpp:468-478,
// FGroupSeqAlgorithm.hpp:343-394, FP2P.hpp:103-171).
//
// A work item is kP2PMutualItem = 64 T targets of big leaf A (T per lane),
// stationary in registers.  Its own leaf and its small neighbours are
// evaluated pair by pair as in k_p2p_mixed_big (the small leaves' own
// targets run in k_p2p_mixed).  Each big forward neighbour B (slots 0..12) is
// streamed in chunks of 64 sources, staged twice in a row in the wave's LDS,
// so that at step k lane i reads source (i + k) mod 64 at a constant offset.
// Lane i adds the pair terms of its T targets into the targets' sums and into
// a source-side sum, which moves one lane down per step with the source it
// belongs to: the rotation (DPP wave_rol:1, lane i takes lane i + 1's value)
// is fused into the add, acc_i <- acc_(i+1) + terms_i, so after step k lane i
// holds the sum of source (i + k) mod 64.  The sums restart every kPhase
// steps (kPhase T terms each, plain FP32); the phase sums go through the LDS
// to the source's own lane, which adds them (compensated) and writes the
// chunk's source-side sums to B's slot for this work item.  k_p2p_combine
// adds a particle's slots to its results in a fixed order: no atomics, the
// same bits on every run.  The big backward neighbours (slots 13..25) are
// skipped here: their work items send their part through those slots.
constexpr int kMutualT = kP2PMutualItem / 64;
constexpr int kPhase = 32;  // 2 phases per chunk: 159.8 ms vs 163.2 ms for 16 on 10M Plummer h8
constexpr int kPhases = 64 / kPhase;

__device__ __forceinline__ float rol1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}

// One pair evaluated for both sides: target sums tp (field at the target,
// -q_s d / r^3), source sums sp (field at the source, +q_t d / r^3), d = t - s.
// FIRST: sp starts with this pair's terms.
template <bool FIRST>
__device__ __forceinline__ void mutual_pair(const int (&t)[3], float qt, int4 s, float qs, float (&tp)[4],
                                            float (&sp)[4]) {
    const float dx = static_cast<float>(t[0] - s.x);
    const float dy = static_cast<float>(t[1] - s.y);
    const float dz = static_cast<float>(t[2] - s.z);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float inv = __builtin_amdgcn_rsqf(r2);
    // 1/r^3 once for both sides: the potentials fold into FMAs, 4 products per pair
    const float inv3 = inv * (inv * inv);
    const float a3 = qs * inv3;
    tp[0] = fmaf(qs, inv, tp[0]);
    tp[1] -= a3 * dx;
    tp[2] -= a3 * dy;
    tp[3] -= a3 * dz;
    const float b3 = qt * inv3;
    if (FIRST) {
        sp[0] = qt * inv;
        sp[1] = b3 * dx;
        sp[2] = b3 * dy;
        sp[3] = b3 * dz;
    } else {
        sp[0] = fmaf(qt, inv, sp[0]);
        sp[1] += b3 * dx;
        sp[2] += b3 * dy;
        sp[3] += b3 * dz;
    }
}

template <int T>
__global__ __launch_bounds__(256, 4) void k_p2p_mixed_mutual(const int2* __restrict__ items,
                                                         const int32_t* __restrict__ nitems,
                                                         const int32_t* __restrict__ lstart,
                                                         const int32_t* __restrict__ slots, const int4* __restrict__ loc,
                                                         const int32_t* __restrict__ csoff,
                                                         const int64_t* __restrict__ csbase, f4* __restrict__ cpart,
                                                         int32_t* __restrict__ next, int64_t n, double inv_u,
                                                         double* __restrict__ res) {
    static_assert(64 * T == kP2PMutualItem, "work item = one wave x T targets");
    __shared__ int4 s_src[4][128];
    __shared__ f4 s_x[4][kPhases * 64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int4* src = s_src[wave];
    f4* xch = s_x[wave];
    const int total = *nitems;
    // work items claimed one at a time (an atomic counter: the items differ in
    // cost by the sizes of their leaves' neighbours, so a fixed share per wave
    // leaves the grid waiting for the slowest); every wave leaves once the
    // counter passes the last item
    for (;;) {
        int claim = 0;
        if (lane == 0) claim = atomicAdd(next, 1);
        const int it = __builtin_amdgcn_readfirstlane(__shfl(claim, 0));
        if (it >= total) break;
        const int2 item = items[it];
        const int leaf = __builtin_amdgcn_readfirstlane(item.x);
        const int first = __builtin_amdgcn_readfirstlane(item.y);
        const int l0 = __builtin_amdgcn_readfirstlane(lstart[leaf]);
        const int tend = min(__builtin_amdgcn_readfirstlane(lstart[leaf + 1]), first + 64 * T);
        const int tile = (first - l0) / (64 * T);
        int ti[T];
        int4 t[T];
        float qt[T];
        float sum[T][4] = {}, comp[T][4] = {};
        load_targets<T>(loc, first, tend, lane, t, qt, ti);
        // ordered part: the leaf itself and its small neighbours
        for (int s = -1; s < kNbrSlots; ++s) {
            const int nb = __builtin_amdgcn_readfirstlane(s < 0 ? leaf : slots[static_cast<int64_t>(leaf) * kNbrSlots + s]);
            if (nb < 0) continue;
            const int j0 = __builtin_amdgcn_readfirstlane(lstart[nb]);
            const int j1 = __builtin_amdgcn_readfirstlane(lstart[nb + 1]);
            if (s >= 0 && j1 - j0 > kP2PBig) continue;
            int tc[T][3];
            offset_targets<T>(t, s, tc);
            if (s < 0)
                p2p_leaf_sources<T, true>(loc, j0, j1, tc, sum, comp);
            else
                p2p_leaf_sources<T, false>(loc, j0, j1, tc, sum, comp);
        }
        // mutual part: the big forward neighbours
        for (int s = 0; s < kNbrForward; ++s) {
            const int nb = __builtin_amdgcn_readfirstlane(slots[static_cast<int64_t>(leaf) * kNbrSlots + s]);
            if (nb < 0) continue;
            const int j0 = __builtin_amdgcn_readfirstlane(lstart[nb]);
            const int j1 = __builtin_amdgcn_readfirstlane(lstart[nb + 1]);
            const int nbn = j1 - j0;
            if (nbn <= kP2PBig) continue;
            int tc[T][3];
            offset_targets<T>(t, s, tc);
            // this item's slot in B (A is B's backward neighbour in slot 25 - s),
            // indexed by the sorted source index
            const int slot = __builtin_amdgcn_readfirstlane(csoff[static_cast<int64_t>(nb) * kNbrForward + 12 - s]) + tile;
            f4* cp = cpart + (csbase[nb] + static_cast<int64_t>(slot) * nbn - j0);
            for (int cb = j0; cb < j1; cb += 64) {
                const int js = cb + lane;
                int4 sv = loc[js < j1 ? js : j0];
                if (js >= j1) sv.w = 0;  // a charge-free copy of the leaf's first source
                src[lane] = sv;
                src[lane + 64] = sv;
                // the wave's LDS accesses run in order; the fence keeps the
                // compiler from moving other lanes' reads above these writes
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                for (int ph = 0; ph < kPhases; ++ph) {
                    const int4* sp = src + lane + kPhase * ph;
                    float part[T][4] = {};
                    float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int k = 0; k < kPhase; ++k) {
                        const int4 sk = sp[k];
                        const float qs = __int_as_float(sk.w);
                        float cs[4];
                        mutual_pair<true>(tc[0], qt[0], sk, qs, part[0], cs);
#pragma unroll
                        for (int u = 1; u < T; ++u) mutual_pair<false>(tc[u], qt[u], sk, qs, part[u], cs);
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[q] = rol1(acc[q]) + cs[q];
                    }
#pragma unroll
                    for (int u = 0; u < T; ++u)
#pragma unroll
                        for (int q = 0; q < 4; ++q) kahan_add(sum[u][q], comp[u][q], part[u][q]);
                    // lane i holds the phase sums of source (i + kPhase ph + kPhase - 1) mod 64
                    xch[ph * 64 + ((lane + kPhase * ph + kPhase - 1) & 63)] = f4{acc[0], acc[1], acc[2], acc[3]};
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (js < j1) {
                    const f4 v0 = xch[lane];
                    float a[4] = {v0[0], v0[1], v0[2], v0[3]}, cmp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int ph = 1; ph < kPhases; ++ph) {
                        const f4 v = xch[ph * 64 + lane];
#pragma unroll
                        for (int q = 0; q < 4; ++q) kahan_add(a[q], cmp[q], v[q]);
                    }
                    cp[js] = f4{a[0] - cmp[0], a[1] - cmp[1], a[2] - cmp[2], a[3] - cmp[3]};
                }
                // the next chunk rewrites src and xch only after these reads
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
#pragma unroll
        for (int u = 0; u < T; ++u)
            if (ti[u] < tend) write_p2p(n, ti[u], sum[u], comp[u], inv_u, res);
    }
}

// Adds the source-side sums of the mutual P2P to the results of the big
// leaves' particles: one thread per particle, its slots in order, in FP64.
__global__ __launch_bounds__(256) void k_p2p_combine(int64_t n, const int32_t* __restrict__ pleaf,
                                                    const int32_t* __restrict__ lstart,
                                                    const int64_t* __restrict__ csbase, const f4* __restrict__ cpart,
                                                    double inv_u, double* __restrict__ res) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = pleaf[i];
    const int j0 = lstart[li], m = lstart[li + 1] - j0;
    if (m <= kP2PBig) return;
    const int64_t b0 = csbase[li];
    const int64_t ns = (csbase[li + 1] - b0) / m;
    if (ns == 0) return;
    const f4* p = cpart + b0 + (i - j0);
    double s[4] = {0., 0., 0., 0.};
    for (int64_t k = 0; k < ns; ++k) {
        const f4 v = p[k * m];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += static_cast<double>(v[q]);
    }
    const double qf = inv_u * inv_u;
    res[i] += s[0] * inv_u;
    res[n + i] += s[1] * qf;
    res[2 * n + i] += s[2] * qf;
    res[3 * n + i] += s[3] * qf;
}

}  // namespace

int run_p2p_mixed(Context* c) {
    const int L = c->height() - 1;
    // kKahanBlock padding entries after the particles: read by the masked tail blocks
    HFMM_TRY(ensure(c, c->d_loc, sizeof(int4) * static_cast<size_t>(c->n + kKahanBlock)));
    const Box bx = make_box(c);
    int4* loc = c->d_loc.as<int4>();
    HFMM_TRY(zero_device(c, loc + c->n, sizeof(int4) * kKahanBlock));
    prof_begin(c, "local_coords");
    hipLaunchKernelGGL(k_local_coords, dim3(blocks_of(c->n)), dim3(256), 0, c->stream, c->n, c->d_pleaf.as<int32_t>(),
                       c->d_cellkey.as<uint64_t>() + c->lvl_off[L], bx, c->d_sorted.as<double>(), loc);
    prof_end(c);
    HFMM_TRY(check_launch(c, "local_coords"));
    const double inv_u = kUnitsPerLeaf / bx.leafw;
    if (!c->p2p_small_done) {
        prof_begin(c, "p2p_mixed");
        hipLaunchKernelGGL(k_p2p_mixed, dim3(blocks_of(c->n)), dim3(256), 0, c->stream, c->n,
                           c->d_pleaf.as<int32_t>(), c->d_leafstart.as<int32_t>(), c->d_p2p.as<int32_t>(), loc, inv_u,
                           sorted_results(c));
        prof_end(c);
        HFMM_TRY(check_launch(c, "p2p_mixed"));
    }
    if (!(c->p2p_mutual && c->p2p_mutual_on)) {
        const unsigned grid = (p2p_tiled_grid(c) + 3) / 4;
        prof_begin(c, "p2p_mixed_big");
        hipLaunchKernelGGL(k_p2p_mixed_big<kP2PItem / 64>, dim3(grid), dim3(256), 0, c->stream,
                           c->d_p2pitems.as<int2>(), p2p_item_count(c), c->d_leafstart.as<int32_t>(),
                           c->d_p2p.as<int32_t>(), loc, c->n, inv_u, sorted_results(c));
        prof_end(c);
        return check_launch(c, "p2p_mixed_big");
    }
    // persistent grid: as many workgroups as fit on the device at once
    static int resident[64] = {0};
    if (c->device < 0 || c->device >= 64) return fail(c, HFMM_ERR_ARG, "device ordinal above 63");
    if (resident[c->device] == 0) {
        int per_cu = 0, cus = 0;
        HFMM_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &per_cu, reinterpret_cast<const void*>(&k_p2p_mixed_mutual<kMutualT>), 256, 0));
        HFMM_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        resident[c->device] = std::max(1, per_cu) * std::max(1, cus);
    }
    const int64_t want = (c->p2p_items_m_cap + 3) / 4;
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(want, resident[c->device])));
    HFMM_TRY(zero_device(c, p2p_mutual_next(c), sizeof(int32_t)));  // a fill kernel: capturable
    prof_begin(c, "p2p_mixed_mutual");
    hipLaunchKernelGGL(k_p2p_mixed_mutual<kMutualT>, dim3(grid), dim3(256), 0, c->stream, c->d_p2pitems_m.as<int2>(),
                       p2p_mutual_item_count(c), c->d_leafstart.as<int32_t>(), c->d_p2p.as<int32_t>(), loc,
                       c->d_csoff.as<int32_t>(), c->d_csbase.as<int64_t>(), c->d_cpart.as<f4>(), p2p_mutual_next(c),
                       c->n, inv_u, sorted_results(c));
    prof_end(c);
    HFMM_TRY(check_launch(c, "p2p_mixed_mutual"));
    prof_begin(c, "p2p_combine");
    hipLaunchKernelGGL(k_p2p_combine, dim3(blocks_of(c->n)), dim3(256), 0, c->stream, c->n, c->d_pleaf.as<int32_t>(),
                       c->d_leafstart.as<int32_t>(), c->d_csbase.as<int64_t>(), c->d_cpart.as<f4>(), inv_u,
                       sorted_results(c));
    prof_end(c);
    return check_launch(c, "p2p_combine");
}

int run_p2p_mixed_small_all(Context* c) {
    const int L = c->height() - 1;
    HFMM_TRY(gather_all_charges(c, true));
    HFMM_TRY(ensure(c, c->d_loc, sizeof(int4) * static_cast<size_t>(c->n + kKahanBlock)));
    const Box bx = make_box(c);
    int4* loc = c->d_loc.as<int4>();
    HFMM_TRY(zero_device(c, loc + c->n, sizeof(int4) * kKahanBlock));
    prof_begin(c, "local_coords");
    hipLaunchKernelGGL(k_local_coords, dim3(blocks_of(c->n)), dim3(256), 0, c->stream, c->n, c->d_pleaf.as<int32_t>(),
                       c->d_cellkey.as<uint64_t>() + c->lvl_off[L], bx, c->d_sorted.as<double>(), loc);
    prof_end(c);
    HFMM_TRY(check_launch(c, "local_coords"));
    const double inv_u = kUnitsPerLeaf / bx.leafw;
    RhsChunk ch[HFMM_MAX_RHS];
    const int nch = rhs_chunks(c->nrhs, ch);
    const int64_t npad = (c->n + 7) & ~int64_t(7);
    const dim3 grid(blocks_of(c->n)), block(256);
    const int32_t* pl = c->d_pleaf.as<int32_t>();
    const int32_t* ls = c->d_leafstart.as<int32_t>();
    const int32_t* sl = c->d_p2p.as<int32_t>();
    prof_begin(c, "p2p_mixed_rhs");
    for (int k = 0; k < nch; ++k) {
        const float* qc = c->d_qall.as<float>() + npad * ch[k].r0;
        double* res = c->d_res_sorted.as<double>() + static_cast<size_t>(ch[k].r0) * 4 * c->n;
        switch (ch[k].R) {
            case 8: hipLaunchKernelGGL(k_p2p_mixed_rhs<8>, grid, block, 0, c->stream, c->n, pl, ls, sl, loc, qc, inv_u, res); break;
            case 4: hipLaunchKernelGGL(k_p2p_mixed_rhs<4>, grid, block, 0, c->stream, c->n, pl, ls, sl, loc, qc, inv_u, res); break;
            case 2: hipLaunchKernelGGL(k_p2p_mixed_rhs<2>, grid, block, 0, c->stream, c->n, pl, ls, sl, loc, qc, inv_u, res); break;
            default: hipLaunchKernelGGL(k_p2p_mixed_rhs<1>, grid, block, 0, c->stream, c->n, pl, ls, sl, loc, qc, inv_u, res); break;
        }
    }
    prof_end(c);
    return check_launch(c, "p2p_mixed_rhs");
}

}  // namespace hfmm
