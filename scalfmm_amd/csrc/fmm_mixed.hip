// fmm_mixed.hip -- the mixed-precision passes of the Chebyshev 1/r FMM
// (HFMM_PRECISION_MIXED).
//
// M2L: the tile scheme of the FP64 kernel, with every operand split into
// three bfloat16 parts (x = hi + mid + lo, 24 significant bits) and every
// product run on v_mfma_f32_16x16x32_bf16 as the six leading partial
// products (hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid) accumulated in
// FP32.  The node permutations of the symmetric scheme are folded into
// position-specific operators (one per far-field position, built on the
// device from the canonical FP64 ones), so both GEMMs stream their A operand
// contiguously.  The M2L output side works in the Chebyshev modal basis
// (coefficients of T_a(x) T_b(y) T_c(z) instead of values at the nodes): the
// FP32 accumulation and the 24-bit operators then err relative to each mode,
// not relative to the cell's mean potential, so the gradient modes that make
// the forces keep their accuracy.  A small FP64 kernel turns the modal locals
// back into node values.  P2M, M2M, L2L and L2P stay FP64.  The mixed P2P
// lives in p2p_mixed.hip.
//
// Reference operator: Src/Kernels/Chebyshev/FChebSymKernel.hpp:210-328 (M2L).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "fmm_common.hpp"
#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf8*>(p); }

typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
// Elements 0..3 from memory (8 bytes), 4..7 zero.
__device__ __forceinline__ bf8 ld4z(const uint16_t* p) {
    const bf4 v = *reinterpret_cast<const bf4*>(p);
    const bf4 z = {};
    return __builtin_shufflevector(v, z, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ f4 mfma_bf16(bf8 a, bf8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The six leading partial products of (ah + am + al) (bh + bm + bl), smallest first.
__device__ __forceinline__ f4 mma6(bf8 ah, bf8 am, bf8 al, bf8 bh, bf8 bm, bf8 bl, f4 c) {
    c = mfma_bf16(ah, bl, c);
    c = mfma_bf16(al, bh, c);
    c = mfma_bf16(am, bm, c);
    c = mfma_bf16(ah, bm, c);
    c = mfma_bf16(am, bh, c);
    return mfma_bf16(ah, bh, c);
}

// NPL = 3: the six leading partial products of a three-part split; NPL = 2:
// the three of a two-part (hi/lo) split, (ah + am)(bh + bm) without am bm,
// the third parts unused (their loads are dead code the compiler removes).
template <int NPL>
__device__ __forceinline__ f4 mmaN(bf8 ah, bf8 am, bf8 al, bf8 bh, bf8 bm, bf8 bl, f4 c) {
    if constexpr (NPL == 3) return mma6(ah, am, al, bh, bm, bl, c);
    c = mfma_bf16(ah, bm, c);
    c = mfma_bf16(am, bh, c);
    return mfma_bf16(ah, bh, c);
}

// mmaN for COLS accumulators sharing the A operand, the partial products
// interleaved over the accumulators (each accumulator receives its products
// in mmaN's order), so that no MFMA waits for the result of the one before.
template <int NPL, int COLS>
__device__ __forceinline__ void mmaN_cols(const bf8 (&a)[3], const bf8 (&b)[3][COLS], f4 (&c)[COLS]) {
    // (A part, B part) of each partial product, smallest first as in mma6
    constexpr int kA6[6] = {0, 2, 1, 0, 1, 0}, kB6[6] = {2, 0, 1, 1, 0, 0};
    constexpr int kA3[3] = {0, 1, 0}, kB3[3] = {1, 0, 0};
#pragma unroll
    for (int q = 0; q < (NPL == 3 ? 6 : 3); ++q) {
        const int ia = NPL == 3 ? kA6[q] : kA3[q], ib = NPL == 3 ? kB6[q] : kB3[q];
#pragma unroll
        for (int ct = 0; ct < COLS; ++ct) c[ct] = mfma_bf16(a[ia], b[ib][ct], c[ct]);
    }
}

__device__ __forceinline__ uint16_t bf16_bits(__bf16 v) { return __builtin_bit_cast(uint16_t, v); }

// x = hi + mid + lo, each part rounded to nearest bfloat16.
__device__ __forceinline__ void split3(float x, __bf16& hi, __bf16& mid, __bf16& lo) {
    hi = static_cast<__bf16>(x);
    const float r1 = x - static_cast<float>(hi);
    mid = static_cast<__bf16>(r1);
    lo = static_cast<__bf16>(r1 - static_cast<float>(mid));
}

__device__ __forceinline__ void split3d(double x, uint16_t& h, uint16_t& m, uint16_t& l) {
    const __bf16 hi = static_cast<__bf16>(static_cast<float>(x));
    const double r1 = x - static_cast<double>(static_cast<float>(hi));
    const __bf16 mid = static_cast<__bf16>(static_cast<float>(r1));
    const double r2 = r1 - static_cast<double>(static_cast<float>(mid));
    h = bf16_bits(hi);
    m = bf16_bits(mid);
    l = bf16_bits(static_cast<__bf16>(static_cast<float>(r2)));
}

// Position-specific operators from the canonical ones (one workgroup per
// far-field position), planes hi/mid/lo:
//   VT_pos[k][n] = V_c[pv[n]][k]                         as [3][kp16][LD]
//   U_pos[m][k]  = sum_n W3[m][n] U_c[pv[n]][k]          as [3][LD][kp32]
// with W3 = W (x) W (x) W the nodal -> modal transform of one dimension,
// W[a][j] = (a ? 2 : 1) / p T_a(x_j) (discrete orthogonality of T at the
// roots), node / mode index n = n1 + p n2 + p^2 n3.  The padded permutation
// entries point at the zero row nn of U_c / V_c.  With `interleave` the rank
// columns of U_pos are stored in the order the wave-level kernel consumes
// them: in each block of 32, rank 16 h + 4 g + e (h < 2, g < 4, e < 4) at
// column 8 g + 4 h + e.
__global__ __launch_bounds__(256) void k_pack_positions(const int16_t* __restrict__ plist, const int32_t* __restrict__ pkt,
                                                       const int64_t* __restrict__ posoff, const uint32_t* __restrict__ pidx,
                                                       const uint16_t* __restrict__ pvec, const int64_t* __restrict__ opoff,
                                                       const double* __restrict__ ops, const double* __restrict__ Troots,
                                                       int p, int LD, int interleave, uint16_t* __restrict__ opm) {
    __shared__ double sW[16 * 16];
    const int nn = p * p * p;
    for (int i = threadIdx.x; i < p * p; i += blockDim.x) sW[i] = (i >= p ? 2. : 1.) / p * Troots[i];
    __syncthreads();
    const int pos = plist[blockIdx.x];
    const int kt = pkt[blockIdx.x];
    const int kp16 = 16 * kt, kp32 = 32 * ((kt + 1) / 2);
    const double* Uc = ops + opoff[pidx[pos]];
    const double* Vc = Uc + static_cast<int64_t>(nn + 1) * kp16;
    const uint16_t* pv = pvec + static_cast<int64_t>(pos) * LD;
    uint16_t* vt = opm + posoff[pos];
    const int64_t nv = static_cast<int64_t>(kp16) * LD;
    for (int64_t e = threadIdx.x; e < nv; e += blockDim.x) {
        const int k = static_cast<int>(e / LD), n = static_cast<int>(e % LD);
        split3d(Vc[static_cast<int64_t>(pv[n]) * kp16 + k], vt[e], vt[nv + e], vt[2 * nv + e]);
    }
    uint16_t* u = vt + 3 * nv;
    const int64_t nu = static_cast<int64_t>(LD) * kp32;
    for (int64_t e = threadIdx.x; e < nu; e += blockDim.x) {
        const int m = static_cast<int>(e / kp32), k = static_cast<int>(e % kp32);
        double v = 0.;
        if (m < nn && k < kp16) {
            const int m1 = m % p, m2 = (m / p) % p, m3 = m / (p * p);
            for (int n3 = 0; n3 < p; ++n3)
                for (int n2 = 0; n2 < p; ++n2) {
                    const double w23 = sW[m3 * p + n3] * sW[m2 * p + n2];
                    const int nb = (n3 * p + n2) * p;
                    double t = 0.;
                    for (int n1 = 0; n1 < p; ++n1) t += sW[m1 * p + n1] * Uc[static_cast<int64_t>(pv[nb + n1]) * kp16 + k];
                    v += w23 * t;
                }
        }
        const int kk = interleave ? 32 * (k / 32) + 8 * ((k % 16) / 4) + 4 * ((k % 32) / 16) + (k % 4) : k;
        const int64_t d = static_cast<int64_t>(m) * kp32 + kk;
        split3d(v, u[d], u[nu + d], u[2 * nu + d]);
    }
}

// Modal -> nodal locals in place, one wave per cell, FP64:
// f[n] = sum_m T_m1(x_n1) T_m2(x_n2) T_m3(x_n3) a[m], three one-dimensional passes.
template <int P>
__global__ __launch_bounds__(256) void k_modal_to_nodal(int64_t ncells, const double* __restrict__ Troots,
                                                       double* __restrict__ L, int LD) {
    constexpr int NN = P * P * P;
    __shared__ double sT[P * P];
    __shared__ double buf[4][2][NN];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < P * P; i += 256) sT[i] = Troots[i];
    const int64_t cell = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    const bool live = cell < ncells;
    double* Lc = L + (live ? cell : 0) * LD;
    double* a = buf[wave][0];
    double* b = buf[wave][1];
    for (int i = lane; i < NN; i += 64) a[i] = live ? Lc[i] : 0.;
    __syncthreads();
    int stride = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int i = lane; i < NN; i += 64) {
            const int nd = (i / stride) % P;
            const int base = i - nd * stride;
            double s = 0.;
#pragma unroll
            for (int m = 0; m < P; ++m) s += sT[m * P + nd] * a[base + m * stride];
            b[i] = s;
        }
        __syncthreads();
        double* t = a;
        a = b;
        b = t;
        stride *= P;
    }
    if (live)
        for (int i = lane; i < NN; i += 64) Lc[i] = a[i];
}

// Multipoles (FP64, [cells + 1][LD]) -> FP32 of the same shape; the M2L splits
// each value into three bfloat16 parts in registers as it loads it.
__global__ __launch_bounds__(256) void k_split_multipoles(int64_t nvals, const double* __restrict__ M,
                                                         float* __restrict__ Mf) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nvals) return;
    Mf[i] = static_cast<float>(M[i]);
}

// ---------------------------------------------------------------------------
// M2L.  A workgroup of 4 waves owns COLS column tiles of 16 targets of one
// parity at one level and walks the 189 slots of that parity, ordered by the
// rank tiles of their operators.  Per slot:
//   GEMM1  C (kp16 x 16 COLS) = VT_pos * M_src, the node dimension split in
//          four over the waves, partial C tiles through LDS;
//   GEMM2  acc (nn x 16 COLS) += U_pos * C, row tiles spread over the waves,
//          B = the sum of the partial C tiles, split again into three parts.
// MFMA 16x16x32 bf16 layout (checked by hfmm_selftest_mfma_bf16): A[i][k] in
// lane i + 16 (k / 8), element k % 8; B[k][j] in lane j + 16 (k / 8),
// element k % 8; register r of lane l holds D[4 (l / 16) + r][l % 16].
template <int P, int COLS>
struct MixState {
    static constexpr int NN = P * P * P;
    static constexpr int NT = (NN + 15) / 16;
    static constexpr int NTW = (NT + kM2LWaves - 1) / kM2LWaves;
    static constexpr int LDP = (NN + 15) / 16 * 16;  // the padded node count (the context's LD)
    f4 acc[NTW][COLS];
    f4 comp[COLS];  // Kahan compensation of the first row tile (the lowest modes)
};

// Multipole fragments (COLS column tiles) of node block kb: eight FP32 values
// per lane (two 16-byte loads, 16-byte aligned: LD and 8 kq are multiples of
// 8), split into the three bfloat16 parts in registers.
template <int COLS>
__device__ __forceinline__ void fetch_mfrag(f4 (&x)[COLS][2], const float* const (&Mrow)[COLS], int kb) {
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct) {
        const float* p = Mrow[ct] + 32 * kb;
        x[ct][0] = *reinterpret_cast<const f4*>(p);
        x[ct][1] = *reinterpret_cast<const f4*>(p + 4);
    }
}

template <int COLS>
__device__ __forceinline__ void split_mcol(bf8 (&b)[3][COLS], const f4 (&x)[COLS][2], int ct) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        __bf16 h, m, l;
        split3(x[ct][0][e], h, m, l);
        b[0][ct][e] = h;
        b[1][ct][e] = m;
        b[2][ct][e] = l;
        split3(x[ct][1][e], h, m, l);
        b[0][ct][4 + e] = h;
        b[1][ct][4 + e] = m;
        b[2][ct][4 + e] = l;
    }
}

template <int COLS>
__device__ __forceinline__ void split_mfrag(bf8 (&b)[3][COLS], const f4 (&x)[COLS][2]) {
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct) split_mcol<COLS>(b, x, ct);
}

template <int COLS>
__device__ __forceinline__ void load_mfrag(bf8 (&b)[3][COLS], const float* const (&Mrow)[COLS], int kb) {
    f4 x[COLS][2];
    fetch_mfrag<COLS>(x, Mrow, kb);
    split_mfrag<COLS>(b, x);
}

// VT rows of rank tile rt, node block kb, three planes.
template <int KP16>
__device__ __forceinline__ void load_vfrag(bf8 (&a)[3], const uint16_t* VT, int LD, int rt, int col, int kq, int kb) {
    const uint16_t* p = VT + static_cast<int64_t>(16 * rt + col) * LD + 32 * kb + 8 * kq;
    a[0] = ld8(p);
    a[1] = ld8(p + KP16 * LD);
    a[2] = ld8(p + 2 * KP16 * LD);
}

// GEMM1 on node block kb with the VT rows rolling one rank tile ahead: va[0]
// holds rank tile 0 of this block on entry and rank tile 0 of block kb + 1 on
// exit (when `more`).  With `more` and SPLIT, the FP32 multipole fragments of
// block kb + 1 (xn) are split into bn alongside the MFMAs, column tile ct in
// the region of rank tile ct % KT, so the conversions overlap the matrix work.
template <int COLS, int KT, int KP16, int NPL, bool SPLIT = true>
__device__ __forceinline__ void g1_step(f4 (&c1)[KT][COLS], const bf8 (&b)[3][COLS], bf8 (&va)[2][3],
                                        const uint16_t* VT, int LD, int col, int kq, int kb, bool more,
                                        bf8 (&bn)[3][COLS], const f4 (&xn)[COLS][2]) {
#pragma unroll
    for (int rt = 0; rt < KT; ++rt) {
        const int cur = rt & 1;
        if (rt + 1 < KT)
            load_vfrag<KP16>(va[cur ^ 1], VT, LD, rt + 1, col, kq, kb);
        else if (more)
            load_vfrag<KP16>(va[cur ^ 1], VT, LD, 0, col, kq, kb + 1);
        mmaN_cols<NPL, COLS>(va[cur], b, c1[rt]);
        if (SPLIT && more) {
#pragma unroll
            for (int ct = 0; ct < COLS; ++ct)
                if (ct % KT == rt) split_mcol<COLS>(bn, xn, ct);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (KT & 1) {  // the next block's rank tile 0 landed in va[1]
        va[0][0] = va[1][0];
        va[0][1] = va[1][1];
        va[0][2] = va[1][2];
    }
}

// U rows of GEMM2 step s = kb NTW + t: row tile wave + W t (a wave with fewer
// real row tiles computes throw-away ones on row 0), K block kb, the planes
// of the split.  U's rank columns 16 KT .. KP32 are zero padding: the lanes
// that hold only padding get zeros without a load (half the bytes of an odd
// KT's last block).
template <int NTW, int KP16, int KP32, int NPL>
__device__ __forceinline__ void load_urows(bf8 (&a)[3], const uint16_t* U, int LD, int wave, int col, int kq,
                                           int s) {
    const int kb = s / NTW, t = s % NTW;
    const int row = 16 * (wave + kM2LWaves * t) + col;
    const uint16_t* p = U + static_cast<int64_t>(row < LD ? row : 0) * KP32 + 32 * kb + 8 * kq;
    a[0] = bf8{};
    a[1] = bf8{};
    a[2] = bf8{};
    if (32 * kb + 8 * kq < KP16) {
        a[0] = ld8(p);
        a[1] = ld8(p + static_cast<int64_t>(LD) * KP32);
        if constexpr (NPL == 3) a[2] = ld8(p + 2 * static_cast<int64_t>(LD) * KP32);
    }
}

// GEMM2's B fragments of K block kb (the split sums of C), from the LDS.
template <int COLS, int NPL>
__device__ __forceinline__ void read_bfrags(bf8 (&b)[3][COLS], const bf8* Bs, int lane, int kb) {
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct) {
        const bf8* q = Bs + ((ct * 2 + kb) * 3) * 64 + lane;
        b[0][ct] = q[0];
        b[1][ct] = q[64];
        b[2][ct] = NPL == 3 ? q[128] : bf8{};
    }
}

// Runs the slots [i0, i1) of the tile's (rank-ordered) slot list, all with
// KT rank tiles.  Per slot:
//   GEMM1   each wave multiplies its quarter of the node dimension into a
//           partial C tile, stored as [wave][target][row] floats;
//   barrier
//   reduce  the (column tile, K block) pairs are spread over the waves: each
//           sums the four partials of its pair once, splits the sums into
//           three bf16 parts and stores them as B fragments;
//   barrier
//   GEMM2   every wave reads the B fragments (16-byte LDS loads) for its row
//           tiles of U.
// The next slot's GEMM1 may overwrite the partials as soon as the second
// barrier is passed (every wave has finished reducing), and the next
// reduction writes the B fragments only after the next first barrier (every
// wave has finished its GEMM2).
template <int P, int COLS, int KT, int NPL>
__device__ __forceinline__ void mix_segment(MixState<P, COLS>& st, int i0, int i1, const int16_t* s_slot,
                                            const int64_t* s_off, const int32_t* lslots, const int (&tgt)[COLS],
                                            int (&nsrc)[COLS], const float* __restrict__ Mbl,
                                            const float* __restrict__ Mbz,
                                            const uint16_t* __restrict__ opm, float* Cp, bf8* Bs, int LD, int CR,
                                            int wave, int lane, int col, int kq, int kb0, int kb1) {
    constexpr int W = kM2LWaves;
    static_assert(COLS == 2 || COLS == 4, "the reduction spreads (column tile, K block) jobs over 4 waves");
    constexpr int NTW = MixState<P, COLS>::NTW;
    constexpr int TC = 16 * COLS;
    constexpr int KP16 = 16 * KT;       // GEMM1 rows: rank padded to 16
    constexpr int KB32 = (KT + 1) / 2;  // GEMM2 K blocks: rank padded to 32
    constexpr int KP32 = 32 * KB32;
    // node blocks per wave: kb1 - kb0 is NBW for every wave when the block
    // count divides over the waves, else NBW or NBW - 1
    constexpr int KBP = MixState<P, COLS>::LDP / 32;
    constexpr int NBW = (KBP + W - 1) / W;
    constexpr bool EVEN = KBP % W == 0;
    const int nbw = EVEN ? NBW : kb1 - kb0;
    for (int i = i0; i < i1; ++i) {
        const float* Mrow[COLS];
        bool any = false;
        bool has[COLS];  // this lane's target has a source in this slot
#pragma unroll
        for (int ct = 0; ct < COLS; ++ct) {
            const int s = nsrc[ct];
            has[ct] = s >= 0;
            any |= s >= 0;
            // a missing source reads the zero multipole (no divergent branches)
            Mrow[ct] = (s >= 0 ? Mbl + static_cast<int64_t>(s) * LD : Mbz) + 8 * kq;
            nsrc[ct] = (tgt[ct] >= 0 && i + 1 < kFarSlots)
                           ? lslots[static_cast<int64_t>(tgt[ct]) * kFarSlots + s_slot[i + 1]]
                           : -1;
        }
        if (__ballot(any) == 0) continue;  // same answer in every wave
        const uint16_t* VT = opm + s_off[i];
        const uint16_t* U = VT + 3 * static_cast<int64_t>(KP16) * LD;
        {
            f4 c1[KT][COLS];
#pragma unroll
            for (int rt = 0; rt < KT; ++rt)
#pragma unroll
                for (int ct = 0; ct < COLS; ++ct) c1[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
            // Software pipeline over (node block, rank tile), so the L2
            // latency overlaps the matrix work: the VT rows of the next rank
            // tile are requested before the current MFMAs (g1_step); the FP32
            // multipole fragments of block j two blocks ahead (xr, a ring of
            // two), split during block j - 1's MFMAs (b3, a ring of two).  The
            // block loop is unrolled: with a loop-carried c1 the compiler moves
            // the accumulators between AGPRs and VGPRs every iteration.
            f4 xr[2][COLS][2];
            bf8 b3[2][3][COLS];
            bf8 va[2][3];
            if (kb0 < kb1) {
                load_mfrag<COLS>(b3[0], Mrow, kb0);
                if (nbw > 1) fetch_mfrag<COLS>(xr[1], Mrow, kb0 + 1);
                if (nbw > 2) fetch_mfrag<COLS>(xr[0], Mrow, kb0 + 2);
                load_vfrag<KP16>(va[0], VT, LD, 0, col, kq, kb0);
            }
#pragma unroll
            for (int j = 0; j < NBW; ++j) {
                if (EVEN || j < nbw) {
                    const bool more = j + 1 < nbw;
                    g1_step<COLS, KT, KP16, NPL>(c1, b3[j & 1], va, VT, LD, col, kq, kb0 + j, more,
                                                 b3[(j + 1) & 1], xr[(j + 1) & 1]);
                    if (j + 3 < nbw) fetch_mfrag<COLS>(xr[(j + 1) & 1], Mrow, kb0 + j + 3);
                }
            }
            // partial C as [wave][target][row]: lane (col, kq) holds rows
            // 16 rt + 4 kq .. +3 of target 16 ct + col
            float* Cw = Cp + static_cast<size_t>(wave) * TC * CR;
#pragma unroll
            for (int rt = 0; rt < KT; ++rt)
#pragma unroll
                for (int ct = 0; ct < COLS; ++ct)
                    *reinterpret_cast<f4*>(Cw + (16 * ct + col) * CR + 16 * rt + 4 * kq) = c1[rt][ct];
        }
        // GEMM2 runs in steps s = kb NTW + t (K block kb, row tile t).  The U
        // rows of a step are requested two steps ahead into a ring of three,
        // those of steps 0 and 1 before the barriers: with one wave per SIMD
        // only this wave's own MFMAs can cover the L2 latency.
        constexpr int NS = KB32 * NTW;
        bf8 ub[3][3];
        load_urows<NTW, KP16, KP32, NPL>(ub[0], U, LD, wave, col, kq, 0);
        if constexpr (NS > 1) load_urows<NTW, KP16, KP32, NPL>(ub[1], U, LD, wave, col, kq, 1);
        __syncthreads();
        // Rows 16 KT .. KP32 of C are zero: the segments run with growing KT,
        // so no earlier slot wrote them, and the LDS was cleared at the start.
        // reduction jobs (column tile ct, K block kb), job = ct + COLS kb, over the waves
#pragma unroll
        for (int job = wave; job < COLS * KB32; job += W) {
            const int ct = job % COLS, kb = job / COLS;
            const float* q = Cp + (16 * ct + col) * CR + 32 * kb + 8 * kq;
            f4 v0 = *reinterpret_cast<const f4*>(q), v1 = *reinterpret_cast<const f4*>(q + 4);
#pragma unroll
            for (int w = 1; w < W; ++w) {
                v0 += *reinterpret_cast<const f4*>(q + w * TC * CR);
                v1 += *reinterpret_cast<const f4*>(q + w * TC * CR + 4);
            }
            bf8 h8, m8, l8;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                __bf16 h, m, l;
                split3(e < 4 ? v0[e] : v1[e - 4], h, m, l);
                h8[e] = h;
                m8[e] = m;
                l8[e] = l;
            }
            bf8* b = Bs + ((ct * 2 + kb) * 3) * 64 + lane;
            b[0] = h8;
            b[64] = m8;
            if constexpr (NPL == 3) b[128] = l8;
        }
        __syncthreads();
        // The first row tile of every wave (modes 0..63: the mean potential and
        // the in-plane gradients, the largest coefficients) is accumulated per
        // slot from zero and added with compensation; the other tiles
        // accumulate directly.
        f4 fresh[COLS];
#pragma unroll
        for (int ct = 0; ct < COLS; ++ct) fresh[ct] = f4{0.f, 0.f, 0.f, 0.f};
        bf8 b[3][COLS];  // B fragments of the current K block
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int kb = s / NTW, t = s % NTW;
            if (s + 2 < NS) load_urows<NTW, KP16, KP32, NPL>(ub[(s + 2) % 3], U, LD, wave, col, kq, s + 2);
            if (t == 0) read_bfrags<COLS, NPL>(b, Bs, lane, kb);
            if (t == 0)
                mmaN_cols<NPL, COLS>(ub[s % 3], b, fresh);
            else
                mmaN_cols<NPL, COLS>(ub[s % 3], b, st.acc[t]);
            __builtin_amdgcn_sched_barrier(0);
        }
        // (a target without a source here keeps its sum and compensation as
        // they are, whatever the other targets of the tile do)
#pragma unroll
        for (int ct = 0; ct < COLS; ++ct) {
            const f4 y = fresh[ct] - st.comp[ct];
            const f4 tt = st.acc[0][ct] + y;
            const f4 cn = (tt - st.acc[0][ct]) - y;
            if (has[ct]) {
                st.comp[ct] = cn;
                st.acc[0][ct] = tt;
            }
        }
    }
}

template <int P, int COLS, int NPL>
__global__ __launch_bounds__(256) void k_m2l_mixed(M2LLevels lv, const int32_t* __restrict__ slots,
                                                  const int32_t* __restrict__ order, const int32_t* __restrict__ pcount,
                                                  const int16_t* __restrict__ slotord, const int32_t* __restrict__ slotseg,
                                                  const int64_t* __restrict__ posoff, const uint16_t* __restrict__ opm,
                                                  const float* __restrict__ Mb, int64_t zero_cell,
                                                  double* __restrict__ L, int LD, int CR) {
    constexpr int NN = P * P * P;
    constexpr int W = kM2LWaves;
    constexpr int NT = MixState<P, COLS>::NT;
    constexpr int NTW = MixState<P, COLS>::NTW;
    constexpr int TC = 16 * COLS;
    __shared__ int16_t s_slot[kFarSlots];
    __shared__ int64_t s_off[kFarSlots];
    // partial C tiles [W][TC][CR] floats (CR = rows padded by 4 against bank
    // conflicts), then the B fragments [COLS][2][3][64 lanes] x 16 bytes
    extern __shared__ float fsm[];
    bf8* Bs = reinterpret_cast<bf8*>(fsm + W * TC * CR);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kq = lane >> 4;
    const int KB = LD / 32;
    const int kb0 = (KB * wave) / W, kb1 = (KB * (wave + 1)) / W;
    const int ntw = (NT - wave + W - 1) / W;

    const int bid = blockIdx.x;
    if (bid >= lv.tile_begin[lv.nlev]) return;
    int li = 0;
    while (bid >= lv.tile_begin[li + 1]) ++li;
    const int lb = bid - lv.tile_begin[li];
    const int e = lb % 8;  // parity: consecutive workgroups land on different XCDs
    const int ptile = lb / 8;
    const int32_t* pc = pcount + lv.pcount_off[li];
    int first = 0;
    for (int q = 0; q < e; ++q) first += pc[q];
    const int cnt = min(TC, pc[e] - ptile * TC);
    if (cnt <= 0) return;
    first += ptile * TC;
    for (int i = tid; i < kFarSlots; i += 64 * W) {
        s_slot[i] = slotord[(e * kFarSlots + i) * 2];
        s_off[i] = posoff[slotord[(e * kFarSlots + i) * 2 + 1]];
    }
    for (int i = tid; i < W * TC * CR; i += 64 * W) fsm[i] = 0.f;
    int seg[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) seg[q] = slotseg[e * 5 + q];
    const int64_t coff = lv.cell_off[li];
    const int32_t* lslots = slots + coff * kFarSlots;
    const float* Mbl = Mb + coff * LD;
    const float* Mbz = Mb + zero_cell * LD;
    double* Ll = L + coff * LD;
    int tgt[COLS];
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct) {
        const int cc = 16 * ct + col;
        tgt[ct] = cc < cnt ? order[coff + first + cc] : -1;
    }
    __syncthreads();
    int nsrc[COLS];
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct)
        nsrc[ct] = tgt[ct] >= 0 ? lslots[static_cast<int64_t>(tgt[ct]) * kFarSlots + s_slot[0]] : -1;
    MixState<P, COLS> st;
#pragma unroll
    for (int t = 0; t < NTW; ++t)
#pragma unroll
        for (int ct = 0; ct < COLS; ++ct) st.acc[t][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct) st.comp[ct] = f4{0.f, 0.f, 0.f, 0.f};
    mix_segment<P, COLS, 1, NPL>(st, seg[0], seg[1], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, fsm, Bs,
                            LD, CR, wave, lane, col, kq, kb0, kb1);
    mix_segment<P, COLS, 2, NPL>(st, seg[1], seg[2], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, fsm, Bs,
                            LD, CR, wave, lane, col, kq, kb0, kb1);
    mix_segment<P, COLS, 3, NPL>(st, seg[2], seg[3], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, fsm, Bs,
                            LD, CR, wave, lane, col, kq, kb0, kb1);
    mix_segment<P, COLS, 4, NPL>(st, seg[3], seg[4], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, fsm, Bs,
                            LD, CR, wave, lane, col, kq, kb0, kb1);
    const double scale = lv.scale[li];
#pragma unroll
    for (int ct = 0; ct < COLS; ++ct) {
        if (tgt[ct] < 0) continue;
#pragma unroll
        for (int t = 0; t < NTW; ++t) {
            if (t < ntw) {  // real row tiles of this wave
                const int nt = wave + W * t;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * nt + 4 * kq + r;
                    if (row < NN)
                        Ll[static_cast<int64_t>(tgt[ct]) * LD + row] =
                            scale * (t == 0 ? static_cast<double>(st.acc[0][ct][r]) - static_cast<double>(st.comp[ct][r])
                                            : static_cast<double>(st.acc[t][ct][r]));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// M2L for small orders (p <= 6): every wave owns CT column tiles of 16
// targets and runs both GEMMs of each slot by itself, with no LDS traffic and
// no barrier.  GEMM1 (C = VT_pos M_src) covers the whole node range, and its
// accumulator registers are GEMM2's B operand as they stand: a lane holds
// rows 4 g .. 4 g + 3 of every rank tile of its target column, and U_pos is
// stored with its rank columns interleaved (k_pack_positions, `interleave`)
// so that K elements 8 g .. 8 g + 7 of a 32-wide block are rows 4 g .. +3 of
// rank tiles 2 s and 2 s + 1.  The first four row tiles of the modal output
// (modes 0..63) are accumulated per slot from zero and added with
// compensation, as in the four-wave kernel.
template <int P, int CT>
struct WaveMixState {
    static constexpr int NN = P * P * P;
    static constexpr int LDC = (NN + 15) / 16 * 16;
    static constexpr int NT = LDC / 16;          // row tiles of the output
    static constexpr int NC = NT < 4 ? NT : 4;   // compensated row tiles
    static constexpr int KB = LDC / 32;          // 32-node blocks of GEMM1
    f4 acc[NT][CT];
    f4 comp[NC][CT];
};

// U rows of the wave-level GEMM2 step s = t KB32 + sb (row tile t, K block
// sb).  With an odd KT the last block's rank tile 2 sb + 1 is zero padding,
// stored (interleaved) in elements 4..7 of every lane: only the first half is
// loaded.
template <int KT, int KP32, int KB32, int NPL>
__device__ __forceinline__ void load_urows_wave(bf8 (&a)[3], const uint16_t* U, int LD, int col, int kq, int s) {
    const int t = s / KB32, sb = s % KB32;
    const uint16_t* p = U + static_cast<int64_t>(16 * t + col) * KP32 + 32 * sb + 8 * kq;
    constexpr bool HALF = (KT & 1) != 0;
    a[2] = bf8{};
    if (HALF && sb == KB32 - 1) {
        a[0] = ld4z(p);
        a[1] = ld4z(p + static_cast<int64_t>(LD) * KP32);
        if constexpr (NPL == 3) a[2] = ld4z(p + 2 * static_cast<int64_t>(LD) * KP32);
    } else {
        a[0] = ld8(p);
        a[1] = ld8(p + static_cast<int64_t>(LD) * KP32);
        if constexpr (NPL == 3) a[2] = ld8(p + 2 * static_cast<int64_t>(LD) * KP32);
    }
}

template <int P, int CT, int KT, int NPL>
__device__ __forceinline__ void wave_segment(WaveMixState<P, CT>& st, int i0, int i1, const int16_t* s_slot,
                                             const int64_t* s_off, const int32_t* lslots, const int (&tgt)[CT],
                                             int (&nsrc)[CT], const float* __restrict__ Mbl,
                                             const float* __restrict__ Mbz,
                                             const uint16_t* __restrict__ opm, int col, int kq) {
    constexpr int NT = WaveMixState<P, CT>::NT;
    constexpr int NC = WaveMixState<P, CT>::NC;
    constexpr int KB = WaveMixState<P, CT>::KB;
    constexpr int LD = WaveMixState<P, CT>::LDC;
    constexpr int KP16 = 16 * KT;       // GEMM1 rows: rank padded to 16
    constexpr int KB32 = (KT + 1) / 2;  // GEMM2 K blocks: rank padded to 32
    constexpr int KP32 = 32 * KB32;
    for (int i = i0; i < i1; ++i) {
        const float* Mrow[CT];
        bool any = false;
        bool has[CT];  // this lane's target has a source in this slot
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int s = nsrc[ct];
            has[ct] = s >= 0;
            any |= s >= 0;
            Mrow[ct] = (s >= 0 ? Mbl + static_cast<int64_t>(s) * LD : Mbz) + 8 * kq;
            nsrc[ct] = (tgt[ct] >= 0 && i + 1 < kFarSlots)
                           ? lslots[static_cast<int64_t>(tgt[ct]) * kFarSlots + s_slot[i + 1]]
                           : -1;
        }
        if (__ballot(any) == 0) continue;  // uniform over the wave
        const uint16_t* VT = opm + s_off[i];
        const uint16_t* U = VT + 3 * static_cast<int64_t>(KP16) * LD;
        f4 c1[KT][CT];
#pragma unroll
        for (int rt = 0; rt < KT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) c1[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
        // GEMM1: the VT rows are requested one rank tile ahead (va).  The
        // FP32 multipole fragments of block kb are requested two blocks ahead
        // (xr) and split during block kb - 1's MFMAs (b3); with two tiles of
        // three parts, whose registers are the tightest, one block ahead and
        // split after block kb - 1's MFMAs.
        constexpr bool X2 = !(CT == 2 && NPL == 3);
        f4 xr[2][CT][2];
        bf8 b3[2][3][CT];
        bf8 va[2][3];
        load_mfrag<CT>(b3[0], Mrow, 0);
        if constexpr (X2 && KB > 1) fetch_mfrag<CT>(xr[1], Mrow, 1);
        if constexpr (X2 && KB > 2) fetch_mfrag<CT>(xr[0], Mrow, 2);
        load_vfrag<KP16>(va[0], VT, LD, 0, col, kq, 0);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            if constexpr (X2) {
                g1_step<CT, KT, KP16, NPL>(c1, b3[kb & 1], va, VT, LD, col, kq, kb, kb + 1 < KB, b3[(kb + 1) & 1],
                                           xr[(kb + 1) & 1]);
                if (kb + 3 < KB) fetch_mfrag<CT>(xr[(kb + 1) & 1], Mrow, kb + 3);
            } else {
                if (kb + 1 < KB) fetch_mfrag<CT>(xr[0], Mrow, kb + 1);
                g1_step<CT, KT, KP16, NPL, false>(c1, b3[kb & 1], va, VT, LD, col, kq, kb, kb + 1 < KB,
                                                  b3[(kb + 1) & 1], xr[0]);
                if (kb + 1 < KB) split_mfrag<CT>(b3[(kb + 1) & 1], xr[0]);
            }
        }
        // GEMM2's B fragments straight from the accumulators, split in three
        bf8 bb[KB32][3][CT];
#pragma unroll
        for (int sb = 0; sb < KB32; ++sb)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int rt = 2 * sb + (e >> 2);
                    const float v = rt < KT ? c1[rt < KT ? rt : 0][ct][e & 3] : 0.f;
                    __bf16 h, m, l;
                    split3(v, h, m, l);
                    bb[sb][0][ct][e] = h;
                    bb[sb][1][ct][e] = m;
                    bb[sb][2][ct][e] = l;
                }
        f4 fresh[NC][CT];
#pragma unroll
        for (int t = 0; t < NC; ++t)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) fresh[t][ct] = f4{0.f, 0.f, 0.f, 0.f};
        // GEMM2 in steps s = t KB32 + sb, the U rows requested UD steps ahead
        constexpr int NS = NT * KB32;
        constexpr int UD = (CT == 2 && NPL == 3) ? 1 : 2;  // one for the tightest registers
        bf8 ub[UD + 1][3];
#pragma unroll
        for (int s = 0; s < UD && s < NS; ++s) load_urows_wave<KT, KP32, KB32, NPL>(ub[s], U, LD, col, kq, s);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int t = s / KB32, sb = s % KB32;
            if (s + UD < NS)
                load_urows_wave<KT, KP32, KB32, NPL>(ub[(s + UD) % (UD + 1)], U, LD, col, kq, s + UD);
            if (t < NC)
                mmaN_cols<NPL, CT>(ub[s % (UD + 1)], bb[sb], fresh[t < NC ? t : 0]);
            else
                mmaN_cols<NPL, CT>(ub[s % (UD + 1)], bb[sb], st.acc[t]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int t = 0; t < NC; ++t)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const f4 y = fresh[t][ct] - st.comp[t][ct];
                const f4 tt = st.acc[t][ct] + y;
                const f4 cn = (tt - st.acc[t][ct]) - y;
                if (has[ct]) {  // targets without a source here stay as they are
                    st.comp[t][ct] = cn;
                    st.acc[t][ct] = tt;
                }
            }
    }
}

// Two waves per SIMD at most 256 registers each: the three-part split with two
// tiles per wave needs 314 and would run one wave per SIMD; held to 256 it
// spills 36 bytes per lane and runs its order-5 trees 1.36x faster (the second
// wave hides the gathers of the first).
template <int P, int CT, int NPL>
__global__ __launch_bounds__(256, 2) void k_m2l_mixed_wave(M2LLevels lv, const int32_t* __restrict__ slots,
                                                       const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ pcount,
                                                       const int16_t* __restrict__ slotord,
                                                       const int32_t* __restrict__ slotseg,
                                                       const int64_t* __restrict__ posoff,
                                                       const uint16_t* __restrict__ opm, const float* __restrict__ Mb,
                                                       int64_t zero_cell, double* __restrict__ L) {
    constexpr int NN = P * P * P;
    constexpr int LD = WaveMixState<P, CT>::LDC;
    constexpr int NT = WaveMixState<P, CT>::NT;
    constexpr int NC = WaveMixState<P, CT>::NC;
    constexpr int W = kM2LWaves;
    constexpr int TCW = 16 * CT;  // targets per wave
    constexpr int TC = W * TCW;   // targets per workgroup
    __shared__ int16_t s_slot[kFarSlots];
    __shared__ int64_t s_off[kFarSlots];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, kq = lane >> 4;

    const int bid = blockIdx.x;
    if (bid >= lv.tile_begin[lv.nlev]) return;
    int li = 0;
    while (bid >= lv.tile_begin[li + 1]) ++li;
    const int lb = bid - lv.tile_begin[li];
    const int e = lb % 8;  // parity: consecutive workgroups land on different XCDs
    const int ptile = lb / 8;
    const int32_t* pc = pcount + lv.pcount_off[li];
    int first = 0;
    for (int q = 0; q < e; ++q) first += pc[q];
    const int cnt = min(TC, pc[e] - ptile * TC);
    if (cnt <= 0) return;
    first += ptile * TC;
    for (int i = tid; i < kFarSlots; i += 64 * W) {
        s_slot[i] = slotord[(e * kFarSlots + i) * 2];
        s_off[i] = posoff[slotord[(e * kFarSlots + i) * 2 + 1]];
    }
    int seg[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) seg[q] = slotseg[e * 5 + q];
    const int64_t coff = lv.cell_off[li];
    const int32_t* lslots = slots + coff * kFarSlots;
    const float* Mbl = Mb + coff * LD;
    const float* Mbz = Mb + zero_cell * LD;
    double* Ll = L + coff * LD;
    int tgt[CT];
    bool mine = false;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int cc = TCW * wave + 16 * ct + col;
        tgt[ct] = cc < cnt ? order[coff + first + cc] : -1;
        mine |= tgt[ct] >= 0;
    }
    __syncthreads();  // the slot tables; no barrier follows
    if (__ballot(mine) == 0) return;  // uniform over the wave
    int nsrc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
        nsrc[ct] = tgt[ct] >= 0 ? lslots[static_cast<int64_t>(tgt[ct]) * kFarSlots + s_slot[0]] : -1;
    WaveMixState<P, CT> st;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) st.acc[t][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < NC; ++t)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) st.comp[t][ct] = f4{0.f, 0.f, 0.f, 0.f};
    wave_segment<P, CT, 1, NPL>(st, seg[0], seg[1], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, col, kq);
    wave_segment<P, CT, 2, NPL>(st, seg[1], seg[2], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, col, kq);
    wave_segment<P, CT, 3, NPL>(st, seg[2], seg[3], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, col, kq);
    wave_segment<P, CT, 4, NPL>(st, seg[3], seg[4], s_slot, s_off, lslots, tgt, nsrc, Mbl, Mbz, opm, col, kq);
    const double scale = lv.scale[li];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        if (tgt[ct] < 0) continue;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * t + 4 * kq + r;
                if (row < NN)
                    Ll[static_cast<int64_t>(tgt[ct]) * LD + row] =
                        scale * (t < NC ? static_cast<double>(st.acc[t][ct][r]) - static_cast<double>(st.comp[t][ct][r])
                                        : static_cast<double>(st.acc[t][ct][r]));
            }
    }
}

// A[i][k] = (7 i + 3 k) mod 11 - 5, B[k][j] = (5 k + 2 j) mod 13 - 6 (exact in
// bfloat16, exact products and sums in FP32); D = A B written row-major
// through the layout the M2L kernel assumes.
__global__ void k_mfma_probe_bf16(float* out) {
    const int l = threadIdx.x;
    bf8 a, b;
    for (int j = 0; j < 8; ++j) {
        const int k = 8 * (l / 16) + j;
        a[j] = static_cast<__bf16>(static_cast<float>((7 * (l % 16) + 3 * k) % 11 - 5));
        b[j] = static_cast<__bf16>(static_cast<float>((5 * k + 2 * (l % 16)) % 13 - 6));
    }
    f4 c = {0.f, 0.f, 0.f, 0.f};
    c = mfma_bf16(a, b, c);
    for (int r = 0; r < 4; ++r) out[(4 * (l / 16) + r) * 16 + (l % 16)] = c[r];
}

inline unsigned blocks_of(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

template <int P, int COLS, int NPL>
int m2l_mixed_launch(Context* c) {
    if (c->LD != MixState<P, COLS>::LDP) return fail(c, HFMM_ERR_ARG, "mixed M2L: padded node count");
    int tiles = 0;
    const M2LLevels lv = make_levels(c, 16 * COLS, tiles);
    if (tiles == 0) return HFMM_OK;  // no M2L level in the working range
    const int CR = ((c->kpad_max + 31) & ~31) + 4;
    const size_t shmem = sizeof(float) * kM2LWaves * static_cast<size_t>(CR) * 16 * COLS + 16 * COLS * 2 * 3 * 64;
    static int lds_limit[64] = {0};
    HFMM_TRY(raise_lds_limit(c, reinterpret_cast<const void*>(&k_m2l_mixed<P, COLS, NPL>), shmem, lds_limit));
    hipLaunchKernelGGL((k_m2l_mixed<P, COLS, NPL>), dim3(static_cast<unsigned>(tiles)), dim3(64 * kM2LWaves), shmem,
                       c->stream, lv, c->d_m2l.as<int32_t>(), c->d_order.as<int32_t>(), c->d_pcount.as<int32_t>(),
                       c->d_slotord.as<int16_t>(), c->d_slotseg.as<int32_t>(), c->d_posoff.as<int64_t>(),
                       c->d_opm.as<uint16_t>(), c->d_Mb.as<float>(), c->total_cells, c->d_L.as<double>(),
                       c->LD, CR);
    return check_launch(c, "m2l_mixed");
}

template <int P, int CT, int NPL>
int m2l_mixed_wave_launch(Context* c) {
    if (c->LD != WaveMixState<P, CT>::LDC) return fail(c, HFMM_ERR_ARG, "wave-level M2L: padded node count");
    int tiles = 0;
    const M2LLevels lv = make_levels(c, 16 * CT * kM2LWaves, tiles);
    if (tiles == 0) return HFMM_OK;  // no M2L level in the working range
    hipLaunchKernelGGL((k_m2l_mixed_wave<P, CT, NPL>), dim3(static_cast<unsigned>(tiles)), dim3(64 * kM2LWaves), 0,
                       c->stream, lv, c->d_m2l.as<int32_t>(), c->d_order.as<int32_t>(), c->d_pcount.as<int32_t>(),
                       c->d_slotord.as<int16_t>(), c->d_slotseg.as<int32_t>(), c->d_posoff.as<int64_t>(),
                       c->d_opm2.as<uint16_t>(), c->d_Mb.as<float>(), c->total_cells, c->d_L.as<double>());
    return check_launch(c, "m2l_mixed");
}

// Mixed M2L variants (the tile shapes the autotuner picks from):
//   0  four waves share 64 targets, GEMM1 split over the waves (every order)
//   1  one 16-target tile per wave, waves independent (p <= 6)
//   2  two 16-target tiles per wave (p = 5)
//   3  four waves share 32 targets (every order): half the registers and LDS
//      of shape 0 for twice its operator reads per target
template <int P, int NPL>
int m2l_mixed_variant(Context* c, int v) {
    switch (v) {
        case 0: return m2l_mixed_launch<P, 4, NPL>(c);
        case 3: return m2l_mixed_launch<P, 2, NPL>(c);
        case 1:
            if constexpr (P <= kWaveM2LMaxOrder) return m2l_mixed_wave_launch<P, 1, NPL>(c);
            break;
        case 2:
            if constexpr (P <= 5) return m2l_mixed_wave_launch<P, 2, NPL>(c);
            break;
        default: break;
    }
    return fail(c, HFMM_ERR_ARG, "mixed M2L variant not available at this order");
}

constexpr int kMixedVariants = 4;
inline bool mixed_variant_ok(int p, int v) {
    return v == 0 || v == 3 || (v == 1 && p <= kWaveM2LMaxOrder) || (v == 2 && p <= 5);
}

// Times every available variant once on the current tree (each writes every
// local of the M2L levels, so the runs are interchangeable) and keeps the
// fastest; the choice holds for the context's later evaluations.
template <int P, int NPL>
int m2l_mixed_autotune(Context* c) {
    hipEvent_t ev[2 * kMixedVariants];
    for (auto& e : ev) HFMM_HIP(c, hipEventCreate(&e));
    float best = 0.f;
    int pick = 0;
    int st = HFMM_OK;
    for (int v = 0; v < kMixedVariants && st == HFMM_OK; ++v) {
        if (!mixed_variant_ok(P, v)) continue;
        (void)hipEventRecord(ev[2 * v], c->stream);
        st = m2l_mixed_variant<P, NPL>(c, v);
        (void)hipEventRecord(ev[2 * v + 1], c->stream);
        if (st != HFMM_OK) break;
        if (hipEventSynchronize(ev[2 * v + 1]) != hipSuccess) {
            st = fail(c, HFMM_ERR_DEVICE, "autotune: event wait failed");
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[2 * v], ev[2 * v + 1]);
        c->m2l_tune_ms[v] = ms;
        if (v == 0 || ms < best) {
            best = ms;
            pick = v;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (st != HFMM_OK) return st;
    c->m2l_variant = pick;
    return HFMM_OK;
}

template <int P, int NPL>
int m2l_mixed_impl(Context* c) {
    if (c->m2l_variant < 0) {
        HFMM_TRY((m2l_mixed_autotune<P, NPL>(c)));
    } else if (!mixed_variant_ok(P, c->m2l_variant)) {
        return fail(c, HFMM_ERR_ARG, "mixed M2L variant not available at this order");
    }
    // (the autotuned run already wrote the locals; the chosen variant runs again
    // so that the profile and the results come from one launch of it)
    prof_begin(c, "m2l_mixed");
    HFMM_TRY((m2l_mixed_variant<P, NPL>(c, c->m2l_variant)));
    prof_end(c);
    // the M2L targets: every cell of levels top..h-1, contiguous in the cell arrays
    const int64_t first = c->lvl_off[c->top()], ncells = c->lvl_off[c->height()] - first;
    prof_begin(c, "modal_to_nodal");
    hipLaunchKernelGGL(k_modal_to_nodal<P>, dim3(static_cast<unsigned>((ncells + 3) / 4)), dim3(256), 0, c->stream,
                       ncells, c->d_Troots.as<double>(), c->d_L.as<double>() + first * c->LD, c->LD);
    prof_end(c);
    return check_launch(c, "modal_to_nodal");
}

}  // namespace

int prepare_mixed_operators(Context* c) {
    const ChebOperators& op = c->ops;
    const int LD = c->LD;
    if (LD % 32) return fail(c, HFMM_ERR_ARG, "mixed precision needs 32 | padded node count");
    bool far[343] = {false};
    for (int e = 0; e < 8; ++e) {
        FarSlot s[kFarSlots];
        far_slots(e, s);
        for (int j = 0; j < kFarSlots; ++j) far[s[j].pos] = true;
    }
    std::vector<int16_t> plist;
    std::vector<int32_t> pkt;
    std::vector<int64_t> posoff(343, 0);
    int64_t total = 0;
    for (int pos = 0; pos < 343; ++pos) {
        if (!far[pos]) continue;
        const int k = op.rank[op.pidx[pos]];
        const int kt = (k + 15) / 16;
        if (kt < 1 || kt > 4) return fail(c, HFMM_ERR_ARG, "operator rank outside 1..64");
        plist.push_back(static_cast<int16_t>(pos));
        pkt.push_back(kt);
        posoff[pos] = total;
        total += 3 * static_cast<int64_t>(16 * kt) * LD + 3 * static_cast<int64_t>(LD) * 32 * ((kt + 1) / 2);
    }
    DevBuf d_plist, d_pkt;
    auto run = [&]() -> int {
        HFMM_TRY(ensure(c, d_plist, sizeof(int16_t) * plist.size()));
        HFMM_TRY(ensure(c, d_pkt, sizeof(int32_t) * pkt.size()));
        HFMM_TRY(ensure(c, c->d_posoff, sizeof(int64_t) * posoff.size()));
        HFMM_TRY(ensure(c, c->d_opm, sizeof(uint16_t) * static_cast<size_t>(total)));
        if (op.p <= kWaveM2LMaxOrder) HFMM_TRY(ensure(c, c->d_opm2, sizeof(uint16_t) * static_cast<size_t>(total)));
        HFMM_HIP(c, hipMemcpy(d_plist.ptr, plist.data(), sizeof(int16_t) * plist.size(), hipMemcpyHostToDevice));
        HFMM_HIP(c, hipMemcpy(d_pkt.ptr, pkt.data(), sizeof(int32_t) * pkt.size(), hipMemcpyHostToDevice));
        HFMM_HIP(c, hipMemcpy(c->d_posoff.ptr, posoff.data(), sizeof(int64_t) * posoff.size(), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_pack_positions, dim3(static_cast<unsigned>(plist.size())), dim3(256), 0, c->stream,
                           d_plist.as<int16_t>(), d_pkt.as<int32_t>(), c->d_posoff.as<int64_t>(),
                           c->d_pidx.as<uint32_t>(), c->d_pvec.as<uint16_t>(), c->d_opoff.as<int64_t>(),
                           c->d_ops.as<double>(), c->d_Troots.as<double>(), op.p, LD, 0, c->d_opm.as<uint16_t>());
        HFMM_TRY(check_launch(c, "pack_positions"));
        if (op.p <= kWaveM2LMaxOrder) {  // the same operators for the wave-level kernel
            hipLaunchKernelGGL(k_pack_positions, dim3(static_cast<unsigned>(plist.size())), dim3(256), 0, c->stream,
                               d_plist.as<int16_t>(), d_pkt.as<int32_t>(), c->d_posoff.as<int64_t>(),
                               c->d_pidx.as<uint32_t>(), c->d_pvec.as<uint16_t>(), c->d_opoff.as<int64_t>(),
                               c->d_ops.as<double>(), c->d_Troots.as<double>(), op.p, LD, 1,
                               c->d_opm2.as<uint16_t>());
            HFMM_TRY(check_launch(c, "pack_positions"));
        }
        HFMM_HIP(c, hipStreamSynchronize(c->stream));
        return HFMM_OK;
    };
    const int st = run();
    release(d_plist);
    release(d_pkt);
    return st;
}

int run_split_multipoles(Context* c) {
    const int64_t nvals = (c->total_cells + 1) * static_cast<int64_t>(c->LD);
    HFMM_TRY(ensure(c, c->d_Mb, sizeof(float) * static_cast<size_t>(nvals)));
    prof_begin(c, "split_multipoles");
    hipLaunchKernelGGL(k_split_multipoles, dim3(blocks_of(nvals)), dim3(256), 0, c->stream, nvals,
                       c->d_M.as<double>(), c->d_Mb.as<float>());
    prof_end(c);
    return check_launch(c, "split_multipoles");
}

int run_m2l_mixed(Context* c) {
    if (c->kpad_max > 64) return fail(c, HFMM_ERR_ARG, "operator rank above 64 is not supported");
    // three-part splits (six products) or, in HFMM_PRECISION_MIXED_HILO, two-part
    // hi/lo splits (three products)
    const bool hilo = c->P.precision == HFMM_PRECISION_MIXED_HILO;
    switch (c->order()) {
        case 5: return hilo ? m2l_mixed_impl<5, 2>(c) : m2l_mixed_impl<5, 3>(c);
        case 6: return hilo ? m2l_mixed_impl<6, 2>(c) : m2l_mixed_impl<6, 3>(c);
        case 7: return hilo ? m2l_mixed_impl<7, 2>(c) : m2l_mixed_impl<7, 3>(c);
        case 8: return hilo ? m2l_mixed_impl<8, 2>(c) : m2l_mixed_impl<8, 3>(c);
        default: return fail(c, HFMM_ERR_ARG, "order not compiled (5..8)");
    }
}

int m2l_variant_count(Context* c) { return is_mixed(c) ? kMixedVariants : 1; }
bool m2l_variant_available(Context* c, int v) {
    if (!is_mixed(c)) return v == 0;
    return v >= 0 && v < kMixedVariants && mixed_variant_ok(c->order(), v);
}

int selftest_mfma_bf16(Context* c, float* out) {
    DevBuf b;
    int st = ensure(c, b, sizeof(float) * 256);
    if (st == HFMM_OK) {
        hipLaunchKernelGGL(k_mfma_probe_bf16, dim3(1), dim3(64), 0, c->stream, b.as<float>());
        st = check_launch(c, "mfma_probe_bf16");
    }
    if (st == HFMM_OK &&
        (hipMemcpyAsync(out, b.ptr, sizeof(float) * 256, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
        st = fail(c, HFMM_ERR_DEVICE, "probe copy failed");
    release(b);
    return st;
}

}  // namespace hfmm
