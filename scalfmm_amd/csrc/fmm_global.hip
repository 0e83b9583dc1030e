// fmm_global.hip -- the passes of the reference's FChebKernel that differ from
// the symmetric kernel's (Src/Kernels/Chebyshev/FChebKernel.hpp:93-200), FP64:
//   compression    M~ = B^T M after P2M and after each M2M      (applyB, overwrite)
//   M2L            L~ += 2/w C_t M~_src over the interaction list (applyC)
//   decompression  L += U L~ before each L2L and before L2P      (applyU, add)
// The compressed expansions have k entries per cell (the global rank).  One
// 64-lane wave per cell; lane r owns rank rows r, r + 64, ... (R of them).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fmm_common.hpp"

namespace hfmm {
namespace {

constexpr int kMaxRankRows = 8;  // ranks up to 512 (the dense scheme's p^3 at order 8)

// M~[cell][a] = sum_n B[n][a] M[cell][n]; Bt is B as [nn][k].
template <int R>
__global__ __launch_bounds__(64) void k_global_compress(int64_t ncell, const double* __restrict__ M, int LD,
                                                        const double* __restrict__ Bt, int nn, int k,
                                                        double* __restrict__ Mt) {
    const int64_t cell = blockIdx.x;
    if (cell >= ncell) return;
    const int lane = threadIdx.x;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.;
    const double* m = M + cell * LD;
    for (int n = 0; n < nn; ++n) {
        const double v = m[n];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int a = lane + 64 * r;
            if (a < k) acc[r] += Bt[static_cast<int64_t>(n) * k + a] * v;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int a = lane + 64 * r;
        if (a < k) Mt[cell * k + a] = acc[r];
    }
}

// L[cell][n] += sum_a U[a][n] L~[cell][a]; U as [k][nn].
__global__ __launch_bounds__(64) void k_global_decompress(int64_t ncell, const double* __restrict__ Lt, int k,
                                                          const double* __restrict__ U, int nn,
                                                          double* __restrict__ L, int LD) {
    const int64_t cell = blockIdx.x;
    if (cell >= ncell) return;
    const double* lt = Lt + cell * k;
    for (int n = threadIdx.x; n < nn; n += 64) {
        double s = 0.;
        for (int a = 0; a < k; ++a) s += U[static_cast<int64_t>(a) * nn + n] * lt[a];
        L[cell * LD + n] += s;
    }
}

// L~[cell] += scale C_pos M~[src] for every source of the cell's 189 slots, in
// slot order; slot s of a cell of parity e sits at relative position
// slotpos[e][s].  C_pos is column-major k x k.
template <int R>
__global__ __launch_bounds__(64) void k_global_m2l(int64_t ncell, const int32_t* __restrict__ slots,
                                                   const uint64_t* __restrict__ keys,
                                                   const int16_t* __restrict__ slotpos,
                                                   const double* __restrict__ C, const double* __restrict__ Mt,
                                                   int k, double scale, double* __restrict__ Lt) {
    const int64_t cell = blockIdx.x;
    if (cell >= ncell) return;
    const int lane = threadIdx.x;
    const int par = static_cast<int>(keys[cell] & 7u);
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.;
    for (int s = 0; s < kFarSlots; ++s) {
        const int src = slots[cell * kFarSlots + s];
        if (src < 0) continue;  // the same for every lane
        const double* Cp = C + static_cast<int64_t>(slotpos[par * kFarSlots + s]) * k * k;
        const double* m = Mt + static_cast<int64_t>(src) * k;
        double t[R];
#pragma unroll
        for (int r = 0; r < R; ++r) t[r] = 0.;
        for (int j = 0; j < k; ++j) {
            const double mj = m[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + 64 * r;
                if (i < k) t[r] += Cp[static_cast<int64_t>(j) * k + i] * mj;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += scale * t[r];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < k) Lt[cell * k + i] += acc[r];
    }
}

// ---------------------------------------------------------------------------
// The same M2L on FP64 MFMA (v_mfma_f64_16x16x4f64), for ranks up to
// 16 kGlobalMfmaMaxTiles.  The operators are used as a padded row-major copy
// Cr[pos][i][j] (KP = 16 NR rows and columns, zeros outside k x k).  A
// workgroup of four waves owns one column tile of 16 targets of one parity at
// one level (the parity-sorted cell order of the symmetric kernel, one parity
// per XCD); wave w takes the tile's slots w, w + 4, ... and keeps all NR row
// tiles of its partial L~ in registers; the four partial sums are added in
// wave order through LDS (the same bits on every run) and added, scaled, to
// L~.  Per slot and 16-wide K block, lane (i, g) = (lane % 16, lane / 16)
// reads four consecutive entries 16 kb + 4 g + s (s < 4) of its target's M~
// (B operand) and of row i of every row tile of C (A operand): a permutation
// of the K order shared by both, so row-major loads of 32 bytes.  Register r
// of lane (i, g) holds D[4 r + g][i] (the FP64 layout the symmetric kernel's
// hfmm_selftest_mfma_f64 checks).
constexpr int kGlobalMfmaMaxTiles = 16;

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_global_pad_ops(const double* __restrict__ C, int k, int KP,
                                                        double* __restrict__ Cr) {
    const int64_t pos = blockIdx.x;
    const double* src = C + pos * k * k;
    double* dst = Cr + pos * KP * KP;
    for (int e = threadIdx.x; e < KP * KP; e += blockDim.x) {
        const int i = e / KP, j = e % KP;
        dst[e] = (i < k && j < k) ? src[static_cast<int64_t>(j) * k + i] : 0.;
    }
}

template <int NR>
__global__ __launch_bounds__(256) void k_global_m2l_mfma(M2LLevels lv, const int32_t* __restrict__ slots,
                                                        const int32_t* __restrict__ order,
                                                        const int32_t* __restrict__ pcount,
                                                        const int16_t* __restrict__ slotpos,
                                                        const double* __restrict__ Cr, const double* __restrict__ Mt,
                                                        int k, double* __restrict__ Lt) {
    constexpr int KP = 16 * NR;
    constexpr int W = 4;
    __shared__ double red[16 * KP];  // [target column][row]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col = lane & 15, g = lane >> 4;
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
    const int cnt = min(16, pc[e] - ptile * 16);
    if (cnt <= 0) return;  // the same for the whole workgroup
    first += ptile * 16;
    const int64_t coff = lv.cell_off[li];
    const int32_t* lslots = slots + coff * kFarSlots;
    const double* Ml = Mt + coff * k;
    const int tgt = col < cnt ? order[coff + first + col] : -1;
    d4 acc[NR];
#pragma unroll
    for (int rt = 0; rt < NR; ++rt) acc[rt] = d4{0., 0., 0., 0.};
    for (int s = wave; s < kFarSlots; s += W) {
        const int src = tgt >= 0 ? lslots[static_cast<int64_t>(tgt) * kFarSlots + s] : -1;
        if (__ballot(src >= 0) == 0) continue;  // uniform over the wave
        const double* C = Cr + static_cast<int64_t>(slotpos[e * kFarSlots + s]) * KP * KP + col * KP + 4 * g;
        const double* m = Ml + static_cast<int64_t>(src >= 0 ? src : 0) * k;
#pragma unroll 2
        for (int kb = 0; kb < NR; ++kb) {
            double b[4];
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2) {
                const int j = 16 * kb + 4 * g + s2;
                b[s2] = (src >= 0 && j < k) ? m[j] : 0.;
            }
#pragma unroll
            for (int rt = 0; rt < NR; ++rt) {
                const d4 a = *reinterpret_cast<const d4*>(C + 16 * rt * KP + 16 * kb);
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s2], b[s2], acc[rt], 0, 0, 0);
            }
        }
    }
    // partial sums in wave order: wave 0 writes, waves 1..3 add
    for (int w = 0; w < W; ++w) {
        if (wave == w) {
#pragma unroll
            for (int rt = 0; rt < NR; ++rt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int idx = col * KP + 16 * rt + 4 * r + g;
                    red[idx] = (w == 0 ? 0. : red[idx]) + acc[rt][r];
                }
        }
        __syncthreads();
    }
    const double scale = lv.scale[li];
    double* Ll = Lt + coff * k;
    for (int idx = tid; idx < 16 * KP; idx += 64 * W) {
        const int cc = idx / KP, i = idx % KP;
        if (cc < cnt && i < k) Ll[static_cast<int64_t>(order[coff + first + cc]) * k + i] += scale * red[idx];
    }
}

// ---------------------------------------------------------------------------
// The uniform scheme (FUnifKernel): expansions in Fourier space.  A cell's p^3
// node values, zero-padded to N^3 (N = 2p-1), are 3D-DFT'd; real input, so
// only kx < p is kept: NF = p N^2 complex entries, index kx + p (ky + N kz).
// One 64-lane wave per cell, the three axes one after the other through LDS.
// tw[j] = (cos 2 pi j/N, sin 2 pi j/N).

// M^ = DFT(pad(M)) (FUnifM2LHandler::applyZeroPaddingAndDFT), overwrite.
template <int P>
__global__ __launch_bounds__(64) void k_unif_forward(int64_t ncell, const double* __restrict__ M, int LD,
                                                     const double2* __restrict__ tw, double2* __restrict__ Mh) {
    constexpr int N = 2 * P - 1, NN = P * P * P, NF = P * N * N;
    __shared__ double sM[NN];
    __shared__ double2 sA[NN];         // [kx][b][c]: kx + P (b + P c)
    __shared__ double2 sB[P * N * P];  // [kx][ky][c]: kx + P (ky + N c)
    __shared__ double2 sW[N];
    const int64_t cell = blockIdx.x;
    if (cell >= ncell) return;
    const int lane = threadIdx.x;
    if (lane < N) sW[lane] = tw[lane];
    for (int i = lane; i < NN; i += 64) sM[i] = M[cell * LD + i];
    __syncthreads();
    for (int i = lane; i < NN; i += 64) {  // x: real -> complex
        const int kx = i % P, bc = i / P;
        double re = 0., im = 0.;
#pragma unroll
        for (int a = 0; a < P; ++a) {
            const double m = sM[a + P * bc];
            const double2 w = sW[(a * kx) % N];
            re += m * w.x;
            im -= m * w.y;
        }
        sA[i] = make_double2(re, im);
    }
    __syncthreads();
    for (int i = lane; i < P * N * P; i += 64) {  // y
        const int kx = i % P, ky = (i / P) % N, cc = i / (P * N);
        double re = 0., im = 0.;
#pragma unroll
        for (int b = 0; b < P; ++b) {
            const double2 v = sA[kx + P * (b + P * cc)];
            const double2 w = sW[(b * ky) % N];
            re += v.x * w.x + v.y * w.y;
            im += v.y * w.x - v.x * w.y;
        }
        sB[i] = make_double2(re, im);
    }
    __syncthreads();
    for (int i = lane; i < NF; i += 64) {  // z
        const int kxy = i % (P * N), kz = i / (P * N);
        double re = 0., im = 0.;
#pragma unroll
        for (int cc = 0; cc < P; ++cc) {
            const double2 v = sB[kxy + P * N * cc];
            const double2 w = sW[(cc * kz) % N];
            re += v.x * w.x + v.y * w.y;
            im += v.y * w.x - v.x * w.y;
        }
        Mh[cell * NF + i] = make_double2(re, im);
    }
}

// L += IDFT(L^) restricted to the p^3 nodes, over N^3
// (FUnifM2LHandler::unapplyZeroPaddingAndDFT + FBlas::add).
template <int P>
__global__ __launch_bounds__(64) void k_unif_inverse(int64_t ncell, const double2* __restrict__ Lh,
                                                     const double2* __restrict__ tw, double* __restrict__ L, int LD) {
    constexpr int N = 2 * P - 1, NN = P * P * P, NF = P * N * N;
    __shared__ double2 sF[NF];
    __shared__ double2 sH[P * N * P];  // [kx][ky][c]
    __shared__ double2 sG[NN];         // [kx][b][c]
    __shared__ double2 sW[N];
    const int64_t cell = blockIdx.x;
    if (cell >= ncell) return;
    const int lane = threadIdx.x;
    if (lane < N) sW[lane] = tw[lane];
    for (int i = lane; i < NF; i += 64) sF[i] = Lh[cell * NF + i];
    __syncthreads();
    for (int i = lane; i < P * N * P; i += 64) {  // z, outputs c < p
        const int kxy = i % (P * N), cc = i / (P * N);
        double re = 0., im = 0.;
#pragma unroll
        for (int kz = 0; kz < N; ++kz) {
            const double2 v = sF[kxy + P * N * kz];
            const double2 w = sW[(cc * kz) % N];
            re += v.x * w.x - v.y * w.y;
            im += v.x * w.y + v.y * w.x;
        }
        sH[i] = make_double2(re, im);
    }
    __syncthreads();
    for (int i = lane; i < NN; i += 64) {  // y, outputs b < p
        const int kx = i % P, b = (i / P) % P, cc = i / (P * P);
        double re = 0., im = 0.;
#pragma unroll
        for (int ky = 0; ky < N; ++ky) {
            const double2 v = sH[kx + P * (ky + N * cc)];
            const double2 w = sW[(b * ky) % N];
            re += v.x * w.x - v.y * w.y;
            im += v.x * w.y + v.y * w.x;
        }
        sG[i] = make_double2(re, im);
    }
    __syncthreads();
    constexpr double inv = 1. / (static_cast<double>(N) * N * N);
    for (int i = lane; i < NN; i += 64) {  // x: the Hermitian half back to real
        const int a = i % P, bc = i / P;
        double v = sG[P * bc].x;
        double h = 0.;
#pragma unroll
        for (int kx = 1; kx < P; ++kx) {
            const double2 g = sG[kx + P * bc];
            const double2 w = sW[(a * kx) % N];
            h += g.x * w.x - g.y * w.y;
        }
        v += 2. * h;
        L[cell * LD + i] += v * inv;
    }
}

// The M2L levels of one uniform-scheme launch: level i owns workgroups
// first[i] .. first[i + 1] - 1, one per target cell.
struct UnifLevels {
    int nlev;
    int64_t first[kMaxLevels + 1];
    const int32_t* slots[kMaxLevels];
    const uint64_t* keys[kMaxLevels];
    const double2* Mh[kMaxLevels];
    double2* Lh[kMaxLevels];
    double scale[kMaxLevels];
};

// L^[cell] += scale K^_pos . M^[src] entrywise (FUnifM2LHandler::applyFC)
// over the cell's 189 slots in slot order, one wave per target cell; every
// M2L level in one launch (per-level launches left the top levels' 64 and
// 512 cells to a nearly idle GPU).
template <int P>
__global__ __launch_bounds__(64) void k_unif_m2l(UnifLevels lv, const int16_t* __restrict__ slotpos,
                                                 const double2* __restrict__ K) {
    constexpr int N = 2 * P - 1, NF = P * N * N, R = (NF + 63) / 64;
    const int64_t bid = blockIdx.x;
    if (bid >= lv.first[lv.nlev]) return;
    int li = 0;
    while (bid >= lv.first[li + 1]) ++li;
    const int64_t cell = bid - lv.first[li];
    const int32_t* __restrict__ slots = lv.slots[li];
    const double2* __restrict__ Mh = lv.Mh[li];
    double2* __restrict__ Lh = lv.Lh[li];
    const int lane = threadIdx.x;
    const int par = static_cast<int>(lv.keys[li][cell] & 7u);
    double2 acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = make_double2(0., 0.);
    for (int s = 0; s < kFarSlots; ++s) {
        const int src = slots[cell * kFarSlots + s];
        if (src < 0) continue;  // the same for every lane
        const double2* Kp = K + static_cast<int64_t>(slotpos[par * kFarSlots + s]) * NF;
        const double2* m = Mh + static_cast<int64_t>(src) * NF;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int f = lane + 64 * r;
            if (f < NF) {
                const double2 k = Kp[f], x = m[f];
                acc[r].x += k.x * x.x - k.y * x.y;
                acc[r].y += k.x * x.y + k.y * x.x;
            }
        }
    }
    const double scale = lv.scale[li];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = lane + 64 * r;
        if (f < NF) {
            double2 o = Lh[cell * NF + f];
            o.x += scale * acc[r].x;
            o.y += scale * acc[r].y;
            Lh[cell * NF + f] = o;
        }
    }
}

#define HFMM_UNIF_ORDERS(X) X(5) X(6) X(7) X(8)

int run_unif_forward(Context* c, int l) {
    const int64_t nc = c->ncell[l];
    if (nc == 0) return HFMM_OK;
    const dim3 grid(static_cast<unsigned>(nc));
    const double* M = c->d_M.as<double>() + c->lvl_off[l] * c->LD;
    double2* Mh = reinterpret_cast<double2*>(c->d_Mt.as<double>() + c->lvl_off[l] * c->glob.k);
    const double2* tw = reinterpret_cast<const double2*>(c->d_utw.as<double>());
    prof_begin(c, "unif_forward");
    switch (c->ops.p) {
#define CASE(PP) case PP: hipLaunchKernelGGL(k_unif_forward<PP>, grid, dim3(64), 0, c->stream, nc, M, c->LD, tw, Mh); break;
        HFMM_UNIF_ORDERS(CASE)
#undef CASE
        default: return fail(c, HFMM_ERR_ARG, "uniform M2L: unsupported order");
    }
    prof_end(c);
    return check_launch(c, "unif_forward");
}

int run_unif_inverse(Context* c, int l) {
    const int64_t nc = c->ncell[l];
    if (nc == 0) return HFMM_OK;
    const dim3 grid(static_cast<unsigned>(nc));
    const double2* Lh = reinterpret_cast<const double2*>(c->d_Lt.as<double>() + c->lvl_off[l] * c->glob.k);
    double* L = c->d_L.as<double>() + c->lvl_off[l] * c->LD;
    const double2* tw = reinterpret_cast<const double2*>(c->d_utw.as<double>());
    prof_begin(c, "unif_inverse");
    switch (c->ops.p) {
#define CASE(PP) case PP: hipLaunchKernelGGL(k_unif_inverse<PP>, grid, dim3(64), 0, c->stream, nc, Lh, tw, L, c->LD); break;
        HFMM_UNIF_ORDERS(CASE)
#undef CASE
        default: return fail(c, HFMM_ERR_ARG, "uniform M2L: unsupported order");
    }
    prof_end(c);
    return check_launch(c, "unif_inverse");
}

int run_unif_m2l(Context* c) {
    const int L = c->height() - 1;
    const int lo = c->work_upper > c->top() ? c->work_upper : c->top();
    const int hi = c->work_lower - 1 < L ? c->work_lower - 1 : L;
    const int k = c->glob.k;
    UnifLevels lv{};
    lv.first[0] = 0;
    for (int l = lo; l <= hi; ++l) {
        const int64_t nc = c->ncell[l];
        if (nc == 0) continue;
        const int i = lv.nlev++;
        lv.scale[i] = 2. / (c->P.box_width / static_cast<double>(int64_t(1) << (l - c->set_levels)));
        lv.slots[i] = c->d_m2l.as<int32_t>() + static_cast<size_t>(c->m2l_off[l]) * kFarSlots;
        lv.keys[i] = c->d_cellkey.as<uint64_t>() + c->lvl_off[l];
        lv.Mh[i] = reinterpret_cast<const double2*>(c->d_Mt.as<double>() + c->lvl_off[l] * k);
        lv.Lh[i] = reinterpret_cast<double2*>(c->d_Lt.as<double>() + c->lvl_off[l] * k);
        lv.first[i + 1] = lv.first[i] + nc;
    }
    if (lv.nlev == 0) return HFMM_OK;
    const int16_t* sp = c->d_slotpos.as<int16_t>();
    const double2* K = reinterpret_cast<const double2*>(c->d_uK.as<double>());
    const dim3 grid(static_cast<unsigned>(lv.first[lv.nlev]));
    prof_begin(c, "unif_m2l");
    switch (c->ops.p) {
#define CASE(PP) case PP: hipLaunchKernelGGL(k_unif_m2l<PP>, grid, dim3(64), 0, c->stream, lv, sp, K); break;
        HFMM_UNIF_ORDERS(CASE)
#undef CASE
        default: prof_end(c); return fail(c, HFMM_ERR_ARG, "uniform M2L: unsupported order");
    }
    HFMM_TRY(check_launch(c, "unif_m2l"));
    prof_end(c);
    return HFMM_OK;
}

}  // namespace

int run_global_compress(Context* c, int l) {
    if (c->m2l_scheme == HFMM_M2L_UNIFORM) return run_unif_forward(c, l);
    const int64_t nc = c->ncell[l];
    if (nc == 0) return HFMM_OK;
    const int k = c->glob.k, R = (k + 63) / 64;
    const double* M = c->d_M.as<double>() + c->lvl_off[l] * c->LD;
    double* Mt = c->d_Mt.as<double>() + c->lvl_off[l] * k;
    prof_begin(c, "global_compress");
    const double* Bt = c->d_gBt.as<double>();
    const int nn = c->ops.nn;
    const dim3 grid(static_cast<unsigned>(nc));
    switch (R) {
#define CASE(RR) case RR: hipLaunchKernelGGL(k_global_compress<RR>, grid, dim3(64), 0, c->stream, nc, M, c->LD, Bt, nn, k, Mt); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default: return fail(c, HFMM_ERR_ARG, "global M2L rank above 512");
    }
    prof_end(c);
    return check_launch(c, "global_compress");
}

int run_global_decompress(Context* c, int l) {
    if (c->m2l_scheme == HFMM_M2L_UNIFORM) return run_unif_inverse(c, l);
    const int64_t nc = c->ncell[l];
    if (nc == 0) return HFMM_OK;
    const int k = c->glob.k;
    prof_begin(c, "global_decompress");
    hipLaunchKernelGGL(k_global_decompress, dim3(static_cast<unsigned>(nc)), dim3(64), 0, c->stream, nc,
                       c->d_Lt.as<double>() + c->lvl_off[l] * k, k, c->d_gU.as<double>(), c->ops.nn,
                       c->d_L.as<double>() + c->lvl_off[l] * c->LD, c->LD);
    prof_end(c);
    return check_launch(c, "global_decompress");
}

// The padded row-major operators of the MFMA M2L, built once per operator set.
static int prepare_global_mfma(Context* c, int KP) {
    if (c->gcr_kp == KP) return HFMM_OK;
    HFMM_TRY(ensure(c, c->d_gCr, sizeof(double) * 343 * static_cast<size_t>(KP) * KP));
    hipLaunchKernelGGL(k_global_pad_ops, dim3(343), dim3(256), 0, c->stream, c->d_gC.as<double>(), c->glob.k, KP,
                       c->d_gCr.as<double>());
    HFMM_TRY(check_launch(c, "global_pad_ops"));
    c->gcr_kp = KP;
    return HFMM_OK;
}

int run_global_m2l(Context* c) {
    if (c->m2l_scheme == HFMM_M2L_UNIFORM) return run_unif_m2l(c);
    const int k = c->glob.k, R = (k + 63) / 64;
    if (R > kMaxRankRows) return fail(c, HFMM_ERR_ARG, "global M2L rank above 512");
    const int NR = (k + 15) / 16;
    if (NR <= kGlobalMfmaMaxTiles) {
        HFMM_TRY(prepare_global_mfma(c, 16 * NR));
        int tiles = 0;
        const M2LLevels lv = make_levels(c, 16, tiles);
        if (tiles == 0) return HFMM_OK;
        prof_begin(c, "global_m2l");
        const dim3 grid(static_cast<unsigned>(tiles));
        const int32_t* slots = c->d_m2l.as<int32_t>();
        const int32_t* ord = c->d_order.as<int32_t>();
        const int32_t* pc = c->d_pcount.as<int32_t>();
        const int16_t* sp = c->d_slotpos.as<int16_t>();
        const double* Cr = c->d_gCr.as<double>();
        const double* Mt = c->d_Mt.as<double>();
        double* Lt = c->d_Lt.as<double>();
        switch (NR) {
#define CASE(N) case N: hipLaunchKernelGGL(k_global_m2l_mfma<N>, grid, dim3(256), 0, c->stream, lv, slots, ord, pc, sp, Cr, Mt, k, Lt); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
            CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
            default: break;
        }
        prof_end(c);
        return check_launch(c, "global_m2l");
    }
    const int L = c->height() - 1;
    // the M2L levels of the working range, from the top
    const int lo = c->work_upper > c->top() ? c->work_upper : c->top();
    const int hi = c->work_lower - 1 < L ? c->work_lower - 1 : L;
    prof_begin(c, "global_m2l");
    for (int l = lo; l <= hi; ++l) {
        const int64_t nc = c->ncell[l];
        if (nc == 0) continue;
        const double scale = 2. / (c->P.box_width / static_cast<double>(int64_t(1) << (l - c->set_levels)));
        const int32_t* slots = c->d_m2l.as<int32_t>() + static_cast<size_t>(c->m2l_off[l]) * kFarSlots;
        const uint64_t* keys = c->d_cellkey.as<uint64_t>() + c->lvl_off[l];
        const double* Mt = c->d_Mt.as<double>() + c->lvl_off[l] * k;
        double* Lt = c->d_Lt.as<double>() + c->lvl_off[l] * k;
        const int16_t* sp = c->d_slotpos.as<int16_t>();
        const double* Cg = c->d_gC.as<double>();
        const dim3 grid(static_cast<unsigned>(nc));
        switch (R) {
#define CASE(RR) case RR: hipLaunchKernelGGL(k_global_m2l<RR>, grid, dim3(64), 0, c->stream, nc, slots, keys, sp, Cg, Mt, k, scale, Lt); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            default: break;
        }
        HFMM_TRY(check_launch(c, "global_m2l"));
    }
    prof_end(c);
    return HFMM_OK;
}

}  // namespace hfmm
