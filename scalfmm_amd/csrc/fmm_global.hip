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

}  // namespace

int run_global_compress(Context* c, int l) {
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

int run_global_m2l(Context* c) {
    const int k = c->glob.k, R = (k + 63) / 64;
    if (R > kMaxRankRows) return fail(c, HFMM_ERR_ARG, "global M2L rank above 512");
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
