// fmm_device.hip -- FP64 FMM passes of the Chebyshev 1/r kernel on CDNA4.
//
//   P2M  per leaf: M_n += sum_p q_p Sx(xi_p, n_x) Sy(.., n_y) Sz(.., n_z)
//        (Src/Kernels/Chebyshev/FChebInterpolator.hpp:625-762)
//   M2M  per parent: three 1-D contractions per child (FChebInterpolator.hpp:560-584)
//   M2L  per level, on MFMA: for tiles of 16 targets of one parity and each
//        of the 189 far-field slots, C = V_c^T P(M_src) then L += U_c P^-1 C
//        (Src/Kernels/Chebyshev/FChebSymKernel.hpp:210-328)
//   L2L  per parent: transposed contractions (FChebInterpolator.hpp:587-611)
//   L2P  per leaf: potential and q * gradient (FChebInterpolator.hpp:1220-1417)
//   P2P  per target particle over the 27-leaf neighbourhood
//        (Src/Kernels/P2P/FP2P.hpp:103-269 conventions)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fmm_common.hpp"
#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// S(xi, root_j) = 1/p + 2/p sum_{o>=1} T_o(xi) T_o(root_j); dS = d/dxi S.
template <int P, bool GRAD>
__device__ __forceinline__ void interp_weights(double xi, const double* __restrict__ sT, double* S, double* dS) {
    double T[P], U[P];
    T[0] = 1.;
    if (P > 1) T[1] = xi;
    U[0] = 1.;
    if (P > 1) U[1] = 2. * xi;
#pragma unroll
    for (int j = 2; j < P; ++j) {
        T[j] = 2. * xi * T[j - 1] - T[j - 2];
        if (GRAD) U[j] = 2. * xi * U[j - 1] - U[j - 2];
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        double s = 1. / P, ds = 0.;
#pragma unroll
        for (int o = 1; o < P; ++o) {
            const double t = sT[o * P + j];
            s += 2. / P * T[o] * t;
            if (GRAD) ds += 2. / P * (o * U[o - 1]) * t;
        }
        S[j] = s;
        if (GRAD) dS[j] = ds;
    }
}

constexpr int kUniformMaxOrder = 8;  // the uniform scheme's dense M2L holds p^3 <= 512 ranks

// FUnifRoots::L and ::dL (Src/Kernels/Uniform/FUnifRoots.hpp:49-118) on the
// equispaced roots r: L_n by the scaled product, dL_n by the sum of products
// over the root table; the argument clamped to [-1, 1].
template <int P, bool GRAD>
__device__ __forceinline__ void lagrange_weights(double xi, const double* __restrict__ r, double* S, double* dS) {
    const double x = fmin(fmax(xi, -1.), 1.);
#pragma unroll
    for (int n = 0; n < P; ++n) {
        const int omn = P - n - 1;
        double fn = 1., fo = 1.;
        for (int i = 2; i <= n; ++i) fn *= i;
        for (int i = 2; i <= omn; ++i) fo *= i;
        double scale = (omn % 2) ? -1. : 1.;
        double p2 = 1.;
        for (int i = 0; i < P - 1; ++i) p2 *= 2.;
        scale /= p2 * fn * fo;
        double L = 1.;
#pragma unroll
        for (int m = 0; m < P; ++m)
            if (m != n) L *= static_cast<double>(P - 1) * (x + 1.) - 2. * m;
        S[n] = L * scale;
        if (GRAD) {
            double num = 0., den = 1.;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                if (q == n) continue;
                double t = 1.;
#pragma unroll
                for (int m = 0; m < P; ++m)
                    if (m != n && m != q) t *= x - r[m];
                num += t;
                den *= r[n] - r[q];
            }
            dS[n] = num / den;
        }
    }
}

// ---------------------------------------------------------------------------
// P2M: one wave per window of 64 sorted particles.  Each lane forms the
// interpolation weights of its particle in the LDS; the wave then walks the
// leaves that start in the window, lanes over the coefficients.  A leaf that
// ends inside the window sums its particles from the LDS in one round (most
// leaves hold a few particles); the last leaf of a window may run past it and
// is summed in rounds of 64 particles from its first one.  A leaf's rounds
// always begin at its first particle, so its sums do not depend on where the
// windows fall.
template <int P, int BASIS>
__device__ __forceinline__ void p2m_weights(const double* __restrict__ sp, int64_t n, int64_t e, Box bx,
                                            uint64_t lkey, const double* __restrict__ sT,
                                            const double* __restrict__ uroots, double* S, double* Q) {
    double a[3], b[3];
    leaf_bounds(bx, lkey, a, b);
    const double pos[3] = {sp[e], sp[n + e], sp[2 * n + e]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double xi = (2. * pos[d] - b[d] - a[d]) / (b[d] - a[d]);
        if constexpr (BASIS == 1)
            lagrange_weights<P, false>(xi, uroots, S + d * P, nullptr);
        else
            interp_weights<P, false>(xi, sT, S + d * P, nullptr);
    }
    *Q = sp[3 * n + e];
}

template <int P, int BASIS>
__global__ __launch_bounds__(64) void k_p2m(int64_t n, const int32_t* __restrict__ pleaf,
                                           const uint64_t* __restrict__ lkeys, const int32_t* __restrict__ lstart,
                                           const double* __restrict__ sp, Box bx, const double* __restrict__ Troots,
                                           const double* __restrict__ uroots, double* __restrict__ M, int LD) {
    constexpr int NN = P * P * P;
    constexpr int PER = (NN + 63) / 64;
    __shared__ double sS[64][3 * P];
    __shared__ double sQ[64];
    __shared__ double sT[P * P];
    __shared__ int sStart[65];
    const int lane = threadIdx.x;
    const int64_t w0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t wend = min(w0 + 64, n);
    for (int i = lane; i < P * P; i += 64) sT[i] = Troots[i];
    // the leaves that start in the window: first .. last
    int first = pleaf[w0];
    if (lstart[first] < w0) ++first;
    const int last = pleaf[wend - 1];
    if (first > last) return;  // inside one long leaf: its first window sums it
    __syncthreads();
    if (w0 + lane < wend) p2m_weights<P, BASIS>(sp, n, w0 + lane, bx, lkeys[pleaf[w0 + lane]], sT, uroots,
                                                &sS[lane][0], &sQ[lane]);
    // the bounds of the window's leaves (at most 64) through the LDS: the leaf
    // loop then holds no vector load, whose wait would also cover the stores
    // of the previous leaves
    if (lane <= last - first) sStart[lane] = lstart[first + lane];
    if (lane == 0) sStart[last - first + 1] = lstart[last + 1];
    __syncthreads();
    for (int leaf = first; leaf <= last; ++leaf) {
        const int s0 = sStart[leaf - first], s1 = sStart[leaf - first + 1];
        double acc[PER];
#pragma unroll
        for (int t = 0; t < PER; ++t) acc[t] = 0.;
        // one round over particles m0 .. m1-1 of the LDS rows; per coefficient the
        // terms are added in particle order
        auto round = [&](int m0, int m1) {
            double sum[PER];
#pragma unroll
            for (int t = 0; t < PER; ++t) sum[t] = 0.;
            for (int m = m0; m < m1; ++m) {
                const double q = sQ[m];
#pragma unroll
                for (int t = 0; t < PER; ++t) {
                    const int idx = lane + 64 * t;
                    const int i = idx % P, j = (idx / P) % P, k = idx / (P * P);
                    if (NN % 64 == 0 || idx < NN) sum[t] += q * sS[m][P + j] * sS[m][2 * P + k] * sS[m][i];
                }
            }
#pragma unroll
            for (int t = 0; t < PER; ++t) acc[t] += sum[t];
        };
        if (s1 <= wend) {
            round(static_cast<int>(s0 - w0), static_cast<int>(s1 - w0));
        } else {
            // the window's last leaf runs past it: rounds of 64 from s0
            for (int base = s0; base < s1; base += 64) {
                const int cnt = min(64, s1 - base);
                __syncthreads();
                if (lane < cnt)
                    p2m_weights<P, BASIS>(sp, n, base + lane, bx, lkeys[leaf], sT, uroots, &sS[lane][0], &sQ[lane]);
                __syncthreads();
                round(0, cnt);
            }
        }
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int idx = lane + 64 * t;
            if (idx < LD) M[static_cast<int64_t>(leaf) * LD + idx] = idx < NN ? acc[t] : 0.;  // zero padding
        }
    }
}

// ---------------------------------------------------------------------------
// M2M and L2L: a workgroup of 4 waves shares the 8 children's interpolation
// matrices in the LDS, one parent per wave.  The waves of a workgroup walk
// different numbers of children, so the separable passes through each wave's
// own LDS buffers are ordered by wave-local fences: the wave's LDS accesses run
// in order, and the fence keeps the compiler from moving other lanes' reads
// above the writes they depend on.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// M2M: child[m] -> parent[n] += Sx[nx][mx] Sy[ny][my] Sz[nz][mz] child[m].
template <int P>
__global__ __launch_bounds__(256) void k_m2m(int np, const int32_t* __restrict__ cstart,
                                            const uint64_t* __restrict__ ckeys, const double* __restrict__ interp,
                                            const double* __restrict__ Mc, double* __restrict__ Mp, int LD) {
    constexpr int NN = P * P * P;
    constexpr int PER = (NN + 63) / 64;
    __shared__ double sI[8 * 3 * P * P];
    __shared__ double sb[4][2][NN];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < 8 * 3 * P * P; i += 256) sI[i] = interp[i];
    __syncthreads();
    const int parent = blockIdx.x * 4 + wave;
    if (parent >= np) return;  // after the workgroup's only barrier
    double* b0 = sb[wave][0];
    double* b1 = sb[wave][1];
    double acc[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) acc[t] = 0.;
    const int c0 = cstart[parent], c1 = cstart[parent + 1];
    for (int ch = c0; ch < c1; ++ch) {
        const int cidx = static_cast<int>(ckeys[ch] & 7u);
        const double* Sx = sI + (cidx * 3 + 0) * P * P;
        const double* Sy = sI + (cidx * 3 + 1) * P * P;
        const double* Sz = sI + (cidx * 3 + 2) * P * P;
        for (int i = lane; i < NN; i += 64) b0[i] = Mc[static_cast<int64_t>(ch) * LD + i];
        wave_lds_order();
        for (int idx = lane; idx < NN; idx += 64) {  // x
            const int nx = idx % P, rest = idx / P;
            double s = 0.;
#pragma unroll
            for (int m = 0; m < P; ++m) s += Sx[nx * P + m] * b0[rest * P + m];
            b1[idx] = s;
        }
        wave_lds_order();
        for (int idx = lane; idx < NN; idx += 64) {  // y
            const int i = idx % P, ny = (idx / P) % P, k = idx / (P * P);
            double s = 0.;
#pragma unroll
            for (int m = 0; m < P; ++m) s += Sy[ny * P + m] * b1[k * P * P + m * P + i];
            b0[idx] = s;
        }
        wave_lds_order();
#pragma unroll
        for (int t = 0; t < PER; ++t) {  // z
            const int idx = lane + 64 * t;
            if (idx < NN) {
                const int i = idx % P, j = (idx / P) % P, nz = idx / (P * P);
                double s = 0.;
#pragma unroll
                for (int m = 0; m < P; ++m) s += Sz[nz * P + m] * b0[m * P * P + j * P + i];
                acc[t] += s;
            }
        }
        wave_lds_order();
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int idx = lane + 64 * t;
        if (idx < LD) Mp[static_cast<int64_t>(parent) * LD + idx] = idx < NN ? acc[t] : 0.;  // zero padding
    }
}

// L2L: child[m] += sum_n Sx[nx][mx] Sy[ny][my] Sz[nz][mz] parent[n].
template <int P>
__global__ __launch_bounds__(256) void k_l2l(int np, const int32_t* __restrict__ cstart,
                                            const uint64_t* __restrict__ ckeys, const double* __restrict__ interp,
                                            const double* __restrict__ Lp, double* __restrict__ Lc, int LD) {
    constexpr int NN = P * P * P;
    __shared__ double sI[8 * 3 * P * P];
    __shared__ double sb[4][3][NN];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < 8 * 3 * P * P; i += 256) sI[i] = interp[i];
    __syncthreads();
    const int parent = blockIdx.x * 4 + wave;
    if (parent >= np) return;  // after the workgroup's only barrier
    double* par = sb[wave][0];
    double* b1 = sb[wave][1];
    double* b2 = sb[wave][2];
    for (int i = lane; i < NN; i += 64) par[i] = Lp[static_cast<int64_t>(parent) * LD + i];
    wave_lds_order();
    const int c0 = cstart[parent], c1 = cstart[parent + 1];
    for (int ch = c0; ch < c1; ++ch) {
        const int cidx = static_cast<int>(ckeys[ch] & 7u);
        const double* Sx = sI + (cidx * 3 + 0) * P * P;
        const double* Sy = sI + (cidx * 3 + 1) * P * P;
        const double* Sz = sI + (cidx * 3 + 2) * P * P;
        for (int idx = lane; idx < NN; idx += 64) {  // x
            const int mx = idx % P, rest = idx / P;
            double s = 0.;
#pragma unroll
            for (int nn = 0; nn < P; ++nn) s += Sx[nn * P + mx] * par[rest * P + nn];
            b1[idx] = s;
        }
        wave_lds_order();
        for (int idx = lane; idx < NN; idx += 64) {  // y
            const int i = idx % P, my = (idx / P) % P, k = idx / (P * P);
            double s = 0.;
#pragma unroll
            for (int nn = 0; nn < P; ++nn) s += Sy[nn * P + my] * b1[k * P * P + nn * P + i];
            b2[idx] = s;
        }
        wave_lds_order();
        for (int idx = lane; idx < NN; idx += 64) {  // z
            const int i = idx % P, j = (idx / P) % P, mz = idx / (P * P);
            double s = 0.;
#pragma unroll
            for (int nn = 0; nn < P; ++nn) s += Sz[nn * P + mz] * b2[nn * P * P + j * P + i];
            Lc[static_cast<int64_t>(ch) * LD + idx] += s;
        }
        wave_lds_order();
    }
}

// ---------------------------------------------------------------------------
// L2P: one lane per sorted particle; adds into the sorted results.
//   phi = sum_k Sz_k sum_j Sy_j sum_i Sx_i L[k][j][i], gradients with dS on one
// axis; Sx, dSx, Sy, dSy stay in registers, Sz and dSz are formed per k.
// Leaves hold from one to thousands of particles (a Plummer tree of 10M
// particles at height 8: 1.13M leaves, 641k of them with one particle), so a
// wave per leaf would leave most lanes idle.  Each lane reads its leaf's local
// expansion straight from memory: the lanes of one leaf read the same
// addresses (one fetch for the wave), the lanes of small leaves their own.
template <int P>
__global__ __launch_bounds__(256) void k_l2p(int64_t n, const int32_t* __restrict__ pleaf,
                                            const uint64_t* __restrict__ lkeys, const double* __restrict__ sp, Box bx,
                                            const double* __restrict__ Troots, const double* __restrict__ L, int LD,
                                            double* __restrict__ res) {
    __shared__ double sT[P * P];
    for (int i = threadIdx.x; i < P * P; i += 256) sT[i] = Troots[i];
    __syncthreads();
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e < n) {
        const int leaf = pleaf[e];
        double a[3], b[3];
        leaf_bounds(bx, lkeys[leaf], a, b);
        const double* __restrict__ sL = L + static_cast<int64_t>(leaf) * LD;
        const double pos[3] = {sp[e], sp[n + e], sp[2 * n + e]};
        double Sx[P], dSx[P], Sy[P], dSy[P];
        const double xi0 = (2. * pos[0] - b[0] - a[0]) / (b[0] - a[0]);
        const double xi1 = (2. * pos[1] - b[1] - a[1]) / (b[1] - a[1]);
        const double xi2 = (2. * pos[2] - b[2] - a[2]) / (b[2] - a[2]);
        interp_weights<P, true>(xi0, sT, Sx, dSx);
        interp_weights<P, true>(xi1, sT, Sy, dSy);
        // T_o(xi2) and o U_{o-1}(xi2) for the z weights
        double Tz[P], dTz[P];
        {
            double t0 = 1., t1 = xi2, u0 = 1., u1 = 2. * xi2;
            Tz[0] = 1.;
            dTz[0] = 0.;
            if (P > 1) { Tz[1] = xi2; dTz[1] = 1.; }
#pragma unroll
            for (int o = 2; o < P; ++o) {
                const double t2 = 2. * xi2 * t1 - t0;
                t0 = t1; t1 = t2;
                Tz[o] = t2;
                dTz[o] = o * u1;
                const double u2 = 2. * xi2 * u1 - u0;
                u0 = u1; u1 = u2;
            }
        }
        double ph = 0., gx = 0., gy = 0., gz = 0.;
#pragma unroll 1
        for (int k = 0; k < P; ++k) {
            double sz = 1. / P, dsz = 0.;
#pragma unroll
            for (int o = 1; o < P; ++o) {
                const double t = sT[o * P + k];
                sz += 2. / P * Tz[o] * t;
                dsz += 2. / P * dTz[o] * t;
            }
            double tp = 0., tx = 0., ty = 0.;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                double va = 0., vb = 0.;
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const double l = sL[(k * P + j) * P + i];
                    va += l * Sx[i];
                    vb += l * dSx[i];
                }
                tp += va * Sy[j];
                tx += vb * Sy[j];
                ty += va * dSy[j];
            }
            ph += tp * sz;
            gx += tx * sz;
            gy += ty * sz;
            gz += tp * dsz;
        }
        res[e] += ph;
        res[n + e] += gx * (2. / (b[0] - a[0]));
        res[2 * n + e] += gy * (2. / (b[1] - a[1]));
        res[3 * n + e] += gz * (2. / (b[2] - a[2]));
    }
}

// L2P on the uniform (Lagrange) basis (FUnifInterpolator.hpp:476-640,
// applyL2P + applyL2PGradient): one lane per sorted particle, the leaf's local
// expansion read directly as in k_l2p.
template <int P>
__global__ __launch_bounds__(256) void k_l2p_lagrange(int64_t n, const int32_t* __restrict__ pleaf,
                                                     const uint64_t* __restrict__ lkeys,
                                                     const double* __restrict__ sp, Box bx,
                                                     const double* __restrict__ uroots, const double* __restrict__ L,
                                                     int LD, double* __restrict__ res) {
    __shared__ double sR[P];
    if (threadIdx.x < P) sR[threadIdx.x] = uroots[threadIdx.x];
    __syncthreads();
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e < n) {
        const int leaf = pleaf[e];
        double a[3], b[3];
        leaf_bounds(bx, lkeys[leaf], a, b);
        const double* __restrict__ sL = L + static_cast<int64_t>(leaf) * LD;
        const double pos[3] = {sp[e], sp[n + e], sp[2 * n + e]};
        double S[3][P], dS[3][P];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double xi = (2. * pos[d] - b[d] - a[d]) / (b[d] - a[d]);
            lagrange_weights<P, true>(xi, sR, S[d], dS[d]);
        }
        double ph = 0., gx = 0., gy = 0., gz = 0.;
        for (int k = 0; k < P; ++k)
#pragma unroll
            for (int j = 0; j < P; ++j) {
                double va = 0., vb = 0.;
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const double l = sL[(k * P + j) * P + i];
                    va += l * S[0][i];
                    vb += l * dS[0][i];
                }
                ph += va * S[1][j] * S[2][k];
                gx += vb * S[1][j] * S[2][k];
                gy += va * dS[1][j] * S[2][k];
                gz += va * S[1][j] * dS[2][k];
            }
        res[e] += ph;
        res[n + e] += gx * (2. / (b[0] - a[0]));
        res[2 * n + e] += gy * (2. / (b[1] - a[1]));
        res[3 * n + e] += gz * (2. / (b[2] - a[2]));
    }
}

// ---------------------------------------------------------------------------
// P2P: one thread per target particle, sources over its own leaf and the 26
// neighbour slots; writes (initialises) the sorted results.
__global__ __launch_bounds__(256) void k_p2p_f64(int64_t n, const int32_t* __restrict__ pleaf,
                                                const int32_t* __restrict__ lstart, const int32_t* __restrict__ slots,
                                                const double* __restrict__ sp, double* __restrict__ res) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = pleaf[i];
    if (lstart[li + 1] - lstart[li] > kP2PBig) return;  // k_p2p_f64_big
    const double tx = sp[i], ty = sp[n + i], tz = sp[2 * n + i];
    double pot = 0., fx = 0., fy = 0., fz = 0.;
    for (int s = -1; s < kNbrSlots; ++s) {
        const int nb = s < 0 ? li : slots[static_cast<int64_t>(li) * kNbrSlots + s];
        if (nb < 0) continue;
        const int j0 = lstart[nb], j1 = lstart[nb + 1];
        for (int j = j0; j < j1; ++j) {
            if (j == i) continue;
            const double dx = tx - sp[j], dy = ty - sp[n + j], dz = tz - sp[2 * n + j];
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double inv3 = inv * inv * inv;
            const double qs = sp[3 * n + j];
            pot += inv * qs;
            fx += -inv3 * dx * qs;
            fy += -inv3 * dy * qs;
            fz += -inv3 * dz * qs;
        }
    }
    res[i] = pot;
    res[n + i] = fx;
    res[2 * n + i] = fy;
    res[3 * n + i] = fz;
}

// k_p2p_f64 for R charge vectors at once: the separation, 1/r and 1/r^3 of a
// pair once, then the terms of each vector, each summed as in k_p2p_f64 (the
// same bits per vector).  qc: the chunk's sorted charges as [n][R]; res: the
// results of its first vector, the next vectors 4n further each.
template <int R>
__global__ __launch_bounds__(256) void k_p2p_f64_rhs(int64_t n, const int32_t* __restrict__ pleaf,
                                                    const int32_t* __restrict__ lstart,
                                                    const int32_t* __restrict__ slots, const double* __restrict__ sp,
                                                    const double* __restrict__ qc, double* __restrict__ res) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = pleaf[i];
    if (lstart[li + 1] - lstart[li] > kP2PBig) return;  // k_p2p_f64_big, per vector
    const double tx = sp[i], ty = sp[n + i], tz = sp[2 * n + i];
    double pot[R], fx[R], fy[R], fz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pot[r] = fx[r] = fy[r] = fz[r] = 0.;
    for (int s = -1; s < kNbrSlots; ++s) {
        const int nb = s < 0 ? li : slots[static_cast<int64_t>(li) * kNbrSlots + s];
        if (nb < 0) continue;
        const int j0 = lstart[nb], j1 = lstart[nb + 1];
        for (int j = j0; j < j1; ++j) {
            if (j == i) continue;
            const double dx = tx - sp[j], dy = ty - sp[n + j], dz = tz - sp[2 * n + j];
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double inv3 = inv * inv * inv;
            const double* q = qc + static_cast<int64_t>(j) * R;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double qs = q[r];
                pot[r] += inv * qs;
                fx[r] += -inv3 * dx * qs;
                fy[r] += -inv3 * dy * qs;
                fz[r] += -inv3 * dz * qs;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double* o = res + static_cast<int64_t>(r) * 4 * n;
        o[i] = pot[r];
        o[n + i] = fx[r];
        o[2 * n + i] = fy[r];
        o[3 * n + i] = fz[r];
    }
}

// P2P of the big leaves (more than kP2PBig particles).  Each wave takes work
// items of kP2PItem consecutive targets of one leaf (two per lane) and walks
// the 27 neighbour leaves in blocks of 8 sources.  Every source address is
// wave-uniform, so the sources arrive through scalar loads into SGPRs and
// feed the FP64 arithmetic as scalar operands (no LDS, no barriers).  1/r is
// v_rsq_f64 refined by two Newton steps, y <- y (3 - r^2 y^2) / 2.  A block
// running past the leaf's end masks the extra sources (the sorted arrays are
// padded, so the reads stay in bounds).  Sources are summed in the same order
// as k_p2p_f64.
constexpr int kP2PBlock64 = 8;

__device__ __forceinline__ double rsqrt_f64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    y = y * fma(-0.5 * r2, y * y, 1.5);
    y = y * fma(-0.5 * r2, y * y, 1.5);
    return y;
}

template <bool MASK>
__device__ __forceinline__ void pair_f64(const double (&t)[3], double sx, double sy, double sz, double sq, bool skip,
                                         double (&acc)[4]) {
    const double dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
    double inv = rsqrt_f64(dx * dx + dy * dy + dz * dz);
    if (MASK) inv = skip ? 0. : inv;
    const double qi = sq * inv;
    const double qi3 = qi * inv * inv;
    acc[0] += qi;
    acc[1] -= qi3 * dx;
    acc[2] -= qi3 * dy;
    acc[3] -= qi3 * dz;
}

template <bool SELF>
__device__ __forceinline__ void p2p_f64_leaf(const double* __restrict__ sp, int64_t n, int j0, int j1,
                                             const double (&t)[2][3], const int (&ti)[2], double (&acc)[2][4]) {
    int jb = j0;
    for (; jb + kP2PBlock64 <= j1; jb += kP2PBlock64) {
#pragma unroll
        for (int jj = 0; jj < kP2PBlock64; ++jj) {
            const int j = jb + jj;
            const double sx = sp[j], sy = sp[n + j], sz = sp[2 * n + j], sq = sp[3 * n + j];
#pragma unroll
            for (int u = 0; u < 2; ++u) pair_f64<SELF>(t[u], sx, sy, sz, sq, j == ti[u], acc[u]);
        }
    }
    if (jb < j1) {
#pragma unroll
        for (int jj = 0; jj < kP2PBlock64; ++jj) {
            const int j = jb + jj;
            const double sx = sp[j], sy = sp[n + j], sz = sp[2 * n + j], sq = sp[3 * n + j];
            const bool past = j >= j1;
#pragma unroll
            for (int u = 0; u < 2; ++u) pair_f64<true>(t[u], sx, sy, sz, sq, past || (SELF && j == ti[u]), acc[u]);
        }
    }
}

__global__ __launch_bounds__(256) void k_p2p_f64_big(const int2* __restrict__ items, const int32_t* __restrict__ nitems,
                                                    const int32_t* __restrict__ lstart,
                                                    const int32_t* __restrict__ slots, const double* __restrict__ sp,
                                                    int64_t n, double* __restrict__ res) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int total = *nitems;
    for (int it = blockIdx.x * 4 + wave; it < total; it += gridDim.x * 4) {
        const int2 item = items[it];
        const int leaf = __builtin_amdgcn_readfirstlane(item.x);
        const int first = __builtin_amdgcn_readfirstlane(item.y);
        const int tend = min(lstart[leaf + 1], first + kP2PItem);
        int ti[2];
        double t[2][3], acc[2][4];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            ti[u] = first + 64 * u + lane;
            const int tt = ti[u] < tend ? ti[u] : first;
            for (int d = 0; d < 3; ++d) t[u][d] = sp[d * n + tt];
            for (int k = 0; k < 4; ++k) acc[u][k] = 0.;
        }
        for (int s = -1; s < kNbrSlots; ++s) {
            const int nb =
                __builtin_amdgcn_readfirstlane(s < 0 ? leaf : slots[static_cast<int64_t>(leaf) * kNbrSlots + s]);
            if (nb < 0) continue;
            const int j0 = __builtin_amdgcn_readfirstlane(lstart[nb]);
            const int j1 = __builtin_amdgcn_readfirstlane(lstart[nb + 1]);
            if (s < 0)
                p2p_f64_leaf<true>(sp, n, j0, j1, t, ti, acc);
            else
                p2p_f64_leaf<false>(sp, n, j0, j1, t, ti, acc);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (ti[u] >= tend) continue;
            res[ti[u]] = acc[u][0];
            res[n + ti[u]] = acc[u][1];
            res[2 * n + ti[u]] = acc[u][2];
            res[3 * n + ti[u]] = acc[u][3];
        }
    }
}

// ---------------------------------------------------------------------------
// M2L on FP64 MFMA (v_mfma_f64_16x16x4f64), all levels in one launch.
//
// A workgroup of 4 waves owns a tile of 32 targets (two 16-column MFMA tiles)
// of one parity at one level, and walks the 189 far-field slots of that
// parity.  For slot j (relative position pos, canonical operator c with rank
// k, node permutation pv):
//   GEMM1  C (kpad x 32) = V_c[pv[n]][r]^T * M_src[n]   work units (rank tile,
//          column tile) spread over the waves; B read as 32-byte vectors
//   GEMM2  acc (nn x 32) += U_c[pv[n]][r] * C          row tiles spread over
//          the waves; A read as 32-byte vectors along r
// The permutations are row gathers of the L2-resident operators; C passes
// through LDS (double buffered, one barrier per slot); every target belongs to
// exactly one tile, so L is written without atomics.  Node indices are padded
// to LD (multiple of 16): padded permutation entries point at a zero operator
// row and the multipole padding is zero, so no lane needs a bounds mask.
// The eight parities run on the eight XCDs side by side, so the XCDs read the
// same source multipoles at about the same time.
// MFMA 16x16x4 f64 operand layout (checked by hfmm_selftest_mfma_f64):
// A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, and register r of lane l
// holds D[4r + l/16][l%16].
constexpr int kM2LCols = 2;       // 16-wide column tiles per workgroup
constexpr int kM2LTile = 16 * kM2LCols;
constexpr int kMaxRankTiles = 4;  // kpad <= 64


// Per-workgroup state of an M2L tile.
template <int P>
struct M2LTileState {
    static constexpr int NN = P * P * P;
    static constexpr int NT = (NN + 15) / 16;
    static constexpr int NTW = (NT + kM2LWaves - 1) / kM2LWaves;
    d4 acc[NTW][kM2LCols];
};

// Runs the slots [i0, i1) of the tile's (rank-ordered) slot list, all with
// KT rank tiles (kpad = 16 KT), so every MFMA accumulator update is
// unconditional.  Per slot: GEMM1 over this wave's node quarter into its
// partial C tile, barrier, GEMM2 over this wave's row tiles.
template <int P, int KT>
__device__ __forceinline__ void m2l_segment(M2LTileState<P>& st, int i0, int i1, const int16_t* s_slot,
                                            const int16_t* s_pos, const int64_t* s_off, const int32_t* lslots,
                                            const int (&tgt)[kM2LCols], int (&nsrc)[kM2LCols],
                                            const double* __restrict__ Ml, const double* __restrict__ Mzero,
                                            const double* __restrict__ ops, const uint16_t* __restrict__ pvec,
                                            uint16_t* pvs, double* Cp, int& nproc, int LD, int kpadmax, int tid,
                                            int wave, int col, int kq, int kbq0, int kbq1) {
    constexpr int NN = P * P * P;
    constexpr int W = kM2LWaves;
    constexpr int NTW = M2LTileState<P>::NTW;
    constexpr int TC = kM2LTile;
    constexpr int kpad = 16 * KT;
    for (int i = i0; i < i1; ++i) {
        const double* Mrow[kM2LCols];
        bool any = false;
#pragma unroll
        for (int ct = 0; ct < kM2LCols; ++ct) {
            const int s = nsrc[ct];
            any |= s >= 0;
            // a missing source reads the zero multipole (no divergent branches)
            Mrow[ct] = (s >= 0 ? Ml + static_cast<int64_t>(s) * LD : Mzero) + 4 * kq;
            nsrc[ct] = (tgt[ct] >= 0 && i + 1 < kFarSlots)
                           ? lslots[static_cast<int64_t>(tgt[ct]) * kFarSlots + s_slot[i + 1]]
                           : -1;
        }
        if (__ballot(any) == 0) continue;  // same answer in every wave
        const int pos = s_pos[i];
        const double* U = ops + s_off[i];
        const double* V = U + static_cast<size_t>(NN + 1) * kpad;
        // alternate the permutation buffer per processed slot: a wave still in
        // the previous slot's GEMM2 reads the other buffer
        uint16_t* pv = pvs + (nproc & 1) * LD;
        ++nproc;
        {
            const ushort4* g = reinterpret_cast<const ushort4*>(pvec + static_cast<size_t>(pos) * LD);
            for (int q = tid; q < LD / 4; q += 64 * W) reinterpret_cast<ushort4*>(pv)[q] = g[q];
        }
        __syncthreads();  // pv ready; every wave is past the previous slot's GEMM2
        {
            d4 c1[KT][kM2LCols];
#pragma unroll
            for (int rt = 0; rt < KT; ++rt)
#pragma unroll
                for (int ct = 0; ct < kM2LCols; ++ct) c1[rt][ct] = d4{0., 0., 0., 0.};
            const double* Vc = V + col;
            for (int kb = kbq0; kb < kbq1; ++kb) {
                const d4 b0 = *reinterpret_cast<const d4*>(Mrow[0] + 16 * kb);
                const d4 b1 = *reinterpret_cast<const d4*>(Mrow[1] + 16 * kb);
                const ushort4 rows = *reinterpret_cast<const ushort4*>(pv + 16 * kb + 4 * kq);
                const double* v0 = Vc + static_cast<int>(rows.x) * kpad;
                const double* v1 = Vc + static_cast<int>(rows.y) * kpad;
                const double* v2 = Vc + static_cast<int>(rows.z) * kpad;
                const double* v3 = Vc + static_cast<int>(rows.w) * kpad;
                double a[KT][4];
#pragma unroll
                for (int rt = 0; rt < KT; ++rt) {
                    a[rt][0] = v0[16 * rt];
                    a[rt][1] = v1[16 * rt];
                    a[rt][2] = v2[16 * rt];
                    a[rt][3] = v3[16 * rt];
                }
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                    for (int rt = 0; rt < KT; ++rt) {
                        c1[rt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][s2], b0[s2], c1[rt][0], 0, 0, 0);
                        c1[rt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][s2], b1[s2], c1[rt][1], 0, 0, 0);
                    }
            }
            // MFMA step s used node 16 kb + 4 kq + s in lane group kq: a
            // permutation of the K order shared by A and B.
            double* Cw = Cp + static_cast<size_t>(wave) * kpadmax * TC;
#pragma unroll
            for (int rt = 0; rt < KT; ++rt)
#pragma unroll
                for (int ct = 0; ct < kM2LCols; ++ct)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Cw[(16 * rt + 4 * r + kq) * TC + 16 * ct + col] = c1[rt][ct][r];
        }
        __syncthreads();
        // GEMM2: row tiles nt = wave + W t for all t < NTW (a wave with fewer
        // real tiles computes throw-away ones on row 0 at no cost: the barrier
        // waits for the busiest wave anyway); B = sum of the W partial C tiles.
        int urow[NTW];
#pragma unroll
        for (int t = 0; t < NTW; ++t) {
            const int row = 16 * (wave + W * t) + col;
            urow[t] = row < LD ? static_cast<int>(pv[row]) * kpad : 0;
        }
#pragma unroll
        for (int kb = 0; kb < KT; ++kb) {
            const int r0 = 16 * kb + 4 * kq;
            d4 av[NTW];
#pragma unroll
            for (int t = 0; t < NTW; ++t) av[t] = *reinterpret_cast<const d4*>(U + urow[t] + r0);
            double b[4][kM2LCols];
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int ct = 0; ct < kM2LCols; ++ct) {
                    const int idx = (r0 + s2) * TC + 16 * ct + col;
                    double v = Cp[idx];
#pragma unroll
                    for (int w = 1; w < W; ++w) v += Cp[static_cast<size_t>(w) * kpadmax * TC + idx];
                    b[s2][ct] = v;
                }
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int t = 0; t < NTW; ++t)
#pragma unroll
                    for (int ct = 0; ct < kM2LCols; ++ct)
                        st.acc[t][ct] =
                            __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][s2], b[s2][ct], st.acc[t][ct], 0, 0, 0);
        }
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_m2l_f64(M2LLevels lv, const int32_t* __restrict__ slots,
                                                const int32_t* __restrict__ order, const int32_t* __restrict__ pcount,
                                                const int16_t* __restrict__ slotord, const int32_t* __restrict__ slotseg,
                                                const uint32_t* __restrict__ pidx, const uint16_t* __restrict__ pvec,
                                                const int64_t* __restrict__ opoff, const double* __restrict__ ops,
                                                const double* __restrict__ M, const double* __restrict__ Mzero,
                                                double* __restrict__ L, int LD, int kpadmax) {
    constexpr int NN = P * P * P;
    constexpr int W = kM2LWaves;
    constexpr int NT = M2LTileState<P>::NT;
    constexpr int NTW = M2LTileState<P>::NTW;
    constexpr int TC = kM2LTile;
    // LDS: rank-ordered slot tables, two node-permutation buffers, W partial C tiles
    __shared__ int16_t s_slot[kFarSlots];
    __shared__ int16_t s_pos[kFarSlots];
    __shared__ int64_t s_off[kFarSlots];
    extern __shared__ double smem[];  // C[W][kpadmax][TC] then pv[2][LD] (uint16)
    double* Cp = smem;
    uint16_t* pvs = reinterpret_cast<uint16_t*>(smem + static_cast<size_t>(W) * kpadmax * TC);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave
    const int col = lane & 15, kq = lane >> 4;
    const int KB = LD / 16;  // 16-node blocks
    // this wave's quarter of the node range in GEMM1
    const int kbq0 = (KB * wave) / W, kbq1 = (KB * (wave + 1)) / W;
    // this wave's real row tiles in GEMM2: nt = wave + W t, t < ntw
    const int ntw = (NT - wave + W - 1) / W;

    // Workgroup i runs on XCD i % 8.  Parity e = i % 8 and the tile index
    // within the parity = i / 8, so the eight XCDs walk the same parents (one
    // child index each) at the same time and share the source multipoles
    // through the memory-side cache.
    const int bid = blockIdx.x;
    if (bid >= lv.tile_begin[lv.nlev]) return;
    int li = 0;
    while (bid >= lv.tile_begin[li + 1]) ++li;
    const int lb = bid - lv.tile_begin[li];
    const int e = lb % 8;
    const int ptile = lb / 8;
    const int32_t* pc = pcount + lv.pcount_off[li];
    int first = 0;
    for (int q = 0; q < e; ++q) first += pc[q];
    const int cnt = min(TC, pc[e] - ptile * TC);
    if (cnt <= 0) return;  // uniform over the workgroup
    first += ptile * TC;
    for (int i = tid; i < kFarSlots; i += 64 * W) {
        const int j = slotord[(e * kFarSlots + i) * 2];
        const int pos = slotord[(e * kFarSlots + i) * 2 + 1];
        s_slot[i] = static_cast<int16_t>(j);
        s_pos[i] = static_cast<int16_t>(pos);
        s_off[i] = opoff[pidx[pos]];
    }
    int seg[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) seg[q] = slotseg[e * 5 + q];
    const int64_t coff = lv.cell_off[li];
    const int32_t* lslots = slots + coff * kFarSlots;
    const double* Ml = M + coff * LD;
    double* Ll = L + coff * LD;
    int tgt[kM2LCols];
#pragma unroll
    for (int ct = 0; ct < kM2LCols; ++ct) {
        const int cc = 16 * ct + col;
        tgt[ct] = cc < cnt ? order[coff + first + cc] : -1;
    }
    __syncthreads();
    int nsrc[kM2LCols];
#pragma unroll
    for (int ct = 0; ct < kM2LCols; ++ct)
        nsrc[ct] = tgt[ct] >= 0 ? lslots[static_cast<int64_t>(tgt[ct]) * kFarSlots + s_slot[0]] : -1;

    M2LTileState<P> st;
#pragma unroll
    for (int t = 0; t < NTW; ++t)
#pragma unroll
        for (int ct = 0; ct < kM2LCols; ++ct) st.acc[t][ct] = d4{0., 0., 0., 0.};
    int nproc = 0;
    m2l_segment<P, 1>(st, seg[0], seg[1], s_slot, s_pos, s_off, lslots, tgt, nsrc, Ml, Mzero, ops, pvec, pvs, Cp,
                      nproc, LD, kpadmax, tid, wave, col, kq, kbq0, kbq1);
    m2l_segment<P, 2>(st, seg[1], seg[2], s_slot, s_pos, s_off, lslots, tgt, nsrc, Ml, Mzero, ops, pvec, pvs, Cp,
                      nproc, LD, kpadmax, tid, wave, col, kq, kbq0, kbq1);
    m2l_segment<P, 3>(st, seg[2], seg[3], s_slot, s_pos, s_off, lslots, tgt, nsrc, Ml, Mzero, ops, pvec, pvs, Cp,
                      nproc, LD, kpadmax, tid, wave, col, kq, kbq0, kbq1);
    m2l_segment<P, 4>(st, seg[3], seg[4], s_slot, s_pos, s_off, lslots, tgt, nsrc, Ml, Mzero, ops, pvec, pvs, Cp,
                      nproc, LD, kpadmax, tid, wave, col, kq, kbq0, kbq1);
    const double scale = lv.scale[li];
#pragma unroll
    for (int ct = 0; ct < kM2LCols; ++ct) {
        if (tgt[ct] < 0) continue;
#pragma unroll
        for (int t = 0; t < NTW; ++t) {
            if (t < ntw) {  // real row tiles of this wave
                const int nt = wave + W * t;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * nt + 4 * r + kq;
                    if (row < NN) Ll[static_cast<int64_t>(tgt[ct]) * LD + row] = scale * st.acc[t][ct][r];
                }
            }
        }
    }
}

__global__ void k_mfma_probe(double* out) {
    const int l = threadIdx.x;
    const double a = 1.0 + (l % 16) * 4 + (l / 16);    // A[l%16][l/16]
    const double b = 0.5 + (l / 16) * 16 + (l % 16);    // B[l/16][l%16]
    d4 c = {0., 0., 0., 0.};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) out[(4 * i + l / 16) * 16 + (l % 16)] = c[i];
}

// Sorted -> input order.  The passes accumulate the potential and the field
// grad(phi); the force is q grad(phi) (the reference's convention,
// FP2P.hpp:80-101: F_t += q_t q_s grad_t(1/r)), the field is kept as well.
// One thread per input index o reads its sorted index i = iperm[o]: the seven
// outputs are written contiguously and the five inputs gathered.
__global__ __launch_bounds__(256) void k_scatter(int64_t n, const int32_t* __restrict__ iperm,
                                                const double* __restrict__ sp, const double* __restrict__ rs,
                                                double* __restrict__ r, double* __restrict__ field) {
    const int64_t o = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int64_t i = iperm[o];
    const double q = sp[3 * n + i];
    r[o] = rs[i];
    for (int d = 1; d < 4; ++d) {
        const double g = rs[d * n + i];
        r[d * n + o] = g * q;
        field[(d - 1) * n + o] = g;
    }
}

inline unsigned blocks256(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Grid-stride fill with zeros (a kernel node, where hipMemsetAsync would be a
// memset node in a captured graph).
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* __restrict__ p, int64_t words) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0u;
}

// Grid-stride copy (ADD = false) or sum (ADD = true) of count doubles.
template <bool ADD>
__global__ __launch_bounds__(256) void k_combine(double* __restrict__ dst, const double* __restrict__ src,
                                                 int64_t count) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = ADD ? dst[i] + src[i] : src[i];
}

}  // namespace

Box make_box(const Context* c) {
    Box b;
    const double half = c->P.box_width / 2;
    for (int d = 0; d < 3; ++d) b.corner[d] = c->P.box_center[d] + (-half);
    const int Lloc = c->P.tree_height - 1;
    b.leafw = c->P.box_width / static_cast<double>(1 << Lloc);
    b.cmask = (1u << Lloc) - 1u;
    return b;
}

M2LLevels make_levels(const Context* c, int tile_targets, int& tiles) {
    const int L = c->height() - 1;
    M2LLevels lv{};
    lv.nlev = 0;
    tiles = 0;
    lv.tile_begin[0] = 0;
    // the M2L levels of the working range (every level top..h-1 by default)
    const int lo = c->work_upper > c->top() ? c->work_upper : c->top();
    const int hi = c->work_lower > 0 && c->work_lower - 1 < L ? c->work_lower - 1 : L;
    for (int l = lo; l <= hi; ++l) {
        const int i = lv.nlev++;
        lv.cell_off[i] = c->lvl_off[l];
        lv.pcount_off[i] = 8 * l;
        // 2^l cells per axis of a set: scale factor 2 / (width / 2^l) (FInterpMatrixKernel.hpp:124-133)
        const int lloc = l - c->set_levels;
        lv.scale[i] = 2. / (c->P.box_width / static_cast<double>(int64_t(1) << lloc));
        // segment of 8 x (tiles of the largest parity); the parity counts are
        // on the device, ceil(ncell/tile) bounds every parity's tile count
        tiles += 8 * static_cast<int>((c->ncell[l] + tile_targets - 1) / tile_targets);
        lv.tile_begin[i + 1] = tiles;
    }
    return lv;
}

namespace {

template <int P>
int p2m_impl(Context* c) {
    const int L = c->height() - 1;
    const int64_t nl = c->ncell[L];
    prof_begin(c, "p2m");
    auto kern = k_p2m<P, 0>;
    if constexpr (P <= kUniformMaxOrder)
        if (c->basis == 1) kern = k_p2m<P, 1>;
    (void)nl;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>((c->n + 63) / 64)), dim3(64), 0, c->stream, c->n,
                       c->d_pleaf.as<int32_t>(), c->d_cellkey.as<uint64_t>() + c->lvl_off[L],
                       c->d_leafstart.as<int32_t>(), c->d_sorted.as<double>(), make_box(c), c->d_Troots.as<double>(),
                       c->d_uroots.as<double>(), c->d_M.as<double>() + static_cast<size_t>(c->lvl_off[L]) * c->LD,
                       c->LD);
    prof_end(c);
    return check_launch(c, "p2m");
}

template <int P>
int m2m_impl(Context* c, int l) {
    const int64_t np = c->ncell[l];
    if (np == 0) return HFMM_OK;
    prof_begin(c, "m2m");
    hipLaunchKernelGGL(k_m2m<P>, dim3(static_cast<unsigned>((np + 3) / 4)), dim3(256), 0, c->stream,
                       static_cast<int>(np), c->d_childstart.as<int32_t>() + c->cs_off[l],
                       c->d_cellkey.as<uint64_t>() + c->lvl_off[l + 1],
                       c->d_child.as<double>(), c->d_M.as<double>() + static_cast<size_t>(c->lvl_off[l + 1]) * c->LD,
                       c->d_M.as<double>() + static_cast<size_t>(c->lvl_off[l]) * c->LD, c->LD);
    prof_end(c);
    return check_launch(c, "m2m");
}

template <int P>
int l2l_impl(Context* c, int l) {
    const int64_t np = c->ncell[l];
    if (np == 0) return HFMM_OK;
    prof_begin(c, "l2l");
    hipLaunchKernelGGL(k_l2l<P>, dim3(static_cast<unsigned>((np + 3) / 4)), dim3(256), 0, c->stream,
                       static_cast<int>(np), c->d_childstart.as<int32_t>() + c->cs_off[l],
                       c->d_cellkey.as<uint64_t>() + c->lvl_off[l + 1],
                       c->d_child.as<double>(), c->d_L.as<double>() + static_cast<size_t>(c->lvl_off[l]) * c->LD,
                       c->d_L.as<double>() + static_cast<size_t>(c->lvl_off[l + 1]) * c->LD, c->LD);
    prof_end(c);
    return check_launch(c, "l2l");
}

template <int P>
int l2p_impl(Context* c) {
    const int L = c->height() - 1;
    prof_begin(c, "l2p");
    if constexpr (P <= kUniformMaxOrder) {
        if (c->basis == 1) {
            hipLaunchKernelGGL(k_l2p_lagrange<P>, dim3(blocks256(c->n)), dim3(256), 0, c->stream, c->n,
                               c->d_pleaf.as<int32_t>(), c->d_cellkey.as<uint64_t>() + c->lvl_off[L],
                               c->d_sorted.as<double>(), make_box(c), c->d_uroots.as<double>(),
                               c->d_L.as<double>() + static_cast<size_t>(c->lvl_off[L]) * c->LD, c->LD,
                               sorted_results(c));
            prof_end(c);
            return check_launch(c, "l2p");
        }
    }
    hipLaunchKernelGGL(k_l2p<P>, dim3(blocks256(c->n)), dim3(256), 0, c->stream, c->n, c->d_pleaf.as<int32_t>(),
                       c->d_cellkey.as<uint64_t>() + c->lvl_off[L], c->d_sorted.as<double>(), make_box(c),
                       c->d_Troots.as<double>(), c->d_L.as<double>() + static_cast<size_t>(c->lvl_off[L]) * c->LD,
                       c->LD, sorted_results(c));
    prof_end(c);
    return check_launch(c, "l2p");
}

template <int P>
int m2l_impl(Context* c) {
    int tiles = 0;
    const M2LLevels lv = make_levels(c, kM2LTile, tiles);
    if (tiles == 0) return HFMM_OK;  // no M2L level in the working range
    const unsigned grid = static_cast<unsigned>(tiles);
    const size_t shmem =
        sizeof(double) * kM2LWaves * static_cast<size_t>(c->kpad_max) * kM2LTile + sizeof(uint16_t) * 2 * c->LD;
    static int lds_limit[64] = {0};
    HFMM_TRY(raise_lds_limit(c, reinterpret_cast<const void*>(&k_m2l_f64<P>), shmem, lds_limit));
    prof_begin(c, "m2l");
    hipLaunchKernelGGL(k_m2l_f64<P>, dim3(grid), dim3(64 * kM2LWaves), shmem, c->stream, lv,
                       c->d_m2l.as<int32_t>(), c->d_order.as<int32_t>(), c->d_pcount.as<int32_t>(),
                       c->d_slotord.as<int16_t>(), c->d_slotseg.as<int32_t>(), c->d_pidx.as<uint32_t>(),
                       c->d_pvec.as<uint16_t>(), c->d_opoff.as<int64_t>(), c->d_ops.as<double>(), c->d_M.as<double>(),
                       c->d_M.as<double>() + static_cast<size_t>(c->total_cells) * c->LD, c->d_L.as<double>(), c->LD,
                       c->kpad_max);
    prof_end(c);
    return check_launch(c, "m2l");
}

#define HFMM_DISPATCH(P_, CALL)                                                   \
    switch (P_) {                                                                 \
        case 5: return CALL(5);                                                   \
        case 6: return CALL(6);                                                   \
        case 7: return CALL(7);                                                   \
        case 8: return CALL(8);                                                   \
        default: return fail(c, HFMM_ERR_ARG, "order not compiled (5..8)");      \
    }

}  // namespace

int run_p2m(Context* c) {
#define CALL(P) p2m_impl<P>(c)
    HFMM_DISPATCH(c->order(), CALL)
#undef CALL
}
int run_m2m(Context* c, int l) {
#define CALL(P) m2m_impl<P>(c, l)
    HFMM_DISPATCH(c->order(), CALL)
#undef CALL
}
int run_l2l(Context* c, int l) {
#define CALL(P) l2l_impl<P>(c, l)
    HFMM_DISPATCH(c->order(), CALL)
#undef CALL
}
int run_l2p(Context* c) {
#define CALL(P) l2p_impl<P>(c)
    HFMM_DISPATCH(c->order(), CALL)
#undef CALL
}
int run_m2l(Context* c) {
    if (c->kpad_max > 16 * kMaxRankTiles) return fail(c, HFMM_ERR_ARG, "operator rank above 64 is not supported");
#define CALL(P) m2l_impl<P>(c)
    HFMM_DISPATCH(c->order(), CALL)
#undef CALL
}

int run_p2p(Context* c) {
    if (!c->p2p_small_done) {
        prof_begin(c, "p2p");
        hipLaunchKernelGGL(k_p2p_f64, dim3(blocks256(c->n)), dim3(256), 0, c->stream, c->n, c->d_pleaf.as<int32_t>(),
                           c->d_leafstart.as<int32_t>(), c->d_p2p.as<int32_t>(), c->d_sorted.as<double>(),
                           sorted_results(c));
        prof_end(c);
        HFMM_TRY(check_launch(c, "p2p"));
    }
    prof_begin(c, "p2p_big");
    hipLaunchKernelGGL(k_p2p_f64_big, dim3((p2p_tiled_grid(c) + 3) / 4), dim3(256), 0, c->stream,
                       c->d_p2pitems.as<int2>(), p2p_item_count(c), c->d_leafstart.as<int32_t>(),
                       c->d_p2p.as<int32_t>(), c->d_sorted.as<double>(), c->n, sorted_results(c));
    prof_end(c);
    return check_launch(c, "p2p_big");
}

int zero_device(Context* c, void* ptr, size_t bytes) {
    if (bytes % 4) return fail(c, HFMM_ERR_ARG, "zero_device: size not a multiple of 4");
    const int64_t words = static_cast<int64_t>(bytes / 4);
    if (words == 0) return HFMM_OK;
    const int64_t blocks = std::min<int64_t>((words + 1023) / 1024, 65536);
    hipLaunchKernelGGL(k_zero_words, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, c->stream,
                       static_cast<uint32_t*>(ptr), words);
    return check_launch(c, "zero");
}

int combine_device(Context* c, double* dst, const double* src, int64_t count, bool add) {
    if (count <= 0) return HFMM_OK;
    const int64_t blocks = std::min<int64_t>((count + 1023) / 1024, 65536);
    if (add)
        hipLaunchKernelGGL(k_combine<true>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, c->stream, dst, src,
                           count);
    else
        hipLaunchKernelGGL(k_combine<false>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, c->stream, dst, src,
                           count);
    return check_launch(c, "combine");
}

int run_p2p_small_all(Context* c) {
    HFMM_TRY(gather_all_charges(c, false));
    RhsChunk ch[HFMM_MAX_RHS];
    const int nch = rhs_chunks(c->nrhs, ch);
    const int64_t npad = (c->n + 7) & ~int64_t(7);
    prof_begin(c, "p2p_rhs");
    for (int k = 0; k < nch; ++k) {
        const double* qc = c->d_qall.as<double>() + npad * ch[k].r0;
        double* res = c->d_res_sorted.as<double>() + static_cast<size_t>(ch[k].r0) * 4 * c->n;
        const dim3 grid(blocks256(c->n)), block(256);
        const int32_t* pl = c->d_pleaf.as<int32_t>();
        const int32_t* ls = c->d_leafstart.as<int32_t>();
        const int32_t* sl = c->d_p2p.as<int32_t>();
        const double* sp = c->d_sorted.as<double>();
        switch (ch[k].R) {
            case 8: hipLaunchKernelGGL(k_p2p_f64_rhs<8>, grid, block, 0, c->stream, c->n, pl, ls, sl, sp, qc, res); break;
            case 4: hipLaunchKernelGGL(k_p2p_f64_rhs<4>, grid, block, 0, c->stream, c->n, pl, ls, sl, sp, qc, res); break;
            case 2: hipLaunchKernelGGL(k_p2p_f64_rhs<2>, grid, block, 0, c->stream, c->n, pl, ls, sl, sp, qc, res); break;
            default: hipLaunchKernelGGL(k_p2p_f64_rhs<1>, grid, block, 0, c->stream, c->n, pl, ls, sl, sp, qc, res); break;
        }
    }
    prof_end(c);
    return check_launch(c, "p2p_rhs");
}

int zero_results(Context* c) { return zero_device(c, sorted_results(c), sizeof(double) * 4 * c->n); }

int scatter_results(Context* c, int rhs) {
    prof_begin(c, "scatter");
    hipLaunchKernelGGL(k_scatter, dim3(blocks256(c->n)), dim3(256), 0, c->stream, c->n, c->d_iperm.as<int32_t>(),
                       c->d_sorted.as<double>(), sorted_results(c),
                       c->d_res.as<double>() + static_cast<size_t>(rhs) * 4 * c->n,
                       c->d_field.as<double>() + static_cast<size_t>(rhs) * 3 * c->n);
    prof_end(c);
    return check_launch(c, "scatter");
}

int selftest_mfma_f64(Context* c, double* out) {
    DevBuf b;
    HFMM_TRY(ensure(c, b, sizeof(double) * 256));
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, c->stream, b.as<double>());
    int st = check_launch(c, "mfma_probe");
    if (st == HFMM_OK) {
        if (hipMemcpyAsync(out, b.ptr, sizeof(double) * 256, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = fail(c, HFMM_ERR_DEVICE, "probe copy failed");
    }
    release(b);
    return st;
}

}  // namespace hfmm
