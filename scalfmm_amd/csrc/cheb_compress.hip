// cheb_compress.hip -- FP64 compression of the 16 canonical M2L operators on
// the device, one workgroup per operator.
//
// Follows (paths relative to the reference's source tree)
//   Src/Kernels/Chebyshev/FChebSymM2LHandler.hpp:39-281  (precompute)
//   Src/Utils/FAca.hpp:141-226                            (partially pivoted ACA)
//   Src/Utils/FSvd.hpp:51-62                              (rank rule)
// For the 16 canonical positions (i in {2,3}, 0 <= j <= i, 0 <= k <= j) of a
// source cell of width 2 at (2i, 2j, 2k) from the target cell at the origin:
//   1. pACA(eps) of the weighted block w_x w_y / |y - x| -> U, V (rank r);
//   2. Householder QR of U and V, phi = R_U R_V^T, one-sided Jacobi SVD of phi;
//   3. k = the smallest rank whose tail energy is <= eps^2 of the total;
//   4. K = [Q_U U_phi S_k | Q_V V_phi] with the rows divided by w (un-weighting).
// The reference calls LAPACK for step 2; these kernels use their own QR and
// SVD.  The pACA's dot products and accumulations run sequentially in one
// thread in index order and the file is built without FMA contraction, so its
// pivots (chosen by exact comparisons, ties included) do not depend on a
// reduction order; the Jacobi SVD's column dot products are 64 lane-strided
// partial sums combined by a fixed xor butterfly, the same bits on every run.
// Parallelism is over lanes, independent columns, entries and the 16
// operators.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

constexpr int kCanon = 16;
constexpr int kCT = 256;  // threads per workgroup

__device__ __forceinline__ double inv_r(const double* t, const double* s) {
    const double dx = t[0] - s[0], dy = t[1] - s[1], dz = t[2] - s[2];
    return 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ double dotp(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// Index of the first maximum of |x[i]| over the free entries, as a left-to-right
// scan with a strict comparison finds it (-1 when every free entry is 0).
__device__ int block_first_max(const double* x, const unsigned char* freef, int n, double* sb, int* si) {
    const int tid = threadIdx.x;
    double best = 0.;
    int idx = -1;
    for (int i = tid; i < n; i += kCT) {
        const double m = fabs(x[i]);
        if (freef[i] && best < m) {
            best = m;
            idx = i;
        }
    }
    sb[tid] = best;
    si[tid] = idx;
    __syncthreads();
    for (int w = kCT / 2; w > 0; w >>= 1) {
        if (tid < w) {
            const double b2 = sb[tid + w];
            const int i2 = si[tid + w];
            if (i2 >= 0 && (si[tid] < 0 || b2 > sb[tid] || (b2 == sb[tid] && i2 < si[tid]))) {
                sb[tid] = b2;
                si[tid] = i2;
            }
        }
        __syncthreads();
    }
    const int r = si[0];
    __syncthreads();
    return r;
}

// Step 1: partially pivoted ACA with the reference's pivoting (first maximum,
// strict comparison) and stopping rule |u_k|^2 |v_k|^2 <= eps^2 |S_k|_F^2.
// U, V: column l at offset l * nn; rank[op] = r, or -1 without convergence.
__global__ __launch_bounds__(kCT) void k_aca(int nn, int maxk, double eps, const double* __restrict__ X,
                                            const double* __restrict__ Ys, const double* __restrict__ w,
                                            double* __restrict__ Us, double* __restrict__ Vs,
                                            int* __restrict__ rank) {
    const int op = blockIdx.x, tid = threadIdx.x;
    const double* Y = Ys + static_cast<size_t>(op) * nn * 3;
    double* U = Us + static_cast<size_t>(op) * nn * maxk;
    double* V = Vs + static_cast<size_t>(op) * nn * maxk;
    __shared__ unsigned char rfree[512], cfree[512];
    __shared__ double sb[kCT];
    __shared__ int si[kCT];
    __shared__ double sprod[512];
    __shared__ double suu, svv, snorm2S, spiv;
    __shared__ int sI, sJ, scont;
    for (int i = tid; i < nn; i += kCT) rfree[i] = cfree[i] = 1;
    if (tid == 0) {
        sI = 0;
        sJ = 0;
        snorm2S = 0.;
    }
    __syncthreads();
    int k = 0;
    for (;;) {
        if (k >= maxk) {
            if (tid == 0) rank[op] = -1;
            return;
        }
        const int I = sI;
        double* u = U + static_cast<size_t>(k) * nn;
        double* v = V + static_cast<size_t>(k) * nn;
        // residual row I
        for (int y = tid; y < nn; y += kCT) {
            double val = w[y] * w[I] * inv_r(Y + 3 * y, X + 3 * I);
            for (int l = 0; l < k; ++l) {
                const double a = -1. * U[static_cast<size_t>(l) * nn + I];
                val += a * V[static_cast<size_t>(l) * nn + y];
            }
            v[y] = val;
        }
        if (tid == 0) rfree[I] = 0;
        __syncthreads();
        const int Jn = block_first_max(v, cfree, nn, sb, si);
        if (tid == 0) {
            if (Jn >= 0) sJ = Jn;
            spiv = 1. / v[sJ];
        }
        __syncthreads();
        const int J = sJ;
        for (int y = tid; y < nn; y += kCT) v[y] *= spiv;
        // residual column J
        for (int x = tid; x < nn; x += kCT) {
            double val = w[J] * w[x] * inv_r(Y + 3 * J, X + 3 * x);
            for (int l = 0; l < k; ++l) {
                const double a = -1. * V[static_cast<size_t>(l) * nn + J];
                val += a * U[static_cast<size_t>(l) * nn + x];
            }
            u[x] = val;
        }
        if (tid == 0) cfree[J] = 0;
        __syncthreads();
        const int In = block_first_max(u, rfree, nn, sb, si);
        if (tid == 0 && In >= 0) sI = In;
        // |S_k|^2 update: cross terms with every earlier column pair
        for (int l = tid; l < k; l += kCT)
            sprod[l] = dotp(nn, u, U + static_cast<size_t>(l) * nn) * dotp(nn, V + static_cast<size_t>(l) * nn, v);
        if (tid == kCT - 1) suu = dotp(nn, u, u);
        if (tid == kCT - 2) svv = dotp(nn, v, v);
        __syncthreads();
        if (tid == 0) {
            double cross = 0.;
            for (int l = 0; l < k; ++l) cross += sprod[l];
            const double norm2uv = suu * svv;
            snorm2S += norm2uv + 2 * cross;
            scont = norm2uv > eps * eps * snorm2S;
        }
        ++k;
        __syncthreads();
        if (!scont) break;
    }
    if (tid == 0) rank[op] = k;
}

// Thin Householder QR in place: A (m x n, column-major) becomes the reflectors
// below the diagonal and R on and above it; Q (m x n) is formed explicitly and
// R copied out (n x n).
// Sum over the 64 lanes by a fixed xor butterfly: every lane ends with the
// same bits, the same on every run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// s + sum_{i = i0}^{m-1} x[i] y[i] over one wave (lane-strided partial sums).
__device__ __forceinline__ double wave_dot_from(double s, int i0, int m, const double* x, const double* y, int lane) {
    double v = 0.;
    for (int i = i0 + lane; i < m; i += 64) v += x[i] * y[i];
    return s + wave_sum(v);
}

__device__ void block_qr(int m, int n, double* A, double* tau, double* Q, double* R, double* sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int kWaves = kCT / 64;
    for (int j = 0; j < n; ++j) {
        double* cj = A + static_cast<size_t>(j) * m;
        if (wave == 0) {
            const double norm = sqrt(wave_dot_from(0., j, m, cj, cj, lane));
            if (lane == 0) {
                if (norm == 0.0) {
                    sh[0] = 0.;
                    tau[j] = 0.;
                } else {
                    const double beta = cj[j] >= 0 ? -norm : norm;
                    const double head = cj[j] - beta;
                    sh[0] = 1.;
                    sh[1] = beta;
                    sh[2] = head;
                    tau[j] = -head / beta;
                }
            }
        }
        __syncthreads();
        if (sh[0] != 0.) {
            const double head = sh[2];
            for (int i = j + 1 + tid; i < m; i += kCT) cj[i] /= head;
            __syncthreads();
            if (tid == 0) cj[j] = sh[1];
            const double tj = tau[j];
            for (int c = j + 1 + wave; c < n; c += kWaves) {  // one wave per column
                double* cc = A + static_cast<size_t>(c) * m;
                const double s = wave_dot_from(cc[j], j + 1, m, cj, cc, lane) * tj;
                if (lane == 0) cc[j] -= s;
                for (int i = j + 1 + lane; i < m; i += 64) cc[i] -= s * cj[i];
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += kCT) {
        const int c = e / n, i = e % n;
        R[e] = i <= c ? A[static_cast<size_t>(c) * m + i] : 0.0;
    }
    for (int e = tid; e < m * n; e += kCT) {
        const int c = e / m, i = e % m;
        Q[e] = i == c ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int j = n - 1; j >= 0; --j) {
        const double tj = tau[j];
        if (tj != 0.0) {
            const double* v = A + static_cast<size_t>(j) * m;
            for (int c = wave; c < n; c += kWaves) {  // one wave per column
                double* qc = Q + static_cast<size_t>(c) * m;
                const double s = wave_dot_from(qc[j], j + 1, m, v, qc, lane) * tj;
                if (lane == 0) qc[j] -= s;
                for (int i = j + 1 + lane; i < m; i += 64) qc[i] -= s * v[i];
            }
        }
        __syncthreads();
    }
}

struct Scratch {
    double *U, *V, *QU, *QV, *RU, *RV, *tau, *phi, *W, *Vw, *Uphi, *Vphi, *S, *norms;
    int* order;
};

// Steps 2-4 for one operator per workgroup.  status[op]: 0 ok, 1 SVD did not
// converge, 2 empty rank; krank[op] = the truncated rank.
__global__ __launch_bounds__(kCT) void k_recompress(int nn, int maxk, double eps, const double* __restrict__ w,
                                                   const int* __restrict__ arank, Scratch sc, size_t per_op,
                                                   size_t per_op_small, double* __restrict__ Kout,
                                                   int* __restrict__ krank, int* __restrict__ status) {
    const int op = blockIdx.x, tid = threadIdx.x;
    const int r = arank[op];
    __shared__ double sh[4];
    __shared__ double sdot[3];
    __shared__ int sflag[2];
    if (r <= 0) return;
    const size_t ob = static_cast<size_t>(op) * per_op, os = static_cast<size_t>(op) * per_op_small;
    double *U = sc.U + ob, *V = sc.V + ob, *QU = sc.QU + ob, *QV = sc.QV + ob;
    double *RU = sc.RU + os, *RV = sc.RV + os, *phi = sc.phi + os, *W = sc.W + os, *Vw = sc.Vw + os;
    double *Uphi = sc.Uphi + os, *Vphi = sc.Vphi + os;
    double* tau = sc.tau + static_cast<size_t>(op) * maxk;
    double* S = sc.S + static_cast<size_t>(op) * maxk;
    double* norms = sc.norms + static_cast<size_t>(op) * maxk;
    int* order = sc.order + static_cast<size_t>(op) * maxk;
    block_qr(nn, r, U, tau, QU, RU, sh);
    block_qr(nn, r, V, tau, QV, RV, sh);
    // phi = R_U R_V^T (column-major r x r)
    for (int e = tid; e < r * r; e += kCT) {
        const int b = e / r, a = e % r;
        double s = 0;
        for (int l = 0; l < r; ++l) s += RU[static_cast<size_t>(l) * r + a] * RV[static_cast<size_t>(l) * r + b];
        phi[e] = s;
        W[e] = s;
        Vw[e] = a == b ? 1.0 : 0.0;
    }
    __syncthreads();
    // one-sided Jacobi SVD of phi, cyclic pair order
    const int n = r;
    const double tol = 2.220446049250313e-16 * n;
    bool done = false;
    for (int sweep = 0; sweep < 100 && !done; ++sweep) {
        done = true;
        for (int a = 0; a < n - 1; ++a)
            for (int b = a + 1; b < n; ++b) {
                double* wa = W + static_cast<size_t>(a) * n;
                double* wb = W + static_cast<size_t>(b) * n;
                // waves 0, 1, 2: wa.wa, wb.wb, wa.wb as 64 lane-strided partial
                // sums and an xor butterfly (every lane ends with the same bits)
                if (tid < 192) {
                    const int wv = tid >> 6, ln = tid & 63;
                    const double* x = wv == 1 ? wb : wa;
                    const double* y = wv == 0 ? wa : wb;
                    double v = 0.;
                    for (int i = ln; i < n; i += 64) v += x[i] * y[i];
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                    if (ln == 0) sdot[wv] = v;
                }
                __syncthreads();
                const double aa = sdot[0], bb = sdot[1], ab = sdot[2];
                __syncthreads();
                if (ab == 0.0 || fabs(ab) <= tol * sqrt(aa * bb)) continue;
                done = false;
                const double zeta = (bb - aa) / (2.0 * ab);
                const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                double* va = Vw + static_cast<size_t>(a) * n;
                double* vb = Vw + static_cast<size_t>(b) * n;
                for (int i = tid; i < n; i += kCT) {
                    const double x = wa[i], y = wb[i];
                    wa[i] = c * x - s * y;
                    wb[i] = s * x + c * y;
                    const double xv = va[i], yv = vb[i];
                    va[i] = c * xv - s * yv;
                    vb[i] = s * xv + c * yv;
                }
                __syncthreads();
            }
    }
    for (int j = tid; j < n; j += kCT)
        norms[j] = sqrt(dotp(n, W + static_cast<size_t>(j) * n, W + static_cast<size_t>(j) * n));
    __syncthreads();
    if (tid == 0) {
        // stable sort of the column indices by norm, descending (insertion sort)
        for (int j = 0; j < n; ++j) order[j] = j;
        for (int j = 1; j < n; ++j) {
            const int v = order[j];
            int q = j - 1;
            while (q >= 0 && norms[order[q]] < norms[v]) {
                order[q + 1] = order[q];
                --q;
            }
            order[q + 1] = v;
        }
        for (int q = 0; q < n; ++q) S[q] = norms[order[q]];
        // smallest k with sum_{i >= k} S_i^2 <= eps^2 sum_i S_i^2 (FSvd::getRank)
        const double total = dotp(n, S, S);
        double tail = 0.;
        int k2 = -1;
        for (int q = n; q > 0; --q) {
            tail += S[q - 1] * S[q - 1];
            if (tail > eps * eps * total) {
                k2 = q;
                break;
            }
        }
        sflag[0] = k2;
        sflag[1] = done ? 0 : 1;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += kCT) {
        const int q = e / n, i = e % n;
        const int j = order[q];
        Uphi[e] = norms[j] > 0 ? W[static_cast<size_t>(j) * n + i] / norms[j] : 0.0;
        Vphi[e] = Vw[static_cast<size_t>(j) * n + i];
    }
    __syncthreads();
    const int k2 = sflag[0];
    if (tid == 0) {
        status[op] = sflag[1] ? 1 : (k2 <= 0 ? 2 : 0);
        krank[op] = k2;
    }
    if (sflag[1] || k2 <= 0) return;
    double* K = Kout + static_cast<size_t>(op) * 2 * nn * maxk;
    for (int e = tid; e < k2 * nn; e += kCT) {
        const int c = e / nn, m = e % nn;
        double su = 0, sv = 0;
        for (int a = 0; a < r; ++a) {
            su += QU[static_cast<size_t>(a) * nn + m] * (Uphi[static_cast<size_t>(c) * r + a] * S[c]);
            sv += QV[static_cast<size_t>(a) * nn + m] * Vphi[static_cast<size_t>(c) * r + a];
        }
        const double iw = 1. / w[m];
        K[e] = su * iw;
        K[static_cast<size_t>(k2) * nn + e] = sv * iw;
    }
}

}  // namespace

int compress_operators_device(Context* c) {
    ChebOperators& op = c->ops;
    const int p = op.p, nn = op.nn, maxk = nn;  // the ACA stops with an error at rank (nx + ny) / 2
    if (nn > 512) return fail(c, HFMM_ERR_ARG, "operator compression supports p^3 <= 512");
    // nodes of the target cell (centre 0, width 2) and of the 16 canonical source
    // cells, and the node weights sqrt(w1 w1 w1) (FChebTensor.hpp:42-55)
    const double cw = 2.0;
    auto nodes = [&](const double ctr[3], double* out) {
        for (int n = 0; n < nn; ++n) {
            const int id[3] = {n % p, (n / p) % p, n / (p * p)};
            for (int d = 0; d < 3; ++d) {
                const double a = ctr[d] - cw / 2., b = ctr[d] + cw / 2.;
                out[3 * n + d] = (a + b) / 2. + (b - a) * op.roots[id[d]] / 2.;
            }
        }
    };
    std::vector<double> X(3 * nn), Ys(static_cast<size_t>(kCanon) * 3 * nn), w1(p), w(nn);
    const double origin[3] = {0., 0., 0.};
    nodes(origin, X.data());
    std::vector<int> canon;
    for (int i = 2; i <= 3; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= j; ++k) {
                const double cy[3] = {cw * i, cw * j, cw * k};
                nodes(cy, Ys.data() + canon.size() * 3 * nn);
                canon.push_back((i + 3) * 49 + (j + 3) * 7 + (k + 3));
            }
    for (int o = 0; o < p; ++o) w1[o] = M_PI / p * std::sqrt(1. - op.roots[o] * op.roots[o]);
    for (int n = 0; n < nn; ++n) w[n] = std::sqrt(w1[n % p] * w1[(n / p) % p] * w1[n / (p * p)]);

    const size_t per_op = static_cast<size_t>(nn) * maxk, per_small = static_cast<size_t>(maxk) * maxk;
    DevBuf dX, dY, dw, dbig, dsmall, dvec, dint, dK;
    auto run = [&]() -> int {
        HFMM_TRY(ensure(c, dX, sizeof(double) * X.size()));
        HFMM_TRY(ensure(c, dY, sizeof(double) * Ys.size()));
        HFMM_TRY(ensure(c, dw, sizeof(double) * nn));
        HFMM_TRY(ensure(c, dbig, sizeof(double) * 4 * kCanon * per_op));    // U, V, QU, QV
        HFMM_TRY(ensure(c, dsmall, sizeof(double) * 7 * kCanon * per_small));  // RU, RV, phi, W, Vw, Uphi, Vphi
        HFMM_TRY(ensure(c, dvec, sizeof(double) * 3 * kCanon * maxk));       // tau, S, norms
        HFMM_TRY(ensure(c, dint, sizeof(int) * (kCanon * maxk + 4 * kCanon)));  // order; ranks, status
        HFMM_TRY(ensure(c, dK, sizeof(double) * kCanon * 2 * per_op));
        HFMM_HIP(c, hipMemcpy(dX.ptr, X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice));
        HFMM_HIP(c, hipMemcpy(dY.ptr, Ys.data(), sizeof(double) * Ys.size(), hipMemcpyHostToDevice));
        HFMM_HIP(c, hipMemcpy(dw.ptr, w.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
        double* big = dbig.as<double>();
        double* small = dsmall.as<double>();
        double* vec = dvec.as<double>();
        int* ints = dint.as<int>();
        int* arank = ints + kCanon * maxk;
        int* krank = arank + kCanon;
        int* status = krank + kCanon;
        Scratch sc;
        sc.U = big;
        sc.V = big + kCanon * per_op;
        sc.QU = big + 2 * kCanon * per_op;
        sc.QV = big + 3 * kCanon * per_op;
        sc.RU = small;
        sc.RV = small + kCanon * per_small;
        sc.phi = small + 2 * kCanon * per_small;
        sc.W = small + 3 * kCanon * per_small;
        sc.Vw = small + 4 * kCanon * per_small;
        sc.Uphi = small + 5 * kCanon * per_small;
        sc.Vphi = small + 6 * kCanon * per_small;
        sc.tau = vec;
        sc.S = vec + kCanon * maxk;
        sc.norms = vec + 2 * kCanon * maxk;
        sc.order = ints;
        hipLaunchKernelGGL(k_aca, dim3(kCanon), dim3(kCT), 0, c->stream, nn, maxk, op.eps, dX.as<double>(),
                           dY.as<double>(), dw.as<double>(), sc.U, sc.V, arank);
        HFMM_TRY(check_launch(c, "aca"));
        hipLaunchKernelGGL(k_recompress, dim3(kCanon), dim3(kCT), 0, c->stream, nn, maxk, op.eps, dw.as<double>(),
                           arank, sc, per_op, per_small, dK.as<double>(), krank, status);
        HFMM_TRY(check_launch(c, "recompress"));
        std::vector<int> hr(3 * kCanon);
        HFMM_HIP(c, hipMemcpyAsync(hr.data(), arank, sizeof(int) * 3 * kCanon, hipMemcpyDeviceToHost, c->stream));
        HFMM_HIP(c, hipStreamSynchronize(c->stream));
        for (int q = 0; q < kCanon; ++q) {
            if (hr[q] < 0) return fail(c, HFMM_ERR_NUMERIC, "ACA did not converge");
            if (hr[2 * kCanon + q] == 1) return fail(c, HFMM_ERR_NUMERIC, "SVD did not converge");
            if (hr[2 * kCanon + q] == 2) return fail(c, HFMM_ERR_NUMERIC, "empty rank");
        }
        for (int q = 0; q < kCanon; ++q) {
            const int k2 = hr[kCanon + q];
            std::vector<double> Kc(static_cast<size_t>(2) * k2 * nn);
            HFMM_HIP(c, hipMemcpy(Kc.data(), dK.as<double>() + static_cast<size_t>(q) * 2 * per_op,
                                  sizeof(double) * Kc.size(), hipMemcpyDeviceToHost));
            op.K[canon[q]] = std::move(Kc);
            op.rank[canon[q]] = k2;
        }
        return HFMM_OK;
    };
    const int st = run();
    for (DevBuf* b : {&dX, &dY, &dw, &dbig, &dsmall, &dvec, &dint, &dK}) release(*b);
    return st;
}

}  // namespace hfmm
