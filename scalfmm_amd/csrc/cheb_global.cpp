// cheb_global.cpp -- the M2L operators of the reference's FChebKernel: one
// global SVD of the 316 far-field operators
// (Src/Kernels/Chebyshev/FChebM2LHandler.hpp:216-315 ComputeAndCompress and
// :437-523 Compress), built on the host in FP64 once per context.
//
//   K_t[m][n] = w_m w_n / |X_m - Y_n(t)|   target roots X of the cell of width 2
//               at the origin, source roots Y at 2 (i, j, k), the 316 offsets
//               with max(|i|, |j|, |k|) > 1 in the order i, j, k = -3..3;
//               w = the square roots of the tensor Chebyshev weights
//   B = the right singular vectors of K_col (rows (t, m), columns n),
//   U = those of K_row^T (rows (t, n), columns m), both truncated to
//   k = min(getRank(sigma_col), getRank(sigma_row)) (Utils/FSvd.hpp:31-48),
//   C_t = U^T K_t B, then U and B divided by w row by row.
// The reference calls LAPACK's dgesvd on the 316 p^3 x p^3 matrices; here each
// tall matrix is reduced by a Householder QR and its triangle by one-sided
// Jacobi rotations, which gives the same singular values and vectors (up to
// the signs of the vectors, which cancel in U C_t B^T).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "hfmm_internal.hpp"

namespace hfmm {
namespace {

// Runs body(i) for i in [0, n) over the host's cores (each index on one thread;
// the results do not depend on the split).
template <class F>
void parallel_for(int n, F body) {
    const int nt = std::max(1, std::min<int>(n, static_cast<int>(std::thread::hardware_concurrency())));
    if (nt <= 1 || n < 2) {
        for (int i = 0; i < n; ++i) body(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) body(i);
        });
    for (auto& th : pool) th.join();
}

// R of the Householder QR of an m x n column-major matrix (m >= n), in place.
void householder_r(int64_t m, int n, std::vector<double>& A, std::vector<double>& R) {
    for (int j = 0; j < n; ++j) {
        double* a = A.data() + static_cast<size_t>(j) * m;
        double s = 0.;
        for (int64_t i = j; i < m; ++i) s += a[i] * a[i];
        const double norm = std::sqrt(s);
        if (norm == 0.) continue;
        const double beta = a[j] >= 0. ? -norm : norm;  // the new diagonal entry
        const double v0 = a[j] - beta;
        for (int64_t i = j + 1; i < m; ++i) a[i] /= v0;  // v = (1, a[j+1..] / v0)
        const double tau = -v0 / beta;
        a[j] = beta;
        // the trailing columns, each updated by one thread (blocks of 8 columns)
        const int nb = (n - j - 1 + 7) / 8;
        parallel_for(nb, [&](int blk) {
            const int c0 = j + 1 + 8 * blk, c1 = std::min(n, c0 + 8);
            for (int c = c0; c < c1; ++c) {
                double* b = A.data() + static_cast<size_t>(c) * m;
                double d = b[j];
                for (int64_t i = j + 1; i < m; ++i) d += a[i] * b[i];
                d *= tau;
                b[j] -= d;
                for (int64_t i = j + 1; i < m; ++i) b[i] -= d * a[i];
            }
        });
    }
    R.assign(static_cast<size_t>(n) * n, 0.);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i <= c; ++i) R[static_cast<size_t>(c) * n + i] = A[static_cast<size_t>(c) * m + i];
}

// One-sided Jacobi SVD of a square n x n column-major matrix: singular values
// (descending) and right singular vectors (column r of V = vector r).
bool jacobi_svd(int n, std::vector<double> W, std::vector<double>& S, std::vector<double>& V) {
    std::vector<double> Q(static_cast<size_t>(n) * n, 0.);
    for (int i = 0; i < n; ++i) Q[static_cast<size_t>(i) * n + i] = 1.;
    auto dotc = [&](const double* x, const double* y) {
        double s = 0.;
        for (int i = 0; i < n; ++i) s += x[i] * y[i];
        return s;
    };
    bool done = false;
    for (int sweep = 0; sweep < 80 && !done; ++sweep) {
        done = true;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double* wp = W.data() + static_cast<size_t>(p) * n;
                double* wq = W.data() + static_cast<size_t>(q) * n;
                const double alpha = dotc(wp, wp), beta = dotc(wq, wq), gamma = dotc(wp, wq);
                if (gamma == 0. || std::fabs(gamma) <= 1e-15 * std::sqrt(alpha * beta)) continue;
                done = false;
                const double zeta = (beta - alpha) / (2. * gamma);
                const double t = (zeta >= 0. ? 1. : -1.) / (std::fabs(zeta) + std::sqrt(1. + zeta * zeta));
                const double cs = 1. / std::sqrt(1. + t * t), sn = cs * t;
                double* qp = Q.data() + static_cast<size_t>(p) * n;
                double* qq = Q.data() + static_cast<size_t>(q) * n;
                for (int i = 0; i < n; ++i) {
                    const double a = wp[i], b = wq[i];
                    wp[i] = cs * a - sn * b;
                    wq[i] = sn * a + cs * b;
                    const double u = qp[i], v = qq[i];
                    qp[i] = cs * u - sn * v;
                    qq[i] = sn * u + cs * v;
                }
            }
    }
    std::vector<double> norm(n);
    std::vector<int> order(n);
    for (int j = 0; j < n; ++j) {
        norm[j] = std::sqrt(dotc(W.data() + static_cast<size_t>(j) * n, W.data() + static_cast<size_t>(j) * n));
        order[j] = j;
    }
    for (int a = 1; a < n; ++a) {  // descending, stable
        const int v = order[a];
        int b = a - 1;
        while (b >= 0 && norm[order[b]] < norm[v]) {
            order[b + 1] = order[b];
            --b;
        }
        order[b + 1] = v;
    }
    S.resize(n);
    V.resize(static_cast<size_t>(n) * n);
    for (int r = 0; r < n; ++r) {
        S[r] = norm[order[r]];
        std::memcpy(V.data() + static_cast<size_t>(r) * n, Q.data() + static_cast<size_t>(order[r]) * n,
                    sizeof(double) * n);
    }
    return done;
}

// FSvd::getRank: the smallest k whose tail sigma_k.. stays within eps^2 of the total.
int frobenius_rank(const std::vector<double>& S, double eps) {
    double total = 0.;
    for (double s : S) total += s * s;
    double tail = 0.;
    for (int k = static_cast<int>(S.size()); k > 0; --k) {
        tail += S[k - 1] * S[k - 1];
        if (tail > eps * eps * total) return k;
    }
    return 0;
}

}  // namespace

int build_global_operators(const ChebOperators& op, GlobalOperators& g, std::string& err) {
    const int p = op.p, nn = op.nn, nint = 316;
    const int64_t m1 = static_cast<int64_t>(nint) * nn;
    std::vector<double> X(3 * nn), w(nn), w1(p);
    for (int o = 0; o < p; ++o) w1[o] = M_PI / p * std::sqrt(1. - op.roots[o] * op.roots[o]);
    for (int n = 0; n < nn; ++n) {
        const int id[3] = {n % p, (n / p) % p, n / (p * p)};
        for (int d = 0; d < 3; ++d) X[3 * n + d] = op.roots[id[d]];  // cell of width 2 at the origin
        w[n] = std::sqrt(w1[id[0]] * w1[id[1]] * w1[id[2]]);
    }
    // weighted operators Kw[t][n][m] = w_n w_m K(X_m, Y_n)
    std::vector<double> Kw(static_cast<size_t>(m1) * nn);
    std::vector<int> tpos;
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k) {
                if (std::abs(i) <= 1 && std::abs(j) <= 1 && std::abs(k) <= 1) continue;
                const size_t t = tpos.size();
                tpos.push_back((i + 3) * 49 + (j + 3) * 7 + (k + 3));
                const double off[3] = {2. * i, 2. * j, 2. * k};
                for (int n = 0; n < nn; ++n)
                    for (int m = 0; m < nn; ++m) {
                        double r2 = 0.;
                        for (int d = 0; d < 3; ++d) {
                            const double dd = X[3 * m + d] - (X[3 * n + d] + off[d]);
                            r2 += dd * dd;
                        }
                        Kw[(t * nn + n) * nn + m] = w[n] * w[m] / std::sqrt(r2);
                    }
            }
    // the two tall matrices, each reduced to its singular values and right vectors
    std::vector<double> T(static_cast<size_t>(m1) * nn), R, S1, S2, V1, V2;
    for (int t = 0; t < nint; ++t)
        for (int n = 0; n < nn; ++n)
            for (int m = 0; m < nn; ++m)
                T[static_cast<size_t>(n) * m1 + static_cast<size_t>(t) * nn + m] =
                    Kw[(static_cast<size_t>(t) * nn + n) * nn + m];
    householder_r(m1, nn, T, R);
    if (!jacobi_svd(nn, R, S1, V1)) {
        err = "global M2L compression: the SVD of K_col did not converge";
        return HFMM_ERR_NUMERIC;
    }
    for (int t = 0; t < nint; ++t)
        for (int n = 0; n < nn; ++n)
            for (int m = 0; m < nn; ++m)
                T[static_cast<size_t>(m) * m1 + static_cast<size_t>(t) * nn + n] =
                    Kw[(static_cast<size_t>(t) * nn + n) * nn + m];
    householder_r(m1, nn, T, R);
    if (!jacobi_svd(nn, R, S2, V2)) {
        err = "global M2L compression: the SVD of K_row did not converge";
        return HFMM_ERR_NUMERIC;
    }
    const int k1 = frobenius_rank(S1, op.eps), k2 = frobenius_rank(S2, op.eps);
    if (k1 <= 0 || k2 <= 0) {
        err = "global M2L compression: rank 0";
        return HFMM_ERR_NUMERIC;
    }
    const int k = k1 < k2 ? k1 : k2;
    g.k = k;
    g.U.assign(V2.begin(), V2.begin() + static_cast<size_t>(nn) * k);
    g.B.assign(V1.begin(), V1.begin() + static_cast<size_t>(nn) * k);
    g.C.assign(static_cast<size_t>(343) * k * k, 0.);
    parallel_for(nint, [&](int t) {
        std::vector<double> G(static_cast<size_t>(nn) * k);
        const double* Kt = Kw.data() + static_cast<size_t>(t) * nn * nn;
        for (int a = 0; a < k; ++a)  // G[:, a] = K_t^T-as-stored U[:, a]: sum over targets m
            for (int n = 0; n < nn; ++n) {
                double s = 0.;
                for (int m = 0; m < nn; ++m) s += Kt[static_cast<size_t>(n) * nn + m] * g.U[static_cast<size_t>(a) * nn + m];
                G[static_cast<size_t>(a) * nn + n] = s;
            }
        double* Ct = g.C.data() + static_cast<size_t>(tpos[t]) * k * k;
        for (int b = 0; b < k; ++b)
            for (int a = 0; a < k; ++a) {
                double s = 0.;
                for (int n = 0; n < nn; ++n) s += G[static_cast<size_t>(a) * nn + n] * g.B[static_cast<size_t>(b) * nn + n];
                Ct[a + static_cast<size_t>(b) * k] = s;
            }
    });
    for (int a = 0; a < k; ++a)
        for (int n = 0; n < nn; ++n) {
            g.U[static_cast<size_t>(a) * nn + n] /= w[n];
            g.B[static_cast<size_t>(a) * nn + n] /= w[n];
        }
    g.S1 = S1;
    g.S2 = S2;
    return HFMM_OK;
}

int build_dense_operators(const ChebOperators& op, GlobalOperators& g) {
    // FChebDenseM2LHandler (Src/Kernels/Chebyshev/FChebDenseM2LHandler.hpp:180-215):
    // the 316 operators K_t[m][n] = K(X_m, Y_n) uncompressed, rank = p^3, so
    // that the global scheme's passes apply them with identity bases
    const int p = op.p, nn = op.nn;
    std::vector<double> X(3 * nn);
    for (int n = 0; n < nn; ++n) {
        const int id[3] = {n % p, (n / p) % p, n / (p * p)};
        for (int d = 0; d < 3; ++d) X[3 * n + d] = op.roots[id[d]];
    }
    g.k = nn;
    g.U.assign(static_cast<size_t>(nn) * nn, 0.);
    for (int n = 0; n < nn; ++n) g.U[static_cast<size_t>(n) * nn + n] = 1.;
    g.B = g.U;
    g.S1.assign(nn, 0.);
    g.S2.assign(nn, 0.);
    g.C.assign(static_cast<size_t>(343) * nn * nn, 0.);
    std::vector<int> tpos;
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k)
                if (std::abs(i) > 1 || std::abs(j) > 1 || std::abs(k) > 1) tpos.push_back((i + 3) * 49 + (j + 3) * 7 + (k + 3));
    parallel_for(static_cast<int>(tpos.size()), [&](int t) {
        const int pos = tpos[t];
        const int off[3] = {2 * (pos / 49 - 3), 2 * ((pos / 7) % 7 - 3), 2 * (pos % 7 - 3)};
        double* Ct = g.C.data() + static_cast<size_t>(pos) * nn * nn;
        for (int n = 0; n < nn; ++n)
            for (int m = 0; m < nn; ++m) {
                double r2 = 0.;
                for (int d = 0; d < 3; ++d) {
                    const double dd = X[3 * m + d] - (X[3 * n + d] + off[d]);
                    r2 += dd * dd;
                }
                Ct[m + static_cast<size_t>(n) * nn] = 1. / std::sqrt(r2);
            }
    });
    return HFMM_OK;
}

int build_uniform_fft_operators(const ChebOperators& op, GlobalOperators& g) {
    const int p = op.p, N = 2 * p - 1, NF = p * N * N;
    const double pi = 3.14159265358979323846;
    g.k = 2 * NF;
    g.U.clear();
    g.B.clear();
    g.C.clear();
    g.S1.clear();
    g.S2.clear();
    g.tw.assign(2 * N, 0.);
    for (int j = 0; j < N; ++j) {
        g.tw[2 * j] = std::cos(2. * pi * j / N);
        g.tw[2 * j + 1] = std::sin(2. * pi * j / N);
    }
    g.Khat.assign(static_cast<size_t>(343) * NF * 2, 0.);
    std::vector<int> tpos;
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k)
                if (std::abs(i) > 1 || std::abs(j) > 1 || std::abs(k) > 1) tpos.push_back((i + 3) * 49 + (j + 3) * 7 + (k + 3));
    parallel_for(static_cast<int>(tpos.size()), [&](int t) {
        const int pos = tpos[t];
        const int off[3] = {2 * (pos / 49 - 3), 2 * ((pos / 7) % 7 - 3), 2 * (pos % 7 - 3)};
        // circulant c[(dx mod N) + N ((dy mod N) + N (dz mod N))] = K for the node
        // index difference d = target - source: target node max(d, 0), source
        // node target - d on every axis (reference cell width 2, source at 2 off)
        std::vector<double> c(static_cast<size_t>(N) * N * N);
        for (int dz = -(p - 1); dz < p; ++dz)
            for (int dy = -(p - 1); dy < p; ++dy)
                for (int dx = -(p - 1); dx < p; ++dx) {
                    const int d[3] = {dx, dy, dz};
                    double r2 = 0.;
                    for (int a = 0; a < 3; ++a) {
                        const int tn = d[a] > 0 ? d[a] : 0, sn = tn - d[a];
                        const double diff = op.roots[tn] - (op.roots[sn] + off[a]);
                        r2 += diff * diff;
                    }
                    const int cx = (dx + N) % N, cy = (dy + N) % N, cz = (dz + N) % N;
                    c[cx + static_cast<size_t>(N) * (cy + static_cast<size_t>(N) * cz)] = 1. / std::sqrt(r2);
                }
        // forward 3D DFT (w^-(a kx + b ky + c kz)), one axis at a time, kx < p kept
        std::vector<double> t1(static_cast<size_t>(p) * N * N * 2), t2(t1.size());
        for (int cz = 0; cz < N; ++cz)
            for (int cy = 0; cy < N; ++cy)
                for (int kx = 0; kx < p; ++kx) {
                    double re = 0., im = 0.;
                    for (int a = 0; a < N; ++a) {
                        const double v = c[a + static_cast<size_t>(N) * (cy + static_cast<size_t>(N) * cz)];
                        const int w = (a * kx) % N;
                        re += v * g.tw[2 * w];
                        im -= v * g.tw[2 * w + 1];
                    }
                    const size_t o = 2 * (kx + static_cast<size_t>(p) * (cy + static_cast<size_t>(N) * cz));
                    t1[o] = re;
                    t1[o + 1] = im;
                }
        for (int cz = 0; cz < N; ++cz)
            for (int ky = 0; ky < N; ++ky)
                for (int kx = 0; kx < p; ++kx) {
                    double re = 0., im = 0.;
                    for (int b = 0; b < N; ++b) {
                        const size_t o = 2 * (kx + static_cast<size_t>(p) * (b + static_cast<size_t>(N) * cz));
                        const int w = (b * ky) % N;
                        const double wr = g.tw[2 * w], wi = g.tw[2 * w + 1];
                        re += t1[o] * wr + t1[o + 1] * wi;
                        im += t1[o + 1] * wr - t1[o] * wi;
                    }
                    const size_t o = 2 * (kx + static_cast<size_t>(p) * (ky + static_cast<size_t>(N) * cz));
                    t2[o] = re;
                    t2[o + 1] = im;
                }
        double* K = g.Khat.data() + static_cast<size_t>(pos) * NF * 2;
        for (int kz = 0; kz < N; ++kz)
            for (int ky = 0; ky < N; ++ky)
                for (int kx = 0; kx < p; ++kx) {
                    double re = 0., im = 0.;
                    for (int cc = 0; cc < N; ++cc) {
                        const size_t o = 2 * (kx + static_cast<size_t>(p) * (ky + static_cast<size_t>(N) * cc));
                        const int w = (cc * kz) % N;
                        const double wr = g.tw[2 * w], wi = g.tw[2 * w + 1];
                        re += t2[o] * wr + t2[o + 1] * wi;
                        im += t2[o + 1] * wr - t2[o] * wi;
                    }
                    const size_t o = 2 * (kx + static_cast<size_t>(p) * (ky + static_cast<size_t>(N) * kz));
                    K[o] = re;
                    K[o + 1] = im;
                }
    });
    return HFMM_OK;
}

}  // namespace hfmm
