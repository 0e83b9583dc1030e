// cheb_setup.cpp -- host FP64 set-up of the Chebyshev 1/r operators.
//
// Follows (paths relative to the reference's source tree):
//   roots            Src/Kernels/Chebyshev/FChebRoots.hpp:90-297
//   T_o(x)           Src/Kernels/Chebyshev/FChebRoots.hpp:46-58
//   interpolators    Src/Kernels/Chebyshev/FChebInterpolator.hpp:300-338, 352-402, 466-485
//   symmetries       Src/Kernels/Interpolation/FInterpSymmetries.hpp:42-125
//   compression      Src/Kernels/Chebyshev/FChebSymM2LHandler.hpp:39-281 with
//                    Src/Utils/FAca.hpp:141-226 (partially pivoted ACA) and
//                    Src/Utils/FSvd.hpp:51-62 (rank rule)
// The reference calls LAPACK (geqrf/orgqr/gesvd) for the recompression; this
// file uses its own Householder QR and one-sided Jacobi SVD.
#include "cheb_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>

#include "../../include/scalfmm_hip.h"

namespace hfmm {
namespace {

// Decimal root tables of the reference (kept digit for digit: the interpolation
// nodes must be the reference's nodes, which differ from the exact cosines by
// up to ~5e-16).
const double kRoots[13][14] = {
    {-0.707106781186548, 0.707106781186547},
    {-8.66025403784439e-01, 0.0, 8.66025403784438e-01},
    {-0.923879532511287, -0.382683432365090, 0.382683432365090, 0.923879532511287},
    {-9.51056516295154e-01, -5.87785252292473e-01, 0.0, 5.87785252292473e-01, 9.51056516295154e-01},
    {-0.965925826289068, -0.707106781186548, -0.258819045102521, 0.258819045102521, 0.707106781186547,
     0.965925826289068},
    {-9.74927912181824e-01, -7.81831482468030e-01, -4.33883739117558e-01, 0.0, 4.33883739117558e-01,
     7.81831482468030e-01, 9.74927912181824e-01},
    {-0.980785280403230, -0.831469612302545, -0.555570233019602, -0.195090322016128, 0.195090322016128,
     0.555570233019602, 0.831469612302545, 0.980785280403230},
    {-9.84807753012208e-01, -8.66025403784439e-01, -6.42787609686539e-01, -3.42020143325669e-01, 0.0,
     3.42020143325669e-01, 6.42787609686539e-01, 8.66025403784438e-01, 9.84807753012208e-01},
    {-0.987688340595138, -0.891006524188368, -0.707106781186548, -0.453990499739547, -0.156434465040231,
     0.156434465040231, 0.453990499739547, 0.707106781186547, 0.891006524188368, 0.987688340595138},
    {-9.89821441880933e-01, -9.09631995354518e-01, -7.55749574354258e-01, -5.40640817455598e-01,
     -2.81732556841430e-01, 0.0, 2.81732556841430e-01, 5.40640817455597e-01, 7.55749574354258e-01,
     9.09631995354518e-01, 9.89821441880933e-01},
    {-0.991444861373810, -0.923879532511287, -0.793353340291235, -0.608761429008721, -0.382683432365090,
     -0.130526192220052, 0.130526192220052, 0.382683432365090, 0.608761429008721, 0.793353340291235,
     0.923879532511287, 0.991444861373810},
    {-9.92708874098054e-01, -9.35016242685415e-01, -8.22983865893656e-01, -6.63122658240795e-01,
     -4.64723172043769e-01, -2.39315664287558e-01, 0.0, 2.39315664287557e-01, 4.64723172043769e-01,
     6.63122658240795e-01, 8.22983865893656e-01, 9.35016242685415e-01, 9.92708874098054e-01},
    {-0.99371220989324258353, -0.94388333030836756290, -0.84672419922828416835, -0.70710678118654752440,
     -0.53203207651533656356, -0.33027906195516708177, -0.11196447610330785847, 0.11196447610330785847,
     0.33027906195516708177, 0.53203207651533656356, 0.70710678118654752440, 0.84672419922828416835,
     0.94388333030836756290, 0.99371220989324258353},
};

double chebT(unsigned n, double x) {
    if (std::fabs(x) > 1.) x = std::min(1., std::max(-1., x));
    return std::cos(n * std::acos(x));
}

struct Vec3 {
    double v[3];
};

// Interpolation nodes of a cell (centre, width): (a+b)/2 + (b-a) * root / 2
// per axis, node n <-> (n % p, (n / p) % p, n / p^2).
std::vector<Vec3> cell_nodes(const ChebOperators& op, const double c[3], double w) {
    std::vector<Vec3> X(op.nn);
    const int p = op.p;
    for (int n = 0; n < op.nn; ++n) {
        const int id[3] = {n % p, (n / p) % p, n / (p * p)};
        for (int d = 0; d < 3; ++d) {
            const double a = c[d] - w / 2., b = c[d] + w / 2.;
            X[n].v[d] = (a + b) / 2. + (b - a) * op.roots[id[d]] / 2.;
        }
    }
    return X;
}

inline double inv_r(const Vec3& t, const Vec3& s) {
    const double dx = t.v[0] - s.v[0], dy = t.v[1] - s.v[1], dz = t.v[2] - s.v[2];
    return 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
}

double dotp(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// Weighted kernel block A[x][y] = w_x w_y K(Y_y, X_x): rows are target nodes
// X, columns source nodes Y (the EntryComputer of FInterpMatrixKernel.hpp:561-608).
struct WeightedBlock {
    const std::vector<Vec3>& X;
    const std::vector<Vec3>& Y;
    const std::vector<double>& w;
    void row(int x, double* out) const {
        for (size_t y = 0; y < Y.size(); ++y) out[y] = w[y] * w[x] * inv_r(Y[y], X[x]);
    }
    void col(int y, double* out) const {
        for (size_t x = 0; x < X.size(); ++x) out[x] = w[y] * w[x] * inv_r(Y[y], X[x]);
    }
};

// Partially pivoted ACA with the reference's pivoting (first maximum, strict
// comparison) and stopping rule |u_k|^2 |v_k|^2 <= eps^2 |S_k|_F^2.
int aca_partial(const WeightedBlock& A, int nx, int ny, double eps, std::vector<double>& U,
                std::vector<double>& V, int& k) {
    const int maxk = (nx + ny) / 2;
    U.assign(static_cast<size_t>(nx) * maxk, 0.0);
    V.assign(static_cast<size_t>(ny) * maxk, 0.0);
    std::vector<char> row_free(nx, 1), col_free(ny, 1);
    double norm2S = 0., norm2uv = 0.;
    int I = 0, J = 0;
    k = 0;
    do {
        if (k >= maxk) return HFMM_ERR_NUMERIC;
        double* u = &U[static_cast<size_t>(nx) * k];
        double* v = &V[static_cast<size_t>(ny) * k];
        A.row(I, v);
        row_free[I] = 0;
        for (int l = 0; l < k; ++l) {
            const double a = -1. * U[static_cast<size_t>(nx) * l + I];
            const double* vl = &V[static_cast<size_t>(ny) * l];
            for (int j = 0; j < ny; ++j) v[j] += a * vl[j];
        }
        double best = 0.;
        for (int j = 0; j < ny; ++j) {
            const double m = std::fabs(v[j]);
            if (col_free[j] && best < m) { best = m; J = j; }
        }
        const double piv = 1. / v[J];
        for (int j = 0; j < ny; ++j) v[j] *= piv;
        A.col(J, u);
        col_free[J] = 0;
        for (int l = 0; l < k; ++l) {
            const double a = -1. * V[static_cast<size_t>(ny) * l + J];
            const double* ul = &U[static_cast<size_t>(nx) * l];
            for (int i = 0; i < nx; ++i) u[i] += a * ul[i];
        }
        best = 0.;
        for (int i = 0; i < nx; ++i) {
            const double m = std::fabs(u[i]);
            if (row_free[i] && best < m) { best = m; I = i; }
        }
        double cross = 0.;
        for (int l = 0; l < k; ++l)
            cross += dotp(nx, u, &U[static_cast<size_t>(nx) * l]) * dotp(ny, &V[static_cast<size_t>(ny) * l], v);
        norm2uv = dotp(nx, u, u) * dotp(ny, v, v);
        norm2S += norm2uv + 2 * cross;
        ++k;
    } while (norm2uv > eps * eps * norm2S);
    return HFMM_OK;
}

// Thin Householder QR of a column-major m x n matrix: Q (m x n), R (n x n).
void householder_qr(int m, int n, const double* A, std::vector<double>& Q, std::vector<double>& R) {
    std::vector<double> W(A, A + static_cast<size_t>(m) * n);
    std::vector<double> tau(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double* cj = &W[static_cast<size_t>(j) * m];
        const double norm = std::sqrt(dotp(m - j, cj + j, cj + j));
        if (norm == 0.0) continue;
        const double beta = cj[j] >= 0 ? -norm : norm;
        const double head = cj[j] - beta;
        for (int i = j + 1; i < m; ++i) cj[i] /= head;
        tau[j] = -head / beta;
        cj[j] = beta;
        for (int c = j + 1; c < n; ++c) {
            double* cc = &W[static_cast<size_t>(c) * m];
            double s = cc[j];
            for (int i = j + 1; i < m; ++i) s += cj[i] * cc[i];
            s *= tau[j];
            cc[j] -= s;
            for (int i = j + 1; i < m; ++i) cc[i] -= s * cj[i];
        }
    }
    R.assign(static_cast<size_t>(n) * n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i <= c; ++i) R[static_cast<size_t>(c) * n + i] = W[static_cast<size_t>(c) * m + i];
    Q.assign(static_cast<size_t>(m) * n, 0.0);
    for (int c = 0; c < n; ++c) Q[static_cast<size_t>(c) * m + c] = 1.0;
    for (int j = n - 1; j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        const double* v = &W[static_cast<size_t>(j) * m];
        for (int c = 0; c < n; ++c) {
            double* qc = &Q[static_cast<size_t>(c) * m];
            double s = qc[j];
            for (int i = j + 1; i < m; ++i) s += v[i] * qc[i];
            s *= tau[j];
            qc[j] -= s;
            for (int i = j + 1; i < m; ++i) qc[i] -= s * v[i];
        }
    }
}

// One-sided Jacobi SVD of a square column-major n x n matrix:
// A = U diag(S) V^T with S descending.  Returns false if it did not converge.
bool jacobi_svd(int n, const std::vector<double>& A, std::vector<double>& U, std::vector<double>& S,
                std::vector<double>& V) {
    std::vector<double> W = A;
    std::vector<double> Vw(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) Vw[static_cast<size_t>(i) * n + i] = 1.0;
    const double tol = 2.220446049250313e-16 * n;
    bool done = false;
    for (int sweep = 0; sweep < 100 && !done; ++sweep) {
        done = true;
        for (int a = 0; a < n - 1; ++a)
            for (int b = a + 1; b < n; ++b) {
                double* wa = &W[static_cast<size_t>(a) * n];
                double* wb = &W[static_cast<size_t>(b) * n];
                const double aa = dotp(n, wa, wa), bb = dotp(n, wb, wb), ab = dotp(n, wa, wb);
                if (ab == 0.0 || std::fabs(ab) <= tol * std::sqrt(aa * bb)) continue;
                done = false;
                const double zeta = (bb - aa) / (2.0 * ab);
                const double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                double* va = &Vw[static_cast<size_t>(a) * n];
                double* vb = &Vw[static_cast<size_t>(b) * n];
                for (int i = 0; i < n; ++i) {
                    const double x = wa[i], y = wb[i];
                    wa[i] = c * x - s * y;
                    wb[i] = s * x + c * y;
                    const double xv = va[i], yv = vb[i];
                    va[i] = c * xv - s * yv;
                    vb[i] = s * xv + c * yv;
                }
            }
    }
    std::vector<double> norms(n);
    for (int j = 0; j < n; ++j) norms[j] = std::sqrt(dotp(n, &W[static_cast<size_t>(j) * n], &W[static_cast<size_t>(j) * n]));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return norms[x] > norms[y]; });
    U.assign(static_cast<size_t>(n) * n, 0.0);
    V.assign(static_cast<size_t>(n) * n, 0.0);
    S.assign(n, 0.0);
    for (int r = 0; r < n; ++r) {
        const int j = order[r];
        S[r] = norms[j];
        for (int i = 0; i < n; ++i) {
            U[static_cast<size_t>(r) * n + i] = norms[j] > 0 ? W[static_cast<size_t>(j) * n + i] / norms[j] : 0.0;
            V[static_cast<size_t>(r) * n + i] = Vw[static_cast<size_t>(j) * n + i];
        }
    }
    return done;
}

// Smallest k with sum_{i>=k} S_i^2 <= eps^2 sum_i S_i^2 (FSvd::getRank).
int truncation_rank(const std::vector<double>& S, double eps) {
    const int n = static_cast<int>(S.size());
    const double total = dotp(n, S.data(), S.data());
    double tail = 0.;
    for (int k = n; k > 0; --k) {
        tail += S[k - 1] * S[k - 1];
        if (tail > eps * eps * total) return k;
    }
    return -1;
}

void build_symmetries(ChebOperators& op) {
    const int p = op.p, nn = op.nn;
    // 8 reflections (quadrants) and 6 axis permutations (cones 0,1,3,4,6,7).
    auto reflect = [p](int q, int i, int j, int k) {
        const int ii = (q & 1) ? p - 1 - i : i;
        const int jj = (q & 2) ? p - 1 - j : j;
        const int kk = (q & 4) ? p - 1 - k : k;
        return std::array<int, 3>{ii, jj, kk};
    };
    // cone c: node (i,j,k) -> linear index with permuted axes (FInterpSymmetries.hpp:62-70)
    auto cone_index = [p](int c, int i, int j, int k) {
        switch (c) {
            case 0: return k * p * p + j * p + i;
            case 1: return j * p * p + k * p + i;
            case 3: return j * p * p + i * p + k;
            case 4: return k * p * p + i * p + j;
            case 6: return i * p * p + k * p + j;
            default: return i * p * p + j * p + k;  // 7
        }
    };
    static const int axis_perm[8][3] = {{0, 1, 2}, {0, 2, 1}, {0, 0, 0}, {2, 0, 1},
                                        {1, 0, 2}, {0, 0, 0}, {1, 2, 0}, {2, 1, 0}};
    op.pvec.assign(static_cast<size_t>(343) * nn, 0);
    op.pidx.fill(0);
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k) {
                if (std::abs(i) <= 1 && std::abs(j) <= 1 && std::abs(k) <= 1) continue;
                const int pos = ((i + 3) * 7 + (j + 3)) * 7 + (k + 3);
                const int q = (i < 0 ? 1 : 0) | (j < 0 ? 2 : 0) | (k < 0 ? 4 : 0);
                const int u[3] = {std::abs(i), std::abs(j), std::abs(k)};
                const int c = ((u[1] > u[0]) << 2) | ((u[2] > u[0]) << 1) | (u[2] > u[1]);
                for (int kk = 0; kk < p; ++kk)
                    for (int jj = 0; jj < p; ++jj)
                        for (int ii = 0; ii < p; ++ii) {
                            const auto r = reflect(q, ii, jj, kk);
                            op.pvec[static_cast<size_t>(pos) * nn + kk * p * p + jj * p + ii] =
                                static_cast<uint16_t>(cone_index(c, r[0], r[1], r[2]));
                        }
                op.pidx[pos] = static_cast<uint32_t>((u[axis_perm[c][0]] + 3) * 49 + (u[axis_perm[c][1]] + 3) * 7 +
                                                     (u[axis_perm[c][2]] + 3));
            }
}

int compress_operators(ChebOperators& op, std::string& err) {
    const int p = op.p, nn = op.nn;
    const double cw = 2.0;  // reference cell width of the homogeneous kernel
    const double origin[3] = {0, 0, 0};
    const auto X = cell_nodes(op, origin, cw);
    std::vector<double> w1(p), w(nn);
    for (int o = 0; o < p; ++o) w1[o] = M_PI / p * std::sqrt(1. - op.roots[o] * op.roots[o]);
    for (int n = 0; n < nn; ++n) w[n] = std::sqrt(w1[n % p] * w1[(n / p) % p] * w1[n / (p * p)]);
    for (int i = 2; i <= 3; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= j; ++k) {
                const double cy[3] = {cw * i, cw * j, cw * k};
                const auto Y = cell_nodes(op, cy, cw);
                WeightedBlock A{X, Y, w};
                std::vector<double> U, V;
                int r = 0;
                if (aca_partial(A, nn, nn, op.eps, U, V, r) != HFMM_OK) {
                    err = "ACA did not converge";
                    return HFMM_ERR_NUMERIC;
                }
                std::vector<double> QU, RU, QV, RV;
                householder_qr(nn, r, U.data(), QU, RU);
                householder_qr(nn, r, V.data(), QV, RV);
                std::vector<double> phi(static_cast<size_t>(r) * r, 0.0);  // R_U R_V^T
                for (int b = 0; b < r; ++b)
                    for (int a = 0; a < r; ++a) {
                        double s = 0;
                        for (int l = 0; l < r; ++l) s += RU[static_cast<size_t>(l) * r + a] * RV[static_cast<size_t>(l) * r + b];
                        phi[static_cast<size_t>(b) * r + a] = s;
                    }
                std::vector<double> Uphi, S, Vphi;
                if (!jacobi_svd(r, phi, Uphi, S, Vphi)) {
                    err = "SVD did not converge";
                    return HFMM_ERR_NUMERIC;
                }
                const int k2 = truncation_rank(S, op.eps);
                if (k2 <= 0) {
                    err = "empty rank";
                    return HFMM_ERR_NUMERIC;
                }
                const int idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3);
                std::vector<double> Kc(static_cast<size_t>(2) * k2 * nn, 0.0);
                for (int c = 0; c < k2; ++c)
                    for (int n = 0; n < nn; ++n) {
                        double su = 0, sv = 0;
                        for (int a = 0; a < r; ++a) {
                            su += QU[static_cast<size_t>(a) * nn + n] * (Uphi[static_cast<size_t>(c) * r + a] * S[c]);
                            sv += QV[static_cast<size_t>(a) * nn + n] * Vphi[static_cast<size_t>(c) * r + a];
                        }
                        const double iw = 1. / w[n];  // un-weighting
                        Kc[static_cast<size_t>(c) * nn + n] = su * iw;
                        Kc[static_cast<size_t>(k2) * nn + static_cast<size_t>(c) * nn + n] = sv * iw;
                    }
                op.K[idx] = std::move(Kc);
                op.rank[idx] = k2;
            }
    return HFMM_OK;
}

}  // namespace

bool cheb_roots(int p, double* out) {
    if (p < kMinOrder || p > kMaxOrder) return false;
    for (int j = 0; j < p; ++j) out[j] = kRoots[p - 2][j];
    return true;
}

int build_cheb_operators(int order, double eps, ChebOperators& op, std::string& err) {
    if (order < kMinOrder || order > kMaxOrder) {
        err = "order out of range";
        return HFMM_ERR_ARG;
    }
    op.p = order;
    op.nn = order * order * order;
    op.eps = eps > 0 ? eps : std::pow(10.0, static_cast<double>(-order));
    cheb_roots(order, op.roots.data());
    const int p = order;
    op.Troots.assign(static_cast<size_t>(p) * p, 1.0);
    for (int o = 1; o < p; ++o)
        for (int j = 0; j < p; ++j) op.Troots[static_cast<size_t>(o) * p + j] = chebT(o, op.roots[j]);
    // child/parent 1-D interpolators (child width 1 in the parent's [-1,1] frame,
    // centre -0.5 or +0.5 per axis by the Morton child bits x=4, y=2, z=1)
    op.child.assign(static_cast<size_t>(8) * 3 * p * p, 0.0);
    for (int ch = 0; ch < 8; ++ch)
        for (int d = 0; d < 3; ++d) {
            const int bit = (ch >> (2 - d)) & 1;
            const double ctr = (bit ? 1.0 : -1.0) * (1. - .5);
            const double a = ctr - 1. / 2., b = ctr + 1. / 2.;
            double* S = &op.child[(static_cast<size_t>(ch) * 3 + d) * p * p];
            for (int m = 0; m < p; ++m) {
                const double xm = (a + b) / 2. + (b - a) * op.roots[m] / 2.;
                double Tx[kMaxOrder];
                for (int o = 1; o < p; ++o) Tx[o] = chebT(o, xm);
                for (int n = 0; n < p; ++n) {
                    double s = 1. / p;
                    for (int o = 1; o < p; ++o) s += 2. / p * Tx[o] * op.Troots[static_cast<size_t>(o) * p + n];
                    S[n * p + m] = s;
                }
            }
        }
    build_symmetries(op);
    op.rank.fill(0);
    return compress_operators(op, err);
}

void far_slots(int parity, FarSlot out[189]) {
    const int bx = (parity >> 2) & 1, by = (parity >> 1) & 1, bz = parity & 1;
    int s = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (!dx && !dy && !dz) continue;
                for (int c = 0; c < 8; ++c) {
                    const int xd = 2 * dx + ((c >> 2) & 1) - bx;
                    const int yd = 2 * dy + ((c >> 1) & 1) - by;
                    const int zd = 2 * dz + (c & 1) - bz;
                    if (std::abs(xd) > 1 || std::abs(yd) > 1 || std::abs(zd) > 1) {
                        out[s].pdx = static_cast<int8_t>(dx);
                        out[s].pdy = static_cast<int8_t>(dy);
                        out[s].pdz = static_cast<int8_t>(dz);
                        out[s].cousin = static_cast<int8_t>(c);
                        out[s].pos = static_cast<int16_t>(((xd + 3) * 7 + (yd + 3)) * 7 + zd + 3);
                        ++s;
                    }
                }
            }
}

}  // namespace hfmm
