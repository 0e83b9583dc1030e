// cheb_setup.cpp -- host FP64 tables of the Chebyshev 1/r operators (the
// compression of the M2L operators runs on the device, cheb_compress.hip).
//
// Follows (paths relative to the reference's source tree):
//   roots            Src/Kernels/Chebyshev/FChebRoots.hpp:90-297
//   T_o(x)           Src/Kernels/Chebyshev/FChebRoots.hpp:46-58
//   interpolators    Src/Kernels/Chebyshev/FChebInterpolator.hpp:300-338, 352-402, 466-485
//   symmetries       Src/Kernels/Interpolation/FInterpSymmetries.hpp:42-125
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
    op.rank.fill(0);  // the compressed operators are built on the device (cheb_compress.hip)
    return HFMM_OK;
}

double unif_lagrange(int p, int n, double x) {
    // FUnifRoots::L (Src/Kernels/Uniform/FUnifRoots.hpp:49-83): the scaled product
    if (x > 1.) x = 1.;
    if (x < -1.) x = -1.;
    const int omn = p - n - 1;
    double fn = 1., fo = 1.;
    for (int i = 2; i <= n; ++i) fn *= i;
    for (int i = 2; i <= omn; ++i) fo *= i;
    double scale = (omn % 2) ? -1. : 1.;
    scale /= std::pow(2., static_cast<double>(p - 1)) * fn * fo;
    double L = 1.;
    for (int m = 0; m < p; ++m)
        if (m != n) L *= static_cast<double>(p - 1) * (x + 1.) - 2. * m;
    return L * scale;
}

int make_uniform_tables(ChebOperators& op) {
    // the equispaced roots as FUnifRootsCore tabulates them (FUnifRoots.hpp:130-240)
    static const double r2[2] = {-1., 1.};
    static const double r3[3] = {-1., 0., 1.};
    static const double r4[4] = {-1., -0.333333333333333, 0.333333333333333, 1.};
    static const double r5[5] = {-1., -0.5, 0., 0.5, 1.};
    static const double r6[6] = {-1., -0.6, -0.2, 0.2, 0.6, 1.};
    static const double r7[7] = {-1., -0.666666666666666, -0.333333333333333, 0., 0.333333333333333,
                                 0.666666666666666, 1.};
    static const double r8[8] = {-1., -0.714285714285714, -0.428571428571429, -0.142857142857143,
                                 0.142857142857143, 0.428571428571429, 0.714285714285714, 1.};
    const int p = op.p;
    const double* r = p == 2 ? r2 : p == 3 ? r3 : p == 4 ? r4 : p == 5 ? r5 : p == 6 ? r6 : p == 7 ? r7 : p == 8 ? r8 : nullptr;
    if (!r) return HFMM_ERR_ARG;
    for (int i = 0; i < p; ++i) op.roots[i] = r[i];
    // child/parent interpolators (FUnifInterpolator.hpp:100-123): child width 1,
    // centre +-0.5, S[n][m] = L_n(child root m)
    for (int ch = 0; ch < 8; ++ch)
        for (int d = 0; d < 3; ++d) {
            const int bit = (ch >> (2 - d)) & 1;
            const double ctr = bit ? 0.5 : -0.5;
            double* S = &op.child[(static_cast<size_t>(ch) * 3 + d) * p * p];
            for (int m = 0; m < p; ++m)
                for (int n = 0; n < p; ++n) S[n * p + m] = unif_lagrange(p, n, ctr + op.roots[m] / 2.);
        }
    return HFMM_OK;
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
