/*
 * oracle.c -- CPU restatement (FP64) of the reference Chebyshev 1/r FMM path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product never links it.
 *
 * Reference paths are relative to the reference's Src/ directory.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - Morton / coordinates: utestMorton known answers, round trips, closed form.
 *   - lists: brute-force definition on small trees, full-level closed-form counts.
 *   - roots / interpolators: closed forms, polynomial reproduction.
 *   - symmetries: permuted canonical blocks equal directly assembled blocks
 *     (the rule of Tests/Utils/testChebSymmetries.cpp).
 *   - QR / SVD: numpy.linalg on random matrices.
 *   - M2L operators: low-rank error bound against the dense kernel block.
 *   - P2M/M2M/L2L/L2P: M2M(P2M child) == P2M parent, L2P(L2L) == L2P parent,
 *     gradient vs finite differences.
 *   - P2P / direct: fixtures Data/UTest/unitCubeRef20kDouble.bfma and
 *     Data/UTest/DirectDouble.bfma.
 *   - whole FMM: the published errors of Data/ResultCompareKernels.txt:28-35
 *     (order 7, height 5, 20k particles).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double wall_seconds(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* ======================================================================== */
/* Morton indexing (Containers/FTreeCoordinate.hpp:91-142)                  */
/* ======================================================================== */

int64_t orc_morton_encode(int x, int y, int z) {
    /* bit 3k <- z_k, bit 3k+1 <- y_k, bit 3k+2 <- x_k (FTreeCoordinate.hpp:91-116) */
    int64_t key = 0;
    for (int b = 0; b < 21; ++b) {
        key |= (int64_t)((z >> b) & 1) << (3 * b);
        key |= (int64_t)((y >> b) & 1) << (3 * b + 1);
        key |= (int64_t)((x >> b) & 1) << (3 * b + 2);
    }
    return key;
}

void orc_morton_decode(int64_t key, int32_t out[3]) {
    /* FTreeCoordinate.hpp:126-142 */
    int x = 0, y = 0, z = 0;
    for (int b = 0; b < 21; ++b) {
        z |= (int)((key >> (3 * b)) & 1) << b;
        y |= (int)((key >> (3 * b + 1)) & 1) << b;
        x |= (int)((key >> (3 * b + 2)) & 1) << b;
    }
    out[0] = x; out[1] = y; out[2] = z;
}

/* FCoordinateComputer::GetTreeCoordinate (FCoordinateComputer.hpp:15-24) */
static int tree_coordinate(double rel, double width, double leafw, int height, int* err) {
    if (!(rel >= 0 && rel <= width)) { *err = 1; return 0; }
    if (rel == width) return (1 << (height - 1)) - 1;
    return (int)(rel / leafw);
}

int orc_particle_keys(int64_t n, const double* x, const double* y, const double* z,
                      const double center[3], double width, int height,
                      int32_t* coords, int64_t* keys) {
    if (n < 0 || height < 2 || height > 20 || !(width > 0)) return ORC_ERR_ARG;
    /* boxCorner(center, -(width/2)) (FGroupTree.hpp:183) */
    const double half = width / 2;
    const double corner[3] = {center[0] + (-half), center[1] + (-half), center[2] + (-half)};
    const double leafw = width / (double)(1 << (height - 1));
    int bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        int err = 0;
        const int cx = tree_coordinate(x[i] - corner[0], width, leafw, height, &err);
        const int cy = tree_coordinate(y[i] - corner[1], width, leafw, height, &err);
        const int cz = tree_coordinate(z[i] - corner[2], width, leafw, height, &err);
        if (err) { bad = 1; continue; }
        if (coords) { coords[3 * i] = cx; coords[3 * i + 1] = cy; coords[3 * i + 2] = cz; }
        keys[i] = orc_morton_encode(cx, cy, cz);
    }
    return bad ? ORC_ERR_OUTSIDE_BOX : ORC_OK;
}

/* FMath::Between(v, lo, hi) = lo <= v < hi (Utils/FMath.hpp:257-259) */
static int between(int v, int lo, int hi) { return lo <= v && v < hi; }

int orc_neighbor_keys(int x, int y, int z, int height, int64_t keys[26], int32_t pos[26]) {
    /* FTreeCoordinate::getNeighborsIndexes (FTreeCoordinate.hpp:189-214) */
    int cnt = 0;
    const int lim = 1 << (height - 1);
    for (int dx = -1; dx <= 1; ++dx) {
        if (!between(x + dx, 0, lim)) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            if (!between(y + dy, 0, lim)) continue;
            for (int dz = -1; dz <= 1; ++dz) {
                if (!between(z + dz, 0, lim)) continue;
                if (dx || dy || dz) {
                    keys[cnt] = orc_morton_encode(x + dx, y + dy, z + dz);
                    pos[cnt] = ((dx + 1) * 3 + (dy + 1)) * 3 + (dz + 1);
                    ++cnt;
                }
            }
        }
    }
    return cnt;
}

int orc_interaction_keys(int x, int y, int z, int level, int64_t keys[216], int32_t pos[216]) {
    /* FTreeCoordinate::getInteractionNeighbors (FTreeCoordinate.hpp:219-264),
     * neighSeparation = 1 */
    const int px = x >> 1, py = y >> 1, pz = z >> 1;
    const int lim = 1 << (level - 1);
    int cnt = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        if (!between(px + dx, 0, lim)) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            if (!between(py + dy, 0, lim)) continue;
            for (int dz = -1; dz <= 1; ++dz) {
                if (!between(pz + dz, 0, lim)) continue;
                if (dx || dy || dz) {
                    const int ox = px + dx, oy = py + dy, oz = pz + dz;
                    const int64_t mo = orc_morton_encode(ox, oy, oz);
                    for (int c = 0; c < 8; ++c) {
                        const int xd = ((ox << 1) | ((c >> 2) & 1)) - x;
                        const int yd = ((oy << 1) | ((c >> 1) & 1)) - y;
                        const int zd = ((oz << 1) | (c & 1)) - z;
                        if (abs(xd) > 1 || abs(yd) > 1 || abs(zd) > 1) {
                            pos[cnt] = (((xd + 3) * 7) + (yd + 3)) * 7 + zd + 3;
                            keys[cnt] = (mo << 3) | c;
                            ++cnt;
                        }
                    }
                }
            }
        }
    }
    return cnt;
}

/* ======================================================================== */
/* Chebyshev roots and polynomials (Kernels/Chebyshev/FChebRoots.hpp)       */
/* ======================================================================== */

/* The reference hard-codes the roots to ~15 significant digits
 * (FChebRoots.hpp:90-297); the same decimal table is used here so that the
 * interpolation nodes are the reference's nodes, not the exact cosines. */
static const double ROOTS_2[] = {-0.707106781186548, 0.707106781186547};
static const double ROOTS_3[] = {-8.66025403784439e-01, 0.0, 8.66025403784438e-01};
static const double ROOTS_4[] = {-0.923879532511287, -0.382683432365090, 0.382683432365090, 0.923879532511287};
static const double ROOTS_5[] = {-9.51056516295154e-01, -5.87785252292473e-01, 0.0, 5.87785252292473e-01,
                                 9.51056516295154e-01};
static const double ROOTS_6[] = {-0.965925826289068, -0.707106781186548, -0.258819045102521,
                                 0.258819045102521,  0.707106781186547,  0.965925826289068};
static const double ROOTS_7[] = {-9.74927912181824e-01, -7.81831482468030e-01, -4.33883739117558e-01, 0.0,
                                 4.33883739117558e-01,  7.81831482468030e-01,  9.74927912181824e-01};
static const double ROOTS_8[] = {-0.980785280403230, -0.831469612302545, -0.555570233019602, -0.195090322016128,
                                 0.195090322016128,  0.555570233019602,  0.831469612302545,  0.980785280403230};
static const double ROOTS_9[] = {-9.84807753012208e-01, -8.66025403784439e-01, -6.42787609686539e-01,
                                 -3.42020143325669e-01, 0.0,                    3.42020143325669e-01,
                                 6.42787609686539e-01,  8.66025403784438e-01,  9.84807753012208e-01};
static const double ROOTS_10[] = {-0.987688340595138, -0.891006524188368, -0.707106781186548, -0.453990499739547,
                                  -0.156434465040231, 0.156434465040231,  0.453990499739547,  0.707106781186547,
                                  0.891006524188368,  0.987688340595138};
static const double ROOTS_11[] = {-9.89821441880933e-01, -9.09631995354518e-01, -7.55749574354258e-01,
                                  -5.40640817455598e-01, -2.81732556841430e-01, 0.0,
                                  2.81732556841430e-01,  5.40640817455597e-01,  7.55749574354258e-01,
                                  9.09631995354518e-01,  9.89821441880933e-01};
static const double ROOTS_12[] = {-0.991444861373810, -0.923879532511287, -0.793353340291235, -0.608761429008721,
                                  -0.382683432365090, -0.130526192220052, 0.130526192220052,  0.382683432365090,
                                  0.608761429008721,  0.793353340291235,  0.923879532511287,  0.991444861373810};
static const double ROOTS_13[] = {-9.92708874098054e-01, -9.35016242685415e-01, -8.22983865893656e-01,
                                  -6.63122658240795e-01, -4.64723172043769e-01, -2.39315664287558e-01,
                                  0.0,                    2.39315664287557e-01,  4.64723172043769e-01,
                                  6.63122658240795e-01,  8.22983865893656e-01,  9.35016242685415e-01,
                                  9.92708874098054e-01};
static const double ROOTS_14[] = {-0.99371220989324258353, -0.94388333030836756290, -0.84672419922828416835,
                                  -0.70710678118654752440, -0.53203207651533656356, -0.33027906195516708177,
                                  -0.11196447610330785847, 0.11196447610330785847,  0.33027906195516708177,
                                  0.53203207651533656356,  0.70710678118654752440,  0.84672419922828416835,
                                  0.94388333030836756290,  0.99371220989324258353};

int orc_cheb_roots(int order, double* roots) {
    const double* t = NULL;
    switch (order) {
        case 2: t = ROOTS_2; break;   case 3: t = ROOTS_3; break;   case 4: t = ROOTS_4; break;
        case 5: t = ROOTS_5; break;   case 6: t = ROOTS_6; break;   case 7: t = ROOTS_7; break;
        case 8: t = ROOTS_8; break;   case 9: t = ROOTS_9; break;   case 10: t = ROOTS_10; break;
        case 11: t = ROOTS_11; break; case 12: t = ROOTS_12; break; case 13: t = ROOTS_13; break;
        case 14: t = ROOTS_14; break;
        default: return ORC_ERR_ARG;
    }
    memcpy(roots, t, sizeof(double) * (size_t)order);
    return ORC_OK;
}

double orc_cheb_T(unsigned n, double x) {
    /* FChebRoots::T (FChebRoots.hpp:46-58) */
    if (fabs(x) > 1.) x = x > 1. ? 1. : (x < -1. ? -1. : x);
    return cos(n * acos(x));
}

double orc_cheb_U(unsigned n, double x) {
    /* FChebRoots::U (FChebRoots.hpp:68-78) */
    if (fabs(x) > 1.) x = x > 1. ? 1. : (x < -1. ? -1. : x);
    return n * (sin(n * acos(x))) / sqrt(1 - x * x);
}

/* ======================================================================== */
/* Dense linear algebra (in-house; the reference calls LAPACK through FBlas) */
/* ======================================================================== */

static double dot(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

int orc_qr(int m, int n, const double* A, double* Q, double* R) {
    /* Householder QR: A = Q R with Q m x n orthonormal columns (the pair
     * geqrf + orgqr of Kernels/Chebyshev/FChebSymM2LHandler.hpp:123-149). */
    if (m < n || n <= 0) return ORC_ERR_ARG;
    double* W = (double*)malloc(sizeof(double) * (size_t)m * n);
    double* tau = (double*)malloc(sizeof(double) * (size_t)n);
    if (!W || !tau) { free(W); free(tau); return ORC_ERR_NOMEM; }
    memcpy(W, A, sizeof(double) * (size_t)m * n);
    for (int j = 0; j < n; ++j) {
        double* col = W + (size_t)j * m;
        double nrm = sqrt(dot(m - j, col + j, col + j));
        if (nrm == 0.0) { tau[j] = 0.0; continue; }
        const double alpha = col[j] >= 0 ? -nrm : nrm;  /* R_jj = alpha */
        const double v0 = col[j] - alpha;
        /* v = [1, col[j+1:]/v0], tau = -v0/alpha */
        for (int i = j + 1; i < m; ++i) col[i] /= v0;
        tau[j] = -v0 / alpha;
        col[j] = alpha;
        for (int k = j + 1; k < n; ++k) {
            double* ck = W + (size_t)k * m;
            double s = ck[j];
            for (int i = j + 1; i < m; ++i) s += col[i] * ck[i];
            s *= tau[j];
            ck[j] -= s;
            for (int i = j + 1; i < m; ++i) ck[i] -= s * col[i];
        }
    }
    if (R) {
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < n; ++i) R[(size_t)k * n + i] = i <= k ? W[(size_t)k * m + i] : 0.0;
    }
    if (Q) {
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < m; ++i) Q[(size_t)k * m + i] = (i == k) ? 1.0 : 0.0;
        for (int j = n - 1; j >= 0; --j) {
            const double* v = W + (size_t)j * m;
            if (tau[j] == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                double* qk = Q + (size_t)k * m;
                double s = qk[j];
                for (int i = j + 1; i < m; ++i) s += v[i] * qk[i];
                s *= tau[j];
                qk[j] -= s;
                for (int i = j + 1; i < m; ++i) qk[i] -= s * v[i];
            }
        }
    }
    free(W); free(tau);
    return ORC_OK;
}

int orc_svd(int n, const double* A, double* U, double* S, double* VT) {
    /* One-sided (Hestenes) Jacobi SVD: rotate column pairs of W = A V until
     * orthogonal; S_i = |W_i|, U_i = W_i / S_i.  Stands in for dgesvd with
     * jobu='O', jobvt='S' (FChebSymM2LHandler.hpp:154-162). */
    if (n <= 0) return ORC_ERR_ARG;
    double* W = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* V = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* nrm = (double*)malloc(sizeof(double) * (size_t)n);
    int* ord = (int*)malloc(sizeof(int) * (size_t)n);
    if (!W || !V || !nrm || !ord) { free(W); free(V); free(nrm); free(ord); return ORC_ERR_NOMEM; }
    memcpy(W, A, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) V[(size_t)j * n + i] = (i == j);
    const double tol = 1e-15;
    int converged = 0;
    for (int sweep = 0; sweep < 80 && !converged; ++sweep) {
        converged = 1;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double* wp = W + (size_t)p * n;
                double* wq = W + (size_t)q * n;
                const double alpha = dot(n, wp, wp);
                const double beta = dot(n, wq, wq);
                const double gamma = dot(n, wp, wq);
                if (gamma == 0.0 || fabs(gamma) <= tol * sqrt(alpha * beta)) continue;
                converged = 0;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = c * t;
                for (int i = 0; i < n; ++i) {
                    const double a = wp[i], b = wq[i];
                    wp[i] = c * a - s * b;
                    wq[i] = s * a + c * b;
                }
                double* vp = V + (size_t)p * n;
                double* vq = V + (size_t)q * n;
                for (int i = 0; i < n; ++i) {
                    const double a = vp[i], b = vq[i];
                    vp[i] = c * a - s * b;
                    vq[i] = s * a + c * b;
                }
            }
        }
    }
    for (int j = 0; j < n; ++j) { nrm[j] = sqrt(dot(n, W + (size_t)j * n, W + (size_t)j * n)); ord[j] = j; }
    /* sort descending (insertion sort, stable) */
    for (int a = 1; a < n; ++a) {
        int v = ord[a], b = a - 1;
        while (b >= 0 && nrm[ord[b]] < nrm[v]) { ord[b + 1] = ord[b]; --b; }
        ord[b + 1] = v;
    }
    for (int r = 0; r < n; ++r) {
        const int j = ord[r];
        S[r] = nrm[j];
        for (int i = 0; i < n; ++i) {
            if (U) U[(size_t)r * n + i] = nrm[j] > 0 ? W[(size_t)j * n + i] / nrm[j] : 0.0;
            if (VT) VT[(size_t)i * n + r] = V[(size_t)j * n + i];
        }
    }
    free(W); free(V); free(nrm); free(ord);
    return converged ? ORC_OK : ORC_ERR_NUMERIC;
}

int orc_svd_rank(const double* S, int n, double eps) {
    /* FSvd::getRank (Utils/FSvd.hpp:51-62) */
    const double nrm2 = dot(n, S, S);
    double nrm2k = 0.;
    for (int k = n; k > 0; --k) {
        nrm2k += S[k - 1] * S[k - 1];
        if (nrm2k > eps * eps * nrm2) return k;
    }
    return -1;
}

/* ======================================================================== */
/* Chebyshev operator set                                                   */
/* ======================================================================== */

struct orc_cheb {
    int p, nn;
    int basis;  /* 0: Chebyshev (FChebInterpolator), 1: Lagrange on equispaced roots (FUnifInterpolator) */
    double eps;
    double roots[ORC_MAX_ORDER];
    double Troots[ORC_MAX_ORDER][ORC_MAX_ORDER]; /* [o][j] = T_o(root_j), o>=1 */
    double* child;                                /* [8][3][p*p] */
    uint32_t pindices[343];
    uint32_t* pvectors;                           /* [343][nn] */
    int rank[343];
    double* K[343];                               /* [U | V], column-major */
};

int orc_cheb_order(const orc_cheb* c) { return c->p; }

/* FInterpSymmetries (Kernels/Interpolation/FInterpSymmetries.hpp:18-128) */
static void build_symmetries(orc_cheb* c) {
    const int p = c->p, nn = c->nn;
    unsigned* quads = (unsigned*)malloc(sizeof(unsigned) * 8 * nn);
    unsigned* cones = (unsigned*)malloc(sizeof(unsigned) * 8 * nn);
    unsigned* perms = (unsigned*)malloc(sizeof(unsigned) * 64 * nn);
    unsigned evn[ORC_MAX_ORDER], odd[ORC_MAX_ORDER];
    for (int o = 0; o < p; ++o) { evn[o] = o; odd[o] = p - 1 - o; }
    const unsigned P = p, PP = p * p;
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j)
            for (int k = 0; k < p; ++k) {
                const int idx = k * p * p + j * p + i;
                quads[0 * nn + idx] = evn[k] * PP + evn[j] * P + evn[i];
                quads[1 * nn + idx] = evn[k] * PP + evn[j] * P + odd[i];
                quads[2 * nn + idx] = evn[k] * PP + odd[j] * P + evn[i];
                quads[3 * nn + idx] = evn[k] * PP + odd[j] * P + odd[i];
                quads[4 * nn + idx] = odd[k] * PP + evn[j] * P + evn[i];
                quads[5 * nn + idx] = odd[k] * PP + evn[j] * P + odd[i];
                quads[6 * nn + idx] = odd[k] * PP + odd[j] * P + evn[i];
                quads[7 * nn + idx] = odd[k] * PP + odd[j] * P + odd[i];
                cones[0 * nn + idx] = evn[k] * PP + evn[j] * P + evn[i];
                cones[1 * nn + idx] = evn[j] * PP + evn[k] * P + evn[i];
                cones[2 * nn + idx] = 0;
                cones[3 * nn + idx] = evn[j] * PP + evn[i] * P + evn[k];
                cones[4 * nn + idx] = evn[k] * PP + evn[i] * P + evn[j];
                cones[5 * nn + idx] = 0;
                cones[6 * nn + idx] = evn[i] * PP + evn[k] * P + evn[j];
                cones[7 * nn + idx] = evn[i] * PP + evn[j] * P + evn[k];
            }
    for (int q = 0; q < 8; ++q)
        for (int cc = 0; cc < 8; ++cc)
            if (cc != 2 && cc != 5)
                for (int n = 0; n < nn; ++n) perms[(q * 8 + cc) * nn + n] = cones[cc * nn + quads[q * nn + n]];
    static const int pm[8][3] = {{0, 1, 2}, {0, 2, 1}, {0, 0, 0}, {2, 0, 1}, {1, 0, 2}, {0, 0, 0}, {1, 2, 0}, {2, 1, 0}};
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k) {
                const int idx = ((i + 3) * 7 + (j + 3)) * 7 + (k + 3);
                c->pindices[idx] = 0;
                for (int n = 0; n < nn; ++n) c->pvectors[(size_t)idx * nn + n] = 0;
                if (abs(i) > 1 || abs(j) > 1 || abs(k) > 1) {
                    /* getPermutationArrayAndIndex (FInterpSymmetries.hpp:98-125) */
                    const unsigned qidx = (unsigned)((i < 0 ? 1 : 0) | (j < 0 ? 2 : 0) | (k < 0 ? 4 : 0));
                    const int u[3] = {abs(i), abs(j), abs(k)};
                    const int cidx = ((u[1] > u[0]) << 2) | ((u[2] > u[0]) << 1) | (u[2] > u[1]);
                    for (int n = 0; n < nn; ++n) c->pvectors[(size_t)idx * nn + n] = perms[(qidx * 8 + cidx) * nn + n];
                    c->pindices[idx] = (unsigned)((u[pm[cidx][0]] + 3) * 49 + (u[pm[cidx][1]] + 3) * 7 + (u[pm[cidx][2]] + 3));
                }
            }
    free(quads); free(cones); free(perms);
}

/* FChebTensor::setRoots (FInterpTensor.hpp:75-87) with map_loc_glob (FInterpMapping.hpp:134-148) */
static void set_roots(const orc_cheb* c, const double center[3], double width, double* X /* nn*3 */) {
    const int p = c->p;
    const double a[3] = {center[0] - width / 2., center[1] - width / 2., center[2] - width / 2.};
    const double b[3] = {center[0] + width / 2., center[1] + width / 2., center[2] + width / 2.};
    for (int n = 0; n < c->nn; ++n) {
        const int id[3] = {n % p, (n / p) % p, n / (p * p)};
        for (int d = 0; d < 3; ++d) X[3 * n + d] = (a[d] + b[d]) / 2. + (b[d] - a[d]) * c->roots[id[d]] / 2.;
    }
}

/* 1/r (FInterpMatrixKernel.hpp:79-87): evaluate(t, s) */
static double kernel_r(const double* t, const double* s) {
    const double dx = t[0] - s[0], dy = t[1] - s[1], dz = t[2] - s[2];
    return 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
}

typedef struct {
    int nn;
    const double* X; /* target roots (the EntryComputer's "ps") */
    const double* Y; /* source roots (the EntryComputer's "pt") */
    const double* w;
} entry_computer;

/* EntryComputer::operator() (FInterpMatrixKernel.hpp:578-590): for j in
 * [tb,te) (outer), i in [sb,se) (inner): w_i w_j K(Y_i, X_j). */
static void entry_compute(const entry_computer* e, int tb, int te, int sb, int se, double* data) {
    int idx = 0;
    for (int j = tb; j < te; ++j)
        for (int i = sb; i < se; ++i) data[idx++] = e->w[i] * e->w[j] * kernel_r(e->Y + 3 * i, e->X + 3 * j);
}

/* FAca::pACA (Utils/FAca.hpp:141-226) */
static int paca(const entry_computer* e, int nx, int ny, double eps, double** Uo, double** Vo, int* ko) {
    char* r = (char*)malloc((size_t)nx);
    char* cu = (char*)malloc((size_t)ny);
    const int maxk = (nx + ny) / 2;
    double* U = (double*)malloc(sizeof(double) * (size_t)nx * maxk);
    double* V = (double*)malloc(sizeof(double) * (size_t)ny * maxk);
    if (!r || !cu || !U || !V) { free(r); free(cu); free(U); free(V); return ORC_ERR_NOMEM; }
    for (int i = 0; i < nx; ++i) r[i] = 1;
    for (int j = 0; j < ny; ++j) cu[j] = 1;
    int k = 0;
    const double eps2 = eps * eps;
    double norm2S = 0., norm2uv = 0.;
    int J = 0, I = 0;
    do {
        if (k >= maxk) { free(r); free(cu); free(U); free(V); return ORC_ERR_NUMERIC; }
        double* colU = U + (size_t)nx * k;
        double* colV = V + (size_t)ny * k;
        entry_compute(e, I, I + 1, 0, ny, colV);
        r[I] = 0;
        for (int l = 0; l < k; ++l) {
            const double* u = U + (size_t)nx * l;
            const double* v = V + (size_t)ny * l;
            const double a = -1. * u[I];
            for (int j = 0; j < ny; ++j) colV[j] += a * v[j];
        }
        double maxval = 0.;
        for (int j = 0; j < ny; ++j) {
            const double av = fabs(colV[j]);
            if (cu[j] && maxval < av) { maxval = av; J = j; }
        }
        const double pivot = 1. / colV[J];
        for (int j = 0; j < ny; ++j) colV[j] *= pivot;
        entry_compute(e, 0, nx, J, J + 1, colU);
        cu[J] = 0;
        for (int l = 0; l < k; ++l) {
            const double* u = U + (size_t)nx * l;
            const double* v = V + (size_t)ny * l;
            const double a = -1. * v[J];
            for (int i = 0; i < nx; ++i) colU[i] += a * u[i];
        }
        maxval = 0.;
        for (int i = 0; i < nx; ++i) {
            const double av = fabs(colU[i]);
            if (r[i] && maxval < av) { maxval = av; I = i; }
        }
        double normuuvv = 0.;
        for (int l = 0; l < k; ++l) normuuvv += dot(nx, colU, U + (size_t)nx * l) * dot(ny, V + (size_t)ny * l, colV);
        norm2uv = dot(nx, colU, colU) * dot(ny, colV, colV);
        norm2S += norm2uv + 2 * normuuvv;
        ++k;
    } while (norm2uv > eps2 * norm2S);
    free(r); free(cu);
    *Uo = U; *Vo = V; *ko = k;
    return ORC_OK;
}

/* precompute (Kernels/Chebyshev/FChebSymM2LHandler.hpp:39-281), partially
 * pivoted ACA + QR + SVD recompression, reference cell width 2. */
static int precompute(orc_cheb* c) {
    const int p = c->p, nn = c->nn;
    const double cw = 2.0;
    double* X = (double*)malloc(sizeof(double) * 3 * nn);
    double* Y = (double*)malloc(sizeof(double) * 3 * nn);
    double* w = (double*)malloc(sizeof(double) * nn);
    const double zero[3] = {0, 0, 0};
    set_roots(c, zero, cw, X);
    /* FChebTensor::setRootOfWeights (FChebTensor.hpp:42-55) */
    double w1[ORC_MAX_ORDER];
    for (int o = 0; o < p; ++o) w1[o] = M_PI / p * sqrt(1. - c->roots[o] * c->roots[o]);
    for (int n = 0; n < nn; ++n) w[n] = sqrt(w1[n % p] * w1[(n / p) % p] * w1[n / (p * p)]);
    int status = ORC_OK;
    for (int i = 2; i <= 3 && status == ORC_OK; ++i)
        for (int j = 0; j <= i && status == ORC_OK; ++j)
            for (int k = 0; k <= j && status == ORC_OK; ++k) {
                const double cy[3] = {cw * i, cw * j, cw * k};
                set_roots(c, cy, cw, Y);
                entry_computer e = {nn, X, Y, w};
                double *UU = NULL, *VV = NULL;
                int rank = 0;
                status = paca(&e, nn, nn, c->eps, &UU, &VV, &rank);
                if (status) break;
                /* QR of U and V, phi = R_U R_V^T */
                double* QU = (double*)malloc(sizeof(double) * (size_t)nn * rank);
                double* QV = (double*)malloc(sizeof(double) * (size_t)nn * rank);
                double* RU = (double*)malloc(sizeof(double) * (size_t)rank * rank);
                double* RV = (double*)malloc(sizeof(double) * (size_t)rank * rank);
                double* phi = (double*)malloc(sizeof(double) * (size_t)rank * rank);
                double* Uphi = (double*)malloc(sizeof(double) * (size_t)rank * rank);
                double* VT = (double*)malloc(sizeof(double) * (size_t)rank * rank);
                double* S = (double*)malloc(sizeof(double) * (size_t)rank);
                orc_qr(nn, rank, UU, QU, RU);
                orc_qr(nn, rank, VV, QV, RV);
                for (int b = 0; b < rank; ++b)
                    for (int a = 0; a < rank; ++a) {
                        double s = 0;
                        for (int l = 0; l < rank; ++l) s += RU[(size_t)l * rank + a] * RV[(size_t)l * rank + b];
                        phi[(size_t)b * rank + a] = s;
                    }
                status = orc_svd(rank, phi, Uphi, S, VT);
                const int aca_rank = rank;
                if (status == ORC_OK) {
                    rank = orc_svd_rank(S, aca_rank, c->eps);
                    if (rank <= 0) status = ORC_ERR_NUMERIC;
                }
                if (status == ORC_OK) {
                    const int idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3);
                    double* K = (double*)calloc((size_t)2 * rank * nn, sizeof(double));
                    c->K[idx] = K;
                    c->rank[idx] = rank;
                    /* K_U = Q_U (U_phi Sigma), K_V = Q_V V_phi (FChebSymM2LHandler.hpp:165-189) */
                    for (int r2 = 0; r2 < rank; ++r2)
                        for (int a = 0; a < aca_rank; ++a) Uphi[(size_t)r2 * aca_rank + a] *= S[r2];
                    for (int r2 = 0; r2 < rank; ++r2)
                        for (int n = 0; n < nn; ++n) {
                            double su = 0, sv = 0;
                            for (int a = 0; a < aca_rank; ++a) {
                                su += QU[(size_t)a * nn + n] * Uphi[(size_t)r2 * aca_rank + a];
                                sv += QV[(size_t)a * nn + n] * VT[(size_t)a * aca_rank + r2];
                            }
                            K[(size_t)r2 * nn + n] = su;
                            K[(size_t)rank * nn + (size_t)r2 * nn + n] = sv;
                        }
                    /* un-weighting (FChebSymM2LHandler.hpp:258-261) */
                    for (int n = 0; n < nn; ++n) {
                        const double iw = 1. / w[n];
                        for (int r2 = 0; r2 < rank; ++r2) {
                            K[(size_t)r2 * nn + n] *= iw;
                            K[(size_t)rank * nn + (size_t)r2 * nn + n] *= iw;
                        }
                    }
                }
                free(QU); free(QV); free(RU); free(RV); free(phi); free(Uphi); free(VT); free(S);
                free(UU); free(VV);
            }
    free(X); free(Y); free(w);
    return status;
}

int orc_cheb_create(int order, double epsilon, orc_cheb** out) {
    if (order < ORC_MIN_ORDER || order > ORC_MAX_ORDER || !out) return ORC_ERR_ARG;
    orc_cheb* c = (orc_cheb*)calloc(1, sizeof(orc_cheb));
    if (!c) return ORC_ERR_NOMEM;
    c->p = order;
    c->nn = order * order * order;
    /* default epsilon 10^-ORDER (FChebSymKernel.hpp:136-140) */
    c->eps = epsilon > 0 ? epsilon : pow(10.0, (double)(-order));
    orc_cheb_roots(order, c->roots);
    for (int o = 1; o < order; ++o)
        for (int j = 0; j < order; ++j) c->Troots[o][j] = orc_cheb_T((unsigned)o, c->roots[j]);
    /* child/parent interpolators at level 2 (FChebInterpolator.hpp:300-338, 466-485):
     * ExtendedCellRatio = 0.5, child width 1, child centre +-0.5 per axis. */
    const int p = order;
    c->child = (double*)malloc(sizeof(double) * 8 * 3 * p * p);
    static const int rel[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                                  {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
    for (int ch = 0; ch < 8; ++ch) {
        for (int d = 0; d < 3; ++d) {
            const double ctr = (double)rel[ch][d] * (1. - .5);
            const double a = ctr - 1. / 2., b = ctr + 1. / 2.;
            double* S = c->child + ((size_t)ch * 3 + d) * p * p;
            for (int m = 0; m < p; ++m) {
                const double xm = (a + b) / 2. + (b - a) * c->roots[m] / 2.;
                double Tx[ORC_MAX_ORDER];
                for (int o = 1; o < p; ++o) Tx[o] = orc_cheb_T((unsigned)o, xm);
                for (int n = 0; n < p; ++n) {
                    double s = 1. / p;
                    for (int o = 1; o < p; ++o) s += 2. / p * Tx[o] * c->Troots[o][n];
                    S[n * p + m] = s;
                }
            }
        }
    }
    c->pvectors = (uint32_t*)malloc(sizeof(uint32_t) * 343 * (size_t)c->nn);
    build_symmetries(c);
    const int st = precompute(c);
    if (st) { orc_cheb_destroy(c); return st; }
    *out = c;
    return ORC_OK;
}

void orc_cheb_destroy(orc_cheb* c) {
    if (!c) return;
    for (int i = 0; i < 343; ++i) free(c->K[i]);
    free(c->child); free(c->pvectors); free(c);
}

/* ======================================================================== */
/* FChebKernel's M2L: one global SVD of the 316 far-field operators         */
/* (Kernels/Chebyshev/FChebM2LHandler.hpp:216-315 ComputeAndCompress,       */
/* :437-523 Compress)                                                      */
/* ======================================================================== */

struct orc_chebg {
    orc_cheb* c;  /* roots, interpolators (P2M, M2M, L2L, L2P are the symmetric kernel's) */
    int k;        /* rank */
    double* U;    /* nn x k, column-major, un-weighted (target side, applyU) */
    double* B;    /* nn x k, column-major, un-weighted (source side, applyB) */
    double* C;    /* [343][k x k] column-major: C_t[i][j] at i + j k; zero for near positions */
    double* S1;   /* singular values of K_col (source side) */
    double* S2;   /* singular values of K_row (target side) */
};

/* Right singular vectors and singular values of a tall m x n matrix (column-
 * major): Householder QR, then the one-sided Jacobi SVD of R.  Stands in for
 * dgesvd on K_col and dgesvd "SO" on K_row (FChebM2LHandler.hpp:459-484); the
 * right singular vectors of K_row^T are the left ones of K_row. */
static int tall_right_svd(int m, int n, const double* A, double* S, double* V /* n x n, column r = vector r */) {
    double* R = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* VT = (double*)malloc(sizeof(double) * (size_t)n * n);
    if (!R || !VT) { free(R); free(VT); return ORC_ERR_NOMEM; }
    int st = orc_qr(m, n, A, NULL, R);
    if (st == ORC_OK) st = orc_svd(n, R, NULL, S, VT);
    if (st == ORC_OK)
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) V[(size_t)r * n + c] = VT[(size_t)c * n + r];
    free(R); free(VT);
    return st;
}

int orc_chebg_create(int order, double epsilon, orc_chebg** out) {
    if (!out) return ORC_ERR_ARG;
    orc_chebg* g = (orc_chebg*)calloc(1, sizeof(orc_chebg));
    if (!g) return ORC_ERR_NOMEM;
    int st = orc_cheb_create(order, epsilon, &g->c);
    if (st) { free(g); return st; }
    const orc_cheb* c = g->c;
    const int p = c->p, nn = c->nn, nint = 316;
    const double eps = c->eps;
    const size_t m1 = (size_t)nint * nn;
    double* X = (double*)malloc(sizeof(double) * 3 * nn);
    double* Y = (double*)malloc(sizeof(double) * 3 * nn);
    double* w = (double*)malloc(sizeof(double) * nn);
    double* Kw = (double*)malloc(sizeof(double) * m1 * nn);  /* [t][n][m] = w_n w_m K(X_m, Y_n) */
    double* T = (double*)malloc(sizeof(double) * m1 * nn);
    double* V1 = (double*)malloc(sizeof(double) * (size_t)nn * nn);
    double* V2 = (double*)malloc(sizeof(double) * (size_t)nn * nn);
    g->S1 = (double*)malloc(sizeof(double) * nn);
    g->S2 = (double*)malloc(sizeof(double) * nn);
    if (!X || !Y || !w || !Kw || !T || !V1 || !V2 || !g->S1 || !g->S2) st = ORC_ERR_NOMEM;
    if (st == ORC_OK) {
        const double zero[3] = {0, 0, 0};
        set_roots(c, zero, 2.0, X);
        double w1[ORC_MAX_ORDER];
        for (int o = 0; o < p; ++o) w1[o] = M_PI / p * sqrt(1. - c->roots[o] * c->roots[o]);
        for (int n = 0; n < nn; ++n) w[n] = sqrt(w1[n % p] * w1[(n / p) % p] * w1[n / (p * p)]);
        /* the 316 operators in the loop order i, j, k (FChebM2LHandler.hpp:236-254) */
        size_t t = 0;
        for (int i = -3; i <= 3; ++i)
            for (int j = -3; j <= 3; ++j)
                for (int kk = -3; kk <= 3; ++kk) {
                    if (abs(i) <= 1 && abs(j) <= 1 && abs(kk) <= 1) continue;
                    const double cy[3] = {2. * i, 2. * j, 2. * kk};
                    set_roots(c, cy, 2.0, Y);
                    for (int n = 0; n < nn; ++n)
                        for (int m = 0; m < nn; ++m)
                            Kw[(t * nn + n) * nn + m] = w[n] * w[m] * kernel_r(X + 3 * m, Y + 3 * n);
                    ++t;
                }
        /* K_col: rows (t, m), column n (FChebM2LHandler.hpp:448-456) -> B */
        for (size_t tt = 0; tt < (size_t)nint; ++tt)
            for (int n = 0; n < nn; ++n)
                for (int m = 0; m < nn; ++m) T[(size_t)n * m1 + tt * nn + m] = Kw[(tt * nn + n) * nn + m];
        st = tall_right_svd((int)m1, nn, T, g->S1, V1);
    }
    if (st == ORC_OK) {
        /* K_row^T: rows (t, n), column m -> U */
        for (size_t tt = 0; tt < (size_t)nint; ++tt)
            for (int n = 0; n < nn; ++n)
                for (int m = 0; m < nn; ++m) T[(size_t)m * m1 + tt * nn + n] = Kw[(tt * nn + n) * nn + m];
        st = tall_right_svd((int)m1, nn, T, g->S2, V2);
    }
    if (st == ORC_OK) {
        const int k1 = orc_svd_rank(g->S1, nn, eps), k2 = orc_svd_rank(g->S2, nn, eps);
        if (k1 <= 0 || k2 <= 0) st = ORC_ERR_NUMERIC;
        g->k = k1 < k2 ? k1 : k2;
    }
    if (st == ORC_OK) {
        const int k = g->k;
        g->U = (double*)malloc(sizeof(double) * (size_t)nn * k);
        g->B = (double*)malloc(sizeof(double) * (size_t)nn * k);
        g->C = (double*)calloc((size_t)343 * k * k, sizeof(double));
        double* G = (double*)malloc(sizeof(double) * (size_t)nn * k);
        if (!g->U || !g->B || !g->C || !G) st = ORC_ERR_NOMEM;
        if (st == ORC_OK) {
            memcpy(g->U, V2, sizeof(double) * (size_t)nn * k);
            memcpy(g->B, V1, sizeof(double) * (size_t)nn * k);
            /* C_t = U^T A_t B with A_t[m][n] = Kw[t][n][m] (the weighted operators) */
            size_t t = 0;
            for (int i = -3; i <= 3; ++i)
                for (int j = -3; j <= 3; ++j)
                    for (int kk = -3; kk <= 3; ++kk) {
                        if (abs(i) <= 1 && abs(j) <= 1 && abs(kk) <= 1) continue;
                        const double* Kt = Kw + t * nn * nn;
                        for (int n = 0; n < nn; ++n)
                            for (int a = 0; a < k; ++a) G[(size_t)a * nn + n] = dot(nn, Kt + (size_t)n * nn, g->U + (size_t)a * nn);
                        double* Ct = g->C + (size_t)((i + 3) * 49 + (j + 3) * 7 + (kk + 3)) * k * k;
                        for (int b = 0; b < k; ++b)
                            for (int a = 0; a < k; ++a) Ct[a + (size_t)b * k] = dot(nn, G + (size_t)a * nn, g->B + (size_t)b * nn);
                        ++t;
                    }
            /* un-weighting (FChebM2LHandler.hpp:299-303) */
            for (int a = 0; a < k; ++a)
                for (int n = 0; n < nn; ++n) {
                    g->U[(size_t)a * nn + n] /= w[n];
                    g->B[(size_t)a * nn + n] /= w[n];
                }
        }
        free(G);
    }
    free(X); free(Y); free(w); free(Kw); free(T); free(V1); free(V2);
    if (st != ORC_OK) { orc_chebg_destroy(g); return st; }
    *out = g;
    return ORC_OK;
}

/* FChebDenseKernel's operators (Kernels/Chebyshev/FChebDenseM2LHandler.hpp:180-215):
 * K_t[m][n] = K(X_m, Y_n) uncompressed, as a "global" set of rank p^3 with
 * identity bases, so the same passes apply them. */
int orc_chebd_create(int order, orc_chebg** out) {
    if (!out) return ORC_ERR_ARG;
    orc_chebg* g = (orc_chebg*)calloc(1, sizeof(orc_chebg));
    if (!g) return ORC_ERR_NOMEM;
    int st = orc_cheb_create(order, 0.0, &g->c);
    if (st) { free(g); return st; }
    const orc_cheb* c = g->c;
    const int nn = c->nn;
    g->k = nn;
    g->U = (double*)calloc((size_t)nn * nn, sizeof(double));
    g->B = (double*)calloc((size_t)nn * nn, sizeof(double));
    g->C = (double*)calloc((size_t)343 * nn * nn, sizeof(double));
    g->S1 = (double*)calloc((size_t)nn, sizeof(double));
    g->S2 = (double*)calloc((size_t)nn, sizeof(double));
    double* X = (double*)malloc(sizeof(double) * 3 * nn);
    double* Y = (double*)malloc(sizeof(double) * 3 * nn);
    if (!g->U || !g->B || !g->C || !g->S1 || !g->S2 || !X || !Y) {
        free(X); free(Y); orc_chebg_destroy(g); return ORC_ERR_NOMEM;
    }
    for (int n = 0; n < nn; ++n) g->U[(size_t)n * nn + n] = g->B[(size_t)n * nn + n] = 1.0;
    const double zero[3] = {0, 0, 0};
    set_roots(c, zero, 2.0, X);
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int kk = -3; kk <= 3; ++kk) {
                if (abs(i) <= 1 && abs(j) <= 1 && abs(kk) <= 1) continue;
                const double cy[3] = {2. * i, 2. * j, 2. * kk};
                set_roots(c, cy, 2.0, Y);
                double* Ct = g->C + (size_t)((i + 3) * 49 + (j + 3) * 7 + (kk + 3)) * nn * nn;
                for (int n = 0; n < nn; ++n)
                    for (int m = 0; m < nn; ++m) Ct[m + (size_t)n * nn] = kernel_r(X + 3 * m, Y + 3 * n);
            }
    free(X); free(Y);
    *out = g;
    return ORC_OK;
}

static double unif_L(int p, int n, double x);

/* The equispaced roots of FUnifRootsCore (Kernels/Uniform/FUnifRoots.hpp:
 * 130-240), as tabulated there (decimal data). */
static int unif_roots(int p, double* r) {
    static const double r2[2] = {-1., 1.};
    static const double r3[3] = {-1., 0., 1.};
    static const double r4[4] = {-1., -0.333333333333333, 0.333333333333333, 1.};
    static const double r5[5] = {-1., -0.5, 0., 0.5, 1.};
    static const double r6[6] = {-1., -0.6, -0.2, 0.2, 0.6, 1.};
    static const double r7[7] = {-1., -0.666666666666666, -0.333333333333333, 0., 0.333333333333333,
                                 0.666666666666666, 1.};
    static const double r8[8] = {-1., -0.714285714285714, -0.428571428571429, -0.142857142857143,
                                 0.142857142857143, 0.428571428571429, 0.714285714285714, 1.};
    const double* t = p == 2 ? r2 : p == 3 ? r3 : p == 4 ? r4 : p == 5 ? r5 : p == 6 ? r6 : p == 7 ? r7 : p == 8 ? r8 : NULL;
    if (!t) return ORC_ERR_ARG;
    memcpy(r, t, sizeof(double) * p);
    return ORC_OK;
}

/* FUnifKernel (Kernels/Uniform/FUnifKernel.hpp:72-175): Lagrange interpolation
 * on the equispaced roots (P2M, M2M, L2L, L2P: FUnifInterpolator.hpp), the
 * M2L applying the kernel between the equispaced roots of target and source
 * cell exactly (the reference does it by a zero-padded DFT,
 * FUnifM2LHandler.hpp, which computes the same product); as a rank-p^3 set
 * with identity bases. */
int orc_unif_create(int order, orc_chebg** out) {
    if (!out) return ORC_ERR_ARG;
    orc_chebg* g = (orc_chebg*)calloc(1, sizeof(orc_chebg));
    if (!g) return ORC_ERR_NOMEM;
    orc_cheb* c = (orc_cheb*)calloc(1, sizeof(orc_cheb));
    if (!c) { free(g); return ORC_ERR_NOMEM; }
    g->c = c;
    const int p = order;
    c->p = p;
    c->nn = p * p * p;
    c->basis = 1;
    c->eps = 0.;
    if (unif_roots(p, c->roots)) { orc_chebg_destroy(g); return ORC_ERR_ARG; }
    /* child/parent interpolators (FUnifInterpolator.hpp:100-123): child width 1,
     * centre +-0.5, S[n][m] = L_n(child root m) */
    c->child = (double*)malloc(sizeof(double) * 8 * 3 * p * p);
    c->pvectors = (uint32_t*)calloc((size_t)343 * c->nn, sizeof(uint32_t));
    static const int rel[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                                  {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
    for (int ch = 0; ch < 8; ++ch)
        for (int d = 0; d < 3; ++d) {
            const double ctr = (double)rel[ch][d] * 0.5;
            double* S = c->child + ((size_t)ch * 3 + d) * p * p;
            for (int m = 0; m < p; ++m) {
                const double xm = ctr + c->roots[m] / 2.;
                for (int n = 0; n < p; ++n) S[n * p + m] = unif_L(p, n, xm);
            }
        }
    const int nn = c->nn;
    g->k = nn;
    g->U = (double*)calloc((size_t)nn * nn, sizeof(double));
    g->B = (double*)calloc((size_t)nn * nn, sizeof(double));
    g->C = (double*)calloc((size_t)343 * nn * nn, sizeof(double));
    g->S1 = (double*)calloc((size_t)nn, sizeof(double));
    g->S2 = (double*)calloc((size_t)nn, sizeof(double));
    double* X = (double*)malloc(sizeof(double) * 3 * nn);
    double* Y = (double*)malloc(sizeof(double) * 3 * nn);
    if (!c->child || !c->pvectors || !g->U || !g->B || !g->C || !g->S1 || !g->S2 || !X || !Y) {
        free(X); free(Y); orc_chebg_destroy(g); return ORC_ERR_NOMEM;
    }
    for (int n = 0; n < nn; ++n) g->U[(size_t)n * nn + n] = g->B[(size_t)n * nn + n] = 1.0;
    const double zero[3] = {0, 0, 0};
    set_roots(c, zero, 2.0, X);
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            for (int kk = -3; kk <= 3; ++kk) {
                if (abs(i) <= 1 && abs(j) <= 1 && abs(kk) <= 1) continue;
                const double cy[3] = {2. * i, 2. * j, 2. * kk};
                set_roots(c, cy, 2.0, Y);
                double* Ct = g->C + (size_t)((i + 3) * 49 + (j + 3) * 7 + (kk + 3)) * nn * nn;
                for (int n = 0; n < nn; ++n)
                    for (int m = 0; m < nn; ++m) Ct[m + (size_t)n * nn] = kernel_r(X + 3 * m, Y + 3 * n);
            }
    free(X); free(Y);
    *out = g;
    return ORC_OK;
}

void orc_chebg_destroy(orc_chebg* g) {
    if (!g) return;
    orc_cheb_destroy(g->c);
    free(g->U); free(g->B); free(g->C); free(g->S1); free(g->S2);
    free(g);
}

int orc_chebg_rank(const orc_chebg* g) { return g->k; }
const orc_cheb* orc_chebg_cheb(const orc_chebg* g) { return g->c; }

/* Replaces the rank and the operators U, B (nn x k, column-major, un-weighted)
 * and C ([343][k x k] column-major) of a handle: the tests apply operators
 * built from another basis of the same singular subspaces (the choice inside
 * a degenerate singular value that a truncation splits is the SVD
 * implementation's, FChebM2LHandler.hpp:459-484). */
int orc_chebg_set(orc_chebg* g, int k, const double* U, const double* B, const double* C) {
    if (!g || k <= 0 || k > g->c->nn || !U || !B || !C) return ORC_ERR_ARG;
    const int nn = g->c->nn;
    double* nU = (double*)malloc(sizeof(double) * (size_t)nn * k);
    double* nB = (double*)malloc(sizeof(double) * (size_t)nn * k);
    double* nC = (double*)malloc(sizeof(double) * (size_t)343 * k * k);
    if (!nU || !nB || !nC) { free(nU); free(nB); free(nC); return ORC_ERR_NOMEM; }
    memcpy(nU, U, sizeof(double) * (size_t)nn * k);
    memcpy(nB, B, sizeof(double) * (size_t)nn * k);
    memcpy(nC, C, sizeof(double) * (size_t)343 * k * k);
    free(g->U); free(g->B); free(g->C);
    g->U = nU; g->B = nB; g->C = nC; g->k = k;
    return ORC_OK;
}

void orc_chebg_get(const orc_chebg* g, double* U, double* B, double* C, double* S1, double* S2) {
    const int nn = g->c->nn, k = g->k;
    if (U) memcpy(U, g->U, sizeof(double) * (size_t)nn * k);
    if (B) memcpy(B, g->B, sizeof(double) * (size_t)nn * k);
    if (C) memcpy(C, g->C, sizeof(double) * (size_t)343 * k * k);
    if (S1) memcpy(S1, g->S1, sizeof(double) * nn);
    if (S2) memcpy(S2, g->S2, sizeof(double) * nn);
}

int orc_cheb_rank(const orc_cheb* c, int idx) { return (idx >= 0 && idx < 343) ? c->rank[idx] : 0; }

int orc_cheb_get_K(const orc_cheb* c, int idx, double* out) {
    if (idx < 0 || idx >= 343 || !c->K[idx]) return ORC_ERR_ARG;
    memcpy(out, c->K[idx], sizeof(double) * 2 * (size_t)c->rank[idx] * c->nn);
    return ORC_OK;
}

void orc_cheb_get_symmetries(const orc_cheb* c, uint32_t* pindices, uint32_t* pvectors) {
    if (pindices) memcpy(pindices, c->pindices, sizeof(c->pindices));
    if (pvectors) memcpy(pvectors, c->pvectors, sizeof(uint32_t) * 343 * (size_t)c->nn);
}

void orc_cheb_get_child_interp(const orc_cheb* c, double* out) {
    memcpy(out, c->child, sizeof(double) * 8 * 3 * (size_t)c->p * c->p);
}

int orc_cheb_apply_m2l(const orc_cheb* c, int idx, double scale, const double* M, double* L) {
    /* FChebSymKernel::M2L for one source (FChebSymKernel.hpp:215-325):
     * mul[pvec[n]] = M[n]; C = K_V^T mul; loc = scale K_U C; L[n] += loc[pvec[n]] */
    if (idx < 0 || idx >= 343) return ORC_ERR_ARG;
    const int cidx = (int)c->pindices[idx];
    const int rank = c->rank[cidx];
    const double* K = c->K[cidx];
    if (!K) return ORC_ERR_ARG;
    const int nn = c->nn;
    const uint32_t* pv = c->pvectors + (size_t)idx * nn;
    double* mul = (double*)malloc(sizeof(double) * nn);
    double* loc = (double*)malloc(sizeof(double) * nn);
    double* C = (double*)malloc(sizeof(double) * rank);
    for (int n = 0; n < nn; ++n) mul[pv[n]] = M[n];
    for (int r = 0; r < rank; ++r) C[r] = dot(nn, K + (size_t)rank * nn + (size_t)r * nn, mul);
    for (int m = 0; m < nn; ++m) {
        double s = 0;
        for (int r = 0; r < rank; ++r) s += K[(size_t)r * nn + m] * C[r];
        loc[m] = scale * s;
    }
    for (int n = 0; n < nn; ++n) L[n] += loc[pv[n]];
    free(mul); free(loc); free(C);
    return ORC_OK;
}

int orc_cheb_dense_block(int order, int i, int j, int k, double* out) {
    if (order < ORC_MIN_ORDER || order > ORC_MAX_ORDER) return ORC_ERR_ARG;
    orc_cheb tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.p = order; tmp.nn = order * order * order;
    orc_cheb_roots(order, tmp.roots);
    double* X = (double*)malloc(sizeof(double) * 3 * tmp.nn);
    double* Y = (double*)malloc(sizeof(double) * 3 * tmp.nn);
    const double zero[3] = {0, 0, 0};
    const double cy[3] = {2. * i, 2. * j, 2. * k};
    set_roots(&tmp, zero, 2., X);
    set_roots(&tmp, cy, 2., Y);
    for (int m = 0; m < tmp.nn; ++m)
        for (int n = 0; n < tmp.nn; ++n) out[(size_t)m * tmp.nn + n] = kernel_r(X + 3 * n, Y + 3 * m);
    free(X); free(Y);
    return ORC_OK;
}

/* ======================================================================== */
/* Per-leaf / per-cell operators                                            */
/* ======================================================================== */

/* S(x, xbar_j) = 1/p + 2/p sum_{o>=1} T_o(x) T_o(xbar_j), T_o(x) by the
 * three-term recurrence (FChebInterpolator.hpp:675-686).  dS = d/dx S using
 * T_o' = o U_{o-1} (FChebInterpolator.hpp:1306-1332). */
/* FUnifRoots::L and ::dL (Kernels/Uniform/FUnifRoots.hpp:49-118): L_n by
 * the scaled product, dL_n by the sum of products over the root table; the
 * argument clamped to [-1, 1]. */
static double unif_L(int p, int n, double x) {
    if (x > 1.) x = 1.;
    if (x < -1.) x = -1.;
    const int omn = p - n - 1;
    double fn = 1., fo = 1.;
    for (int i = 2; i <= n; ++i) fn *= i;
    for (int i = 2; i <= omn; ++i) fo *= i;
    double scale = (omn % 2) ? -1. : 1.;
    scale /= pow(2., (double)(p - 1)) * fn * fo;
    double L = 1.;
    for (int m = 0; m < p; ++m)
        if (m != n) L *= (double)(p - 1) * (x + 1.) - 2. * m;
    return L * scale;
}

static double unif_dL(const double* roots, int p, int n, double x) {
    if (x > 1.) x = 1.;
    if (x < -1.) x = -1.;
    double num = 0., den = 1.;
    for (int q = 0; q < p; ++q) {
        if (q == n) continue;
        double t = 1.;
        for (int m = 0; m < p; ++m)
            if (m != n && m != q) t *= x - roots[m];
        num += t;
        den *= roots[n] - roots[q];
    }
    return num / den;
}

static void interp_1d(const orc_cheb* c, double xi, double* S, double* dS) {
    const int p = c->p;
    if (c->basis == 1) {
        for (int j = 0; j < p; ++j) {
            S[j] = unif_L(p, j, xi);
            if (dS) dS[j] = unif_dL(c->roots, p, j, xi);
        }
        return;
    }
    double T[ORC_MAX_ORDER], U[ORC_MAX_ORDER];
    T[0] = 1.; T[1] = xi;
    U[0] = 1.; U[1] = 2. * xi;
    for (int j = 2; j < p; ++j) {
        T[j] = 2. * xi * T[j - 1] - T[j - 2];
        U[j] = 2. * xi * U[j - 1] - U[j - 2];
    }
    for (int j = 0; j < p; ++j) {
        double s = 1. / p, ds = 0.;
        for (int o = 1; o < p; ++o) {
            s += 2. / p * T[o] * c->Troots[o][j];
            ds += 2. / p * (o * U[o - 1]) * c->Troots[o][j];
        }
        S[j] = s;
        if (dS) dS[j] = ds;
    }
}

void orc_p2m(const orc_cheb* c, const double center[3], double width, int64_t n,
             const double* x, const double* y, const double* z, const double* q, double* M) {
    /* applyP2M (FChebInterpolator.hpp:625-762) with map_glob_loc (FInterpMapping.hpp:97-103) */
    const int p = c->p;
    double a[3], b[3];
    for (int d = 0; d < 3; ++d) { a[d] = center[d] - width / 2.; b[d] = center[d] + width / 2.; }
    for (int64_t i = 0; i < n; ++i) {
        const double pos[3] = {x[i], y[i], z[i]};
        double S[3][ORC_MAX_ORDER];
        for (int d = 0; d < 3; ++d) {
            const double xi = (2. * pos[d] - b[d] - a[d]) / (b[d] - a[d]);
            interp_1d(c, xi, S[d], NULL);
        }
        for (int kk = 0; kk < p; ++kk)
            for (int jj = 0; jj < p; ++jj) {
                const double qyz = q[i] * S[1][jj] * S[2][kk];
                for (int ii = 0; ii < p; ++ii) M[kk * p * p + jj * p + ii] += qyz * S[0][ii];
            }
    }
}

void orc_l2p(const orc_cheb* c, const double center[3], double width, const double* L, int64_t n,
             const double* x, const double* y, const double* z, const double* q,
             double* pot, double* fx, double* fy, double* fz) {
    /* applyL2PTotal (FChebInterpolator.hpp:1220-1417): potential and q*grad */
    const int p = c->p;
    double a[3], b[3], jac[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = center[d] - width / 2.; b[d] = center[d] + width / 2.;
        jac[d] = 2. / (b[d] - a[d]);
    }
    for (int64_t i = 0; i < n; ++i) {
        const double pos[3] = {x[i], y[i], z[i]};
        double S[3][ORC_MAX_ORDER], dS[3][ORC_MAX_ORDER];
        for (int d = 0; d < 3; ++d) {
            const double xi = (2. * pos[d] - b[d] - a[d]) / (b[d] - a[d]);
            interp_1d(c, xi, S[d], dS[d]);
        }
        double ph = 0, gx = 0, gy = 0, gz = 0;
        for (int kk = 0; kk < p; ++kk)
            for (int jj = 0; jj < p; ++jj) {
                double s0 = 0, s1 = 0;
                for (int ii = 0; ii < p; ++ii) {
                    const double l = L[kk * p * p + jj * p + ii];
                    s0 += l * S[0][ii];
                    s1 += l * dS[0][ii];
                }
                ph += s0 * S[1][jj] * S[2][kk];
                gx += s1 * S[1][jj] * S[2][kk];
                gy += s0 * dS[1][jj] * S[2][kk];
                gz += s0 * S[1][jj] * dS[2][kk];
            }
        pot[i] += ph;
        fx[i] += gx * jac[0] * q[i];
        fy[i] += gy * jac[1] * q[i];
        fz[i] += gz * jac[2] * q[i];
    }
}

void orc_m2m(const orc_cheb* c, int child, const double* Mc, double* Mp) {
    /* applyM2M (FChebInterpolator.hpp:560-584): parent[n] += sum_m
     * Sx[n_x][m_x] Sy[n_y][m_y] Sz[n_z][m_z] child[m] */
    const int p = c->p, nn = c->nn;
    const double* Sx = c->child + ((size_t)child * 3 + 0) * p * p;
    const double* Sy = c->child + ((size_t)child * 3 + 1) * p * p;
    const double* Sz = c->child + ((size_t)child * 3 + 2) * p * p;
    double* A = (double*)malloc(sizeof(double) * nn);
    double* B = (double*)malloc(sizeof(double) * nn);
    /* contract x: A[k][j][nx] = sum_mx Sx[nx][mx] child[k][j][mx] */
    for (int k = 0; k < p; ++k)
        for (int j = 0; j < p; ++j)
            for (int nx = 0; nx < p; ++nx) {
                double s = 0;
                for (int m = 0; m < p; ++m) s += Sx[nx * p + m] * Mc[k * p * p + j * p + m];
                A[k * p * p + j * p + nx] = s;
            }
    for (int k = 0; k < p; ++k)
        for (int ny = 0; ny < p; ++ny)
            for (int i = 0; i < p; ++i) {
                double s = 0;
                for (int m = 0; m < p; ++m) s += Sy[ny * p + m] * A[k * p * p + m * p + i];
                B[k * p * p + ny * p + i] = s;
            }
    for (int nz = 0; nz < p; ++nz)
        for (int j = 0; j < p; ++j)
            for (int i = 0; i < p; ++i) {
                double s = 0;
                for (int m = 0; m < p; ++m) s += Sz[nz * p + m] * B[m * p * p + j * p + i];
                Mp[nz * p * p + j * p + i] += s;
            }
    free(A); free(B);
}

void orc_l2l(const orc_cheb* c, int child, const double* Lp, double* Lc) {
    /* applyL2L (FChebInterpolator.hpp:587-611): child[m] += sum_n
     * Sx[n_x][m_x] Sy[n_y][m_y] Sz[n_z][m_z] parent[n] */
    const int p = c->p, nn = c->nn;
    const double* Sx = c->child + ((size_t)child * 3 + 0) * p * p;
    const double* Sy = c->child + ((size_t)child * 3 + 1) * p * p;
    const double* Sz = c->child + ((size_t)child * 3 + 2) * p * p;
    double* A = (double*)malloc(sizeof(double) * nn);
    double* B = (double*)malloc(sizeof(double) * nn);
    for (int k = 0; k < p; ++k)
        for (int j = 0; j < p; ++j)
            for (int mx = 0; mx < p; ++mx) {
                double s = 0;
                for (int n = 0; n < p; ++n) s += Sx[n * p + mx] * Lp[k * p * p + j * p + n];
                A[k * p * p + j * p + mx] = s;
            }
    for (int k = 0; k < p; ++k)
        for (int my = 0; my < p; ++my)
            for (int i = 0; i < p; ++i) {
                double s = 0;
                for (int n = 0; n < p; ++n) s += Sy[n * p + my] * A[k * p * p + n * p + i];
                B[k * p * p + my * p + i] = s;
            }
    for (int mz = 0; mz < p; ++mz)
        for (int j = 0; j < p; ++j)
            for (int i = 0; i < p; ++i) {
                double s = 0;
                for (int n = 0; n < p; ++n) s += Sz[n * p + mz] * B[n * p * p + j * p + i];
                Lc[mz * p * p + j * p + i] += s;
            }
    free(A); free(B);
}

/* ======================================================================== */
/* Tree (GroupTree/Core/FGroupTree.hpp:178-360)                             */
/* ======================================================================== */

struct orc_tree {
    int height;
    int64_t n;
    int64_t* skeys;  /* sorted particle keys */
    int64_t* perm;   /* sorted -> original index */
    int64_t ncell[21];
    int64_t* keys[21];    /* per level sorted cell keys */
    int64_t* lstart;      /* leaf particle start */
    int64_t* lcount;
    /* lists */
    int64_t* p2p_off; int64_t* p2p_nbr; int32_t* p2p_pos;
    int64_t* m2l_off[21]; int64_t* m2l_src[21]; int32_t* m2l_pos[21];
};

typedef struct { int64_t key, idx; } kv;
static int kv_cmp(const void* a, const void* b) {
    const kv* x = (const kv*)a; const kv* y = (const kv*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int64_t find_key(const int64_t* keys, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < n && keys[lo] == key) ? lo : -1;
}

int orc_tree_build(int64_t n, const double* x, const double* y, const double* z,
                   const double center[3], double width, int height, orc_tree** out) {
    if (n < 0 || height < 2 || height > 20 || !out) return ORC_ERR_ARG;
    int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    const int st = orc_particle_keys(n, x, y, z, center, width, height, NULL, keys);
    if (st) { free(keys); return st; }
    orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
    t->height = height;
    t->n = n;
    kv* arr = (kv*)malloc(sizeof(kv) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) { arr[i].key = keys[i]; arr[i].idx = i; }
    qsort(arr, (size_t)n, sizeof(kv), kv_cmp);
    t->skeys = keys;
    t->perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) { t->skeys[i] = arr[i].key; t->perm[i] = arr[i].idx; }
    free(arr);
    /* leaves: run-length encoding of the sorted keys (FGroupTree.hpp:224-289) */
    const int L = height - 1;
    int64_t nl = 0;
    for (int64_t i = 0; i < n; ++i) if (i == 0 || t->skeys[i] != t->skeys[i - 1]) ++nl;
    t->ncell[L] = nl;
    t->keys[L] = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nl > 0 ? nl : 1));
    t->lstart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nl > 0 ? nl : 1));
    t->lcount = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nl > 0 ? nl : 1));
    nl = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i == 0 || t->skeys[i] != t->skeys[i - 1]) {
            t->keys[L][nl] = t->skeys[i]; t->lstart[nl] = i; t->lcount[nl] = 0; ++nl;
        }
        t->lcount[nl - 1]++;
    }
    /* upper levels: key >> 3 deduplicated (FGroupTree.hpp:293-360) */
    for (int l = L - 1; l >= 0; --l) {
        const int64_t nc = t->ncell[l + 1];
        int64_t cnt = 0;
        t->keys[l] = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
        for (int64_t i = 0; i < nc; ++i) {
            const int64_t pk = t->keys[l + 1][i] >> 3;
            if (cnt == 0 || t->keys[l][cnt - 1] != pk) t->keys[l][cnt++] = pk;
        }
        t->ncell[l] = cnt;
    }
    /* P2P list of every leaf */
    t->p2p_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(t->ncell[L] + 1));
    t->p2p_off[0] = 0;
    for (int64_t i = 0; i < t->ncell[L]; ++i) {
        int32_t c3[3]; int64_t nk[26]; int32_t np[26];
        orc_morton_decode(t->keys[L][i], c3);
        const int m = orc_neighbor_keys(c3[0], c3[1], c3[2], height, nk, np);
        int64_t e = 0;
        for (int a = 0; a < m; ++a) if (find_key(t->keys[L], t->ncell[L], nk[a]) >= 0) ++e;
        t->p2p_off[i + 1] = t->p2p_off[i] + e;
    }
    t->p2p_nbr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(t->p2p_off[t->ncell[L]] + 1));
    t->p2p_pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(t->p2p_off[t->ncell[L]] + 1));
    for (int64_t i = 0; i < t->ncell[L]; ++i) {
        int32_t c3[3]; int64_t nk[26]; int32_t np[26];
        orc_morton_decode(t->keys[L][i], c3);
        const int m = orc_neighbor_keys(c3[0], c3[1], c3[2], height, nk, np);
        int64_t e = t->p2p_off[i];
        for (int a = 0; a < m; ++a) {
            const int64_t f = find_key(t->keys[L], t->ncell[L], nk[a]);
            if (f >= 0) { t->p2p_nbr[e] = f; t->p2p_pos[e] = np[a]; ++e; }
        }
    }
    /* M2L lists for levels 2..h-1 (FGroupTaskAlgorithm.hpp:417-510) */
    for (int l = 2; l <= L; ++l) {
        const int64_t nc = t->ncell[l];
        t->m2l_off[l] = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc + 1));
        t->m2l_off[l][0] = 0;
        int64_t* cnts = (int64_t*)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int64_t));
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            int32_t c3[3]; int64_t ik[216]; int32_t ip[216];
            orc_morton_decode(t->keys[l][i], c3);
            const int m = orc_interaction_keys(c3[0], c3[1], c3[2], l, ik, ip);
            int64_t e = 0;
            for (int a = 0; a < m; ++a) if (find_key(t->keys[l], nc, ik[a]) >= 0) ++e;
            cnts[i] = e;
        }
        for (int64_t i = 0; i < nc; ++i) t->m2l_off[l][i + 1] = t->m2l_off[l][i] + cnts[i];
        free(cnts);
        t->m2l_src[l] = (int64_t*)malloc(sizeof(int64_t) * (size_t)(t->m2l_off[l][nc] + 1));
        t->m2l_pos[l] = (int32_t*)malloc(sizeof(int32_t) * (size_t)(t->m2l_off[l][nc] + 1));
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i) {
            int32_t c3[3]; int64_t ik[216]; int32_t ip[216];
            orc_morton_decode(t->keys[l][i], c3);
            const int m = orc_interaction_keys(c3[0], c3[1], c3[2], l, ik, ip);
            int64_t e = t->m2l_off[l][i];
            for (int a = 0; a < m; ++a) {
                const int64_t f = find_key(t->keys[l], nc, ik[a]);
                if (f >= 0) { t->m2l_src[l][e] = f; t->m2l_pos[l][e] = ip[a]; ++e; }
            }
        }
    }
    *out = t;
    return ORC_OK;
}

void orc_tree_destroy(orc_tree* t) {
    if (!t) return;
    free(t->skeys); free(t->perm); free(t->lstart); free(t->lcount);
    free(t->p2p_off); free(t->p2p_nbr); free(t->p2p_pos);
    for (int l = 0; l < 21; ++l) { free(t->keys[l]); free(t->m2l_off[l]); free(t->m2l_src[l]); free(t->m2l_pos[l]); }
    free(t);
}

void orc_tree_sorted(const orc_tree* t, int64_t* keys, int64_t* perm) {
    if (keys) memcpy(keys, t->skeys, sizeof(int64_t) * (size_t)t->n);
    if (perm) memcpy(perm, t->perm, sizeof(int64_t) * (size_t)t->n);
}
int64_t orc_tree_num_cells(const orc_tree* t, int level) {
    return (level >= 0 && level < t->height) ? t->ncell[level] : 0;
}
void orc_tree_cell_keys(const orc_tree* t, int level, int64_t* keys) {
    if (level >= 0 && level < t->height) memcpy(keys, t->keys[level], sizeof(int64_t) * (size_t)t->ncell[level]);
}
void orc_tree_leaf_ranges(const orc_tree* t, int64_t* start, int64_t* count) {
    const int L = t->height - 1;
    if (start) memcpy(start, t->lstart, sizeof(int64_t) * (size_t)t->ncell[L]);
    if (count) memcpy(count, t->lcount, sizeof(int64_t) * (size_t)t->ncell[L]);
}
int64_t orc_tree_p2p_list_size(const orc_tree* t) { return t->p2p_off[t->ncell[t->height - 1]]; }
void orc_tree_p2p_list(const orc_tree* t, int64_t* off, int64_t* nbr, int32_t* pos) {
    const int64_t nl = t->ncell[t->height - 1], m = t->p2p_off[nl];
    if (off) memcpy(off, t->p2p_off, sizeof(int64_t) * (size_t)(nl + 1));
    if (nbr) memcpy(nbr, t->p2p_nbr, sizeof(int64_t) * (size_t)m);
    if (pos) memcpy(pos, t->p2p_pos, sizeof(int32_t) * (size_t)m);
}
int64_t orc_tree_m2l_list_size(const orc_tree* t, int level) {
    if (level < 2 || level >= t->height) return 0;
    return t->m2l_off[level][t->ncell[level]];
}
void orc_tree_m2l_list(const orc_tree* t, int level, int64_t* off, int64_t* src, int32_t* pos) {
    if (level < 2 || level >= t->height) return;
    const int64_t nc = t->ncell[level], m = t->m2l_off[level][nc];
    if (off) memcpy(off, t->m2l_off[level], sizeof(int64_t) * (size_t)(nc + 1));
    if (src) memcpy(src, t->m2l_src[level], sizeof(int64_t) * (size_t)m);
    if (pos) memcpy(pos, t->m2l_pos[level], sizeof(int32_t) * (size_t)m);
}

/* ======================================================================== */
/* FMM passes (GroupTree/Core/FGroupSeqAlgorithm.hpp:42-415)                */
/* ======================================================================== */

struct orc_run {
    const orc_cheb* c;
    orc_tree* t;
    double* M[21];
    double* L[21];
    /* sorted particle SoA */
    double *x, *y, *z, *q, *pot, *fx, *fy, *fz;
    orc_stats st;
    double center[3], width;
    /* FChebKernel: compressed multipoles / locals (k per cell) beside M and L */
    const orc_chebg* g;
    double* Mt[21];
    double* Lt[21];
};

/* y = B^T x (applyB, FChebM2LHandler.hpp:194-197: overwrite) */
static void chebg_apply_b(const orc_chebg* g, const double* x, double* y) {
    const int nn = g->c->nn;
    for (int a = 0; a < g->k; ++a) y[a] = dot(nn, g->B + (size_t)a * nn, x);
}

/* x += U X (applyU, FChebM2LHandler.hpp:150-153: gemva adds) */
static void chebg_apply_u(const orc_chebg* g, const double* X, double* x) {
    const int nn = g->c->nn;
    for (int n = 0; n < nn; ++n) {
        double s = 0;
        for (int a = 0; a < g->k; ++a) s += g->U[(size_t)a * nn + n] * X[a];
        x[n] += s;
    }
}

/* X += scale C_t Y (applyC, FChebM2LHandler.hpp:166-173) */
static void chebg_apply_c(const orc_chebg* g, int pos, double scale, const double* Y, double* X) {
    const int k = g->k;
    const double* Ct = g->C + (size_t)pos * k * k;
    for (int i = 0; i < k; ++i) {
        double s = 0;
        for (int j = 0; j < k; ++j) s += Ct[i + (size_t)j * k] * Y[j];
        X[i] += scale * s;
    }
}

/* P2P for one target leaf against the particles of another leaf, one-way.
 * FP2P::GenericFullMutual / GenericInner (Kernels/P2P/FP2P.hpp:103-269)
 * compute each unordered pair once and update both sides; the one-way form
 * gives every target the same set of contributions. */
static void p2p_oneway(const struct orc_run* r, int64_t t0, int64_t tn, int64_t s0, int64_t sn, int same) {
    for (int64_t i = t0; i < t0 + tn; ++i) {
        const double tx = r->x[i], ty = r->y[i], tz = r->z[i], tv = r->q[i];
        double fxs = 0, fys = 0, fzs = 0, pos = 0;
        for (int64_t j = s0; j < s0 + sn; ++j) {
            if (same && j == i) continue;
            const double dx = tx - r->x[j], dy = ty - r->y[j], dz = tz - r->z[j];
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double inv3 = inv * inv * inv;
            const double coef = tv * r->q[j];
            fxs += -inv3 * dx * coef;
            fys += -inv3 * dy * coef;
            fzs += -inv3 * dz * coef;
            pos += inv * r->q[j];
        }
        r->fx[i] += fxs; r->fy[i] += fys; r->fz[i] += fzs; r->pot[i] += pos;
    }
}

/* Working levels [upper, lower) as FAbstractAlgorithm::execute(ops, upper, lower)
 * sets them (Core/FCoreCommon.hpp:81-86), with the loop bounds of
 * FGroupSeqAlgorithm (GroupTree/Core/FGroupSeqAlgorithm.hpp:88, 137, 238): M2M
 * parents from min(h-2, lower-1) down to upper, M2L levels lower-1 down to
 * upper, L2L parents upper .. lower-2; P2M, L2P and P2P work on the leaves. */
static void run_passes(struct orc_run* r, unsigned ops, const double center[3], double width, int upper,
                       int lower) {
    const orc_cheb* c = r->c;
    const orc_tree* t = r->t;
    const int h = t->height, L = h - 1, nn = c->nn;
    const double corner[3] = {center[0] - width / 2., center[1] - width / 2., center[2] - width / 2.};
    const double leafw = width / pow(2., (double)(h - 1));
    double t0;
    /* bottom pass: P2M (FGroupSeqAlgorithm bottomPass -> FChebSymKernel::P2M :181-188) */
    t0 = wall_seconds();
    if (ops & ORC_P2M) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < t->ncell[L]; ++i) {
            int32_t c3[3];
            orc_morton_decode(t->keys[L][i], c3);
            const double ctr[3] = {corner[0] + ((double)c3[0] + .5) * leafw, corner[1] + ((double)c3[1] + .5) * leafw,
                                   corner[2] + ((double)c3[2] + .5) * leafw};
            const int64_t s = t->lstart[i];
            orc_p2m(c, ctr, leafw, t->lcount[i], r->x + s, r->y + s, r->z + s, r->q + s, r->M[L] + (size_t)i * nn);
            if (r->g) chebg_apply_b(r->g, r->M[L] + (size_t)i * nn, r->Mt[L] + (size_t)i * r->g->k);
        }
    }
    r->st.t_p2m += wall_seconds() - t0;
    /* upward pass: M2M levels h-2..2 */
    t0 = wall_seconds();
    if (ops & ORC_M2M) {
        for (int l = (L - 1 < lower - 1 ? L - 1 : lower - 1); l >= upper; --l) {
            const int64_t nc = t->ncell[l], nch = t->ncell[l + 1];
            const int64_t* ck = t->keys[l + 1];
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t i = 0; i < nc; ++i) {
                const int64_t pk = t->keys[l][i];
                int64_t lo = 0, hi = nch;
                while (lo < hi) { const int64_t mid = (lo + hi) / 2; if ((ck[mid] >> 3) < pk) lo = mid + 1; else hi = mid; }
                for (int64_t j = lo; j < nch && (ck[j] >> 3) == pk; ++j)
                    orc_m2m(c, (int)(ck[j] & 7), r->M[l + 1] + (size_t)j * nn, r->M[l] + (size_t)i * nn);
                if (r->g) chebg_apply_b(r->g, r->M[l] + (size_t)i * nn, r->Mt[l] + (size_t)i * r->g->k);
            }
        }
    }
    r->st.t_m2m += wall_seconds() - t0;
    /* transfer pass: M2L levels 2..h-1 */
    t0 = wall_seconds();
    if (ops & ORC_M2L) {
        for (int l = lower - 1; l >= upper; --l) {
            /* getScaleFactor(RootCellWidth, level) = 2/(W/2^l) (FInterpMatrixKernel.hpp:124-133) */
            const double scale = 2. / (width / pow(2., (double)l));
            const int64_t nc = t->ncell[l];
            r->st.n_m2l += t->m2l_off[l][nc];
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t i = 0; i < nc; ++i)
                for (int64_t e = t->m2l_off[l][i]; e < t->m2l_off[l][i + 1]; ++e)
                    if (r->g)
                        chebg_apply_c(r->g, t->m2l_pos[l][e], scale, r->Mt[l] + (size_t)t->m2l_src[l][e] * r->g->k,
                                      r->Lt[l] + (size_t)i * r->g->k);
                    else
                        orc_cheb_apply_m2l(c, t->m2l_pos[l][e], scale, r->M[l] + (size_t)t->m2l_src[l][e] * nn,
                                           r->L[l] + (size_t)i * nn);
        }
    }
    r->st.t_m2l += wall_seconds() - t0;
    /* downward pass: L2L levels 2..h-2 */
    t0 = wall_seconds();
    if (ops & ORC_L2L) {
        for (int l = upper; l < lower - 1; ++l) {
            const int64_t nc = t->ncell[l], nch = t->ncell[l + 1];
            const int64_t* ck = t->keys[l + 1];
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t i = 0; i < nc; ++i) {
                const int64_t pk = t->keys[l][i];
                int64_t lo = 0, hi = nch;
                while (lo < hi) { const int64_t mid = (lo + hi) / 2; if ((ck[mid] >> 3) < pk) lo = mid + 1; else hi = mid; }
                if (r->g) chebg_apply_u(r->g, r->Lt[l] + (size_t)i * r->g->k, r->L[l] + (size_t)i * nn);
                for (int64_t j = lo; j < nch && (ck[j] >> 3) == pk; ++j)
                    orc_l2l(c, (int)(ck[j] & 7), r->L[l] + (size_t)i * nn, r->L[l + 1] + (size_t)j * nn);
            }
        }
    }
    r->st.t_l2l += wall_seconds() - t0;
    /* near field: P2P over the 27-leaf neighbourhood */
    t0 = wall_seconds();
    if (ops & ORC_P2P) {
        int64_t pairs = 0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : pairs)
        for (int64_t i = 0; i < t->ncell[L]; ++i) {
            const int64_t s = t->lstart[i], cnt = t->lcount[i];
            p2p_oneway(r, s, cnt, s, cnt, 1);
            pairs += cnt * (cnt - 1);
            for (int64_t e = t->p2p_off[i]; e < t->p2p_off[i + 1]; ++e) {
                const int64_t j = t->p2p_nbr[e];
                p2p_oneway(r, s, cnt, t->lstart[j], t->lcount[j], 0);
                pairs += cnt * t->lcount[j];
            }
        }
        r->st.n_p2p_pairs += pairs;
    }
    r->st.t_p2p += wall_seconds() - t0;
    /* merge pass: L2P (FChebSymKernel::L2P :432-451) */
    t0 = wall_seconds();
    if (ops & ORC_L2P) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < t->ncell[L]; ++i) {
            int32_t c3[3];
            orc_morton_decode(t->keys[L][i], c3);
            const double ctr[3] = {corner[0] + ((double)c3[0] + .5) * leafw, corner[1] + ((double)c3[1] + .5) * leafw,
                                   corner[2] + ((double)c3[2] + .5) * leafw};
            const int64_t s = t->lstart[i];
            if (r->g) chebg_apply_u(r->g, r->Lt[L] + (size_t)i * r->g->k, r->L[L] + (size_t)i * nn);
            orc_l2p(c, ctr, leafw, r->L[L] + (size_t)i * nn, t->lcount[i], r->x + s, r->y + s, r->z + s, r->q + s,
                    r->pot + s, r->fx + s, r->fy + s, r->fz + s);
        }
    }
    r->st.t_l2p += wall_seconds() - t0;
}

static int run_setup(const orc_cheb* c, int64_t n, const double* x, const double* y, const double* z,
                     const double* q, const double center[3], double width, int height, struct orc_run** out) {
    double t0 = wall_seconds();
    struct orc_run* r = (struct orc_run*)calloc(1, sizeof(struct orc_run));
    r->c = c;
    for (int d = 0; d < 3; ++d) r->center[d] = center[d];
    r->width = width;
    int st = orc_tree_build(n, x, y, z, center, width, height, &r->t);
    if (st) { free(r); return st; }
    const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
    r->x = (double*)malloc(nb); r->y = (double*)malloc(nb); r->z = (double*)malloc(nb); r->q = (double*)malloc(nb);
    r->pot = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    r->fx = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    r->fy = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    r->fz = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = r->t->perm[i];
        r->x[i] = x[o]; r->y[i] = y[o]; r->z[i] = z[o]; r->q[i] = q[o];
    }
    for (int l = 2; l < height; ++l) {
        r->M[l] = (double*)calloc((size_t)(r->t->ncell[l] > 0 ? r->t->ncell[l] : 1) * c->nn, sizeof(double));
        r->L[l] = (double*)calloc((size_t)(r->t->ncell[l] > 0 ? r->t->ncell[l] : 1) * c->nn, sizeof(double));
    }
    r->st.t_tree = wall_seconds() - t0;
    *out = r;
    return ORC_OK;
}

void orc_run_destroy(orc_run* r) {
    if (!r) return;
    orc_tree_destroy(r->t);
    free(r->x); free(r->y); free(r->z); free(r->q); free(r->pot); free(r->fx); free(r->fy); free(r->fz);
    for (int l = 0; l < 21; ++l) { free(r->M[l]); free(r->L[l]); free(r->Mt[l]); free(r->Lt[l]); }
    free(r);
}

int orc_run_create(const orc_cheb* c, int64_t n, const double* x, const double* y, const double* z,
                   const double* q, const double center[3], double width, int height, unsigned ops, orc_run** out) {
    if (!c || !out) return ORC_ERR_ARG;
    int st = run_setup(c, n, x, y, z, q, center, width, height, out);
    if (st) return st;
    run_passes(*out, ops, center, width, 2, height);
    return ORC_OK;
}

const orc_tree* orc_run_tree(const orc_run* r) { return r->t; }

/* Another execute(ops, upper, lower) on the run's tree: the passes add into
 * the cells and the particle results left by earlier executions, as the
 * reference's tree keeps them (every kernel operator accumulates:
 * FChebInterpolator.hpp:627-700, 560-582; FChebSymKernel.hpp:193-206,
 * 210-328, 331-350, 432-486).  pot..fz receive the accumulated particle
 * results in input order (NULL: not copied). */
int orc_run_execute(orc_run* r, unsigned ops, int upper, int lower, double* pot, double* fx, double* fy,
                    double* fz) {
    if (!r) return ORC_ERR_ARG;
    const int h = r->t->height;
    if (upper < 2 || upper > lower || lower > h) return ORC_ERR_ARG;
    run_passes(r, ops, r->center, r->width, upper, lower);
    for (int64_t i = 0; i < r->t->n; ++i) {
        const int64_t o = r->t->perm[i];
        if (pot) pot[o] = r->pot[i];
        if (fx) fx[o] = r->fx[i];
        if (fy) fy[o] = r->fy[i];
        if (fz) fz[o] = r->fz[i];
    }
    return ORC_OK;
}

/* tree.forEachCell(resetToInitialState) (FChebCell.hpp:151-152): zero
 * multipoles and locals; the particles keep their results. */
void orc_run_reset_cells(orc_run* r) {
    if (!r) return;
    for (int l = 2; l < r->t->height; ++l) {
        const size_t sz = sizeof(double) * (size_t)(r->t->ncell[l] > 0 ? r->t->ncell[l] : 1) * r->c->nn;
        memset(r->M[l], 0, sz);
        memset(r->L[l], 0, sz);
    }
}

void orc_run_expansions(const orc_run* r, int level, double* M, double* L) {
    if (level < 2 || level >= r->t->height) return;
    const size_t sz = sizeof(double) * (size_t)r->t->ncell[level] * r->c->nn;
    if (M) memcpy(M, r->M[level], sz);
    if (L) memcpy(L, r->L[level], sz);
}

int orc_fmm_with(const orc_cheb* c, int64_t n, const double* x, const double* y, const double* z,
                 const double* q, const double center[3], double width, int height,
                 unsigned ops, int nthreads,
                 double* pot, double* fx, double* fy, double* fz, orc_stats* stats) {
    return orc_fmm_levels(c, n, x, y, z, q, center, width, height, ops, 2, height, nthreads, pot, fx, fy, fz,
                          stats);
}

int orc_fmm_levels(const orc_cheb* c, int64_t n, const double* x, const double* y, const double* z,
                   const double* q, const double center[3], double width, int height, unsigned ops, int upper,
                   int lower, int nthreads, double* pot, double* fx, double* fy, double* fz, orc_stats* stats) {
    if (!c || n < 0) return ORC_ERR_ARG;
    if (upper < 2 || upper > lower || lower > height) return ORC_ERR_ARG;  /* validateLevels, FCoreCommon.hpp:49-52 */
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double t0 = wall_seconds();
    struct orc_run* r = NULL;
    int st = run_setup(c, n, x, y, z, q, center, width, height, &r);
    if (st == ORC_OK) {
        run_passes(r, ops, center, width, upper, lower);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t o = r->t->perm[i];
            pot[o] = r->pot[i]; fx[o] = r->fx[i]; fy[o] = r->fy[i]; fz[o] = r->fz[i];
        }
        r->st.t_total = wall_seconds() - t0;
        if (stats) *stats = r->st;
        orc_run_destroy(r);
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return st;
}

int orc_fmm_global(const orc_chebg* g, int64_t n, const double* x, const double* y, const double* z,
                   const double* q, const double center[3], double width, int height, unsigned ops, int upper,
                   int lower, int nthreads, double* pot, double* fx, double* fy, double* fz, orc_stats* stats) {
    if (!g || n < 0) return ORC_ERR_ARG;
    if (upper < 2 || upper > lower || lower > height) return ORC_ERR_ARG;
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double t0 = wall_seconds();
    struct orc_run* r = NULL;
    int st = run_setup(g->c, n, x, y, z, q, center, width, height, &r);
    if (st == ORC_OK) {
        r->g = g;
        for (int l = 2; l < height; ++l) {
            const size_t cells = (size_t)(r->t->ncell[l] > 0 ? r->t->ncell[l] : 1);
            r->Mt[l] = (double*)calloc(cells * g->k, sizeof(double));
            r->Lt[l] = (double*)calloc(cells * g->k, sizeof(double));
        }
        run_passes(r, ops, center, width, upper, lower);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t o = r->t->perm[i];
            pot[o] = r->pot[i]; fx[o] = r->fx[i]; fy[o] = r->fy[i]; fz[o] = r->fz[i];
        }
        r->st.t_total = wall_seconds() - t0;
        if (stats) *stats = r->st;
        orc_run_destroy(r);
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return st;
}

int orc_fmm(int64_t n, const double* x, const double* y, const double* z, const double* q,
            const double center[3], double width, int height, int order, double epsilon,
            unsigned ops, int nthreads,
            double* pot, double* fx, double* fy, double* fz, orc_stats* stats) {
    orc_cheb* c = NULL;
    int st = orc_cheb_create(order, epsilon, &c);
    if (st) return st;
    st = orc_fmm_with(c, n, x, y, z, q, center, width, height, ops, nthreads, pot, fx, fy, fz, stats);
    orc_cheb_destroy(c);
    return st;
}

void orc_direct(int64_t n, const double* x, const double* y, const double* z, const double* q,
                int64_t nt, const int64_t* targets, int nthreads,
                double* pot, double* fx, double* fy, double* fz) {
    /* FP2PR::MutualParticles conventions (Kernels/P2P/FP2PR.hpp:14-42) */
    const int64_t m = targets ? nt : n;
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t a = 0; a < m; ++a) {
        const int64_t i = targets ? targets[a] : a;
        double ph = 0, gx = 0, gy = 0, gz = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            const double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double inv3 = inv * inv * inv;
            ph += q[j] * inv;
            gx += -inv3 * dx * q[j];
            gy += -inv3 * dy * q[j];
            gz += -inv3 * dz * q[j];
        }
        pot[a] = ph; fx[a] = gx * q[i]; fy[a] = gy * q[i]; fz[a] = gz * q[i];
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
}
