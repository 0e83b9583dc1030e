"""The oracle's restatement of FChebKernel's M2L (one global SVD of the 316
far-field operators, Kernels/Chebyshev/FChebM2LHandler.hpp:216-523, used by
Kernels/Chebyshev/FChebKernel.hpp:93-200), pinned without the reference:

- the weighted operators are reproduced by U C_t B^T within the truncation
  the rank rule allows (FSvd::getRank: Frobenius tail below eps, per side);
- weighted U and B have orthonormal columns;
- K_col and K_row^T hold the same rows up to order (the node set and the 316
  offsets are symmetric under reflection), so their singular values agree;
- the rank is getRank of those singular values;
- the FMM against the fixture's direct sums, against the symmetric kernel
  (P2P identical; far fields within the two compressions' errors) and the
  working levels (2, h) bit for bit.
"""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def g5(oracle):
    return oracle.ChebGlobal(5)


def _nodes(oracle, p):
    r = np.asarray(oracle.roots(p))
    ids = np.array([[n % p, (n // p) % p, n // (p * p)] for n in range(p ** 3)])
    return r[ids]


def _weights(oracle, p):
    r = np.asarray(oracle.roots(p))
    w1 = np.pi / p * np.sqrt(1 - r * r)
    return np.array([np.sqrt(w1[n % p] * w1[(n // p) % p] * w1[n // (p * p)]) for n in range(p ** 3)])


def test_global_operators_reconstruct(oracle, g5):
    p, eps = 5, 1e-5
    U, B, Cm, S1, S2 = g5.operators()
    X = _nodes(oracle, p)
    w = _weights(oracle, p)
    err2 = tot2 = 0.0
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                pos = (i + 3) * 49 + (j + 3) * 7 + (k + 3)
                if max(abs(i), abs(j), abs(k)) <= 1:
                    assert not Cm[pos].any()
                    continue
                Y = X + 2.0 * np.array([i, j, k])
                A = 1.0 / np.linalg.norm(X[:, None, :] - Y[None, :, :], axis=2)  # A[m][n]: target m, source n
                Aw = w[:, None] * A * w[None, :]
                Rw = w[:, None] * (U @ Cm[pos] @ B.T) * w[None, :]
                err2 += np.sum((Rw - Aw) ** 2)
                tot2 += np.sum(Aw ** 2)
    rel = np.sqrt(err2 / tot2)
    assert rel < 2 * eps, rel  # two projections, each with a tail below eps


def test_global_operators_orthonormal_and_ranked(oracle, g5):
    p, eps = 5, 1e-5
    U, B, Cm, S1, S2 = g5.operators()
    w = _weights(oracle, p)
    k = g5.rank
    for M in (U, B):
        G = (w[:, None] * M).T @ (w[:, None] * M)
        assert np.abs(G - np.eye(k)).max() < 1e-12
    assert np.allclose(S1, S2, rtol=1e-10, atol=1e-12 * S1[0])
    assert np.all(np.diff(S1) <= 0)
    # FSvd::getRank: the smallest k whose tail (k .. nn-1) stays within eps^2 of the total
    tail = np.cumsum((S1 ** 2)[::-1])[::-1]
    expect = int(np.max(np.nonzero(tail > eps * eps * np.sum(S1 ** 2))[0])) + 1
    assert k == expect


def test_global_fmm_accuracy_and_symmetric_kernel(oracle, g5, fixture20k):
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    got = oracle.fmm_global(g5, x, y, z, q, c, w, 5)
    assert oracle.accuracy(d[:, 4], got[0])["rel_l2"] < 2e-6  # measured 1.0e-6
    for k in range(3):
        assert oracle.accuracy(d[:, 5 + k], got[1 + k])["rel_l2"] < 2e-4  # measured 8.6e-5
    near_g = oracle.fmm_global(g5, x, y, z, q, c, w, 5, ops=oracle.P2P)
    near_s = oracle.fmm(x, y, z, q, c, w, 5, 5, ops=oracle.P2P)
    assert all(np.array_equal(a, b) for a, b in zip(near_g[:4], near_s[:4]))
    far = oracle.ALL & ~oracle.P2P
    far_g = oracle.fmm_global(g5, x, y, z, q, c, w, 5, ops=far)
    far_s = oracle.fmm(x, y, z, q, c, w, 5, 5, ops=far)
    assert oracle.accuracy(far_s[0], far_g[0])["rel_l2"] < 5e-3
    same = oracle.fmm_global(g5, x, y, z, q, c, w, 5, levels=(2, 5))
    assert all(np.array_equal(a, b) for a, b in zip(same[:4], got[:4]))


def test_dense_operators_are_the_kernel(oracle):
    """FChebDenseKernel's operators (FChebDenseM2LHandler.hpp:180-215): C_t is
    1/r between the target roots and the source roots at 2 (i, j, k), column
    n for source node n; identity bases; the FMM then carries only the
    interpolation error, and the global SVD's compression adds at most its
    eps on top."""
    p = 4
    g = oracle.ChebGlobal(p, dense=True)
    U, B, Cm, _, _ = g.operators()
    nn = p ** 3
    assert g.rank == nn
    assert np.array_equal(U, np.eye(nn)) and np.array_equal(B, np.eye(nn))
    X = _nodes(oracle, p)
    for (i, j, k) in [(2, 0, 0), (-3, 1, 2), (0, -2, 3), (3, 3, 3)]:
        Y = X + 2.0 * np.array([i, j, k])
        A = 1.0 / np.linalg.norm(X[:, None, :] - Y[None, :, :], axis=2)
        assert np.allclose(Cm[(i + 3) * 49 + (j + 3) * 7 + (k + 3)], A, rtol=1e-14, atol=0)
    assert not Cm[(1 + 3) * 49 + (1 + 3) * 7 + (0 + 3)].any()


def test_dense_fmm_against_global_and_direct(oracle, g5, fixture20k):
    d = fixture20k["data"]
    c, w = fixture20k["center"], fixture20k["width"]
    x, y, z, q = d[:, 0], d[:, 1], d[:, 2], d[:, 3]
    dense = oracle.fmm_global(oracle.ChebGlobal(5, dense=True), x, y, z, q, c, w, 5)
    glob = oracle.fmm_global(g5, x, y, z, q, c, w, 5)
    assert oracle.accuracy(d[:, 4], dense[0])["rel_l2"] < 2e-6  # measured 1.0027e-6
    far = oracle.ALL & ~oracle.P2P
    fd = oracle.fmm_global(oracle.ChebGlobal(5, dense=True), x, y, z, q, c, w, 5, ops=far)
    fg = oracle.fmm_global(g5, x, y, z, q, c, w, 5, ops=far)
    assert oracle.accuracy(fd[0], fg[0])["rel_l2"] < 2 * 1e-5  # the compression's eps (order 5)
    assert oracle.accuracy(dense[0], glob[0])["rel_l2"] < 5e-7  # measured 1.7e-7


def test_uniform_kernel_reproduces_published_errors(oracle, fixture20k):
    """FUnifKernel of order 7 on the 20k fixture, height 5: the errors against
    the fixture's direct sums equal the reference's published ones
    (Data/ResultCompareKernels.txt, "Lagrange FMM ... ORDER 7") to the six
    printed digits."""
    d = fixture20k["data"]
    g = oracle.ChebGlobal(7, uniform=True)
    got = oracle.fmm_global(g, d[:, 0], d[:, 1], d[:, 2], d[:, 3], fixture20k["center"], fixture20k["width"], 5)
    published = [2.11079e-08, 2.23423e-06, 2.36648e-06, 2.36806e-06]
    for k in range(4):
        err = oracle.accuracy(d[:, 4 + k], got[k])["rel_l2"]
        assert abs(err - published[k]) <= 5e-6 * published[k], (k, err, published[k])


# Data/ResultCompareKernels.txt:18-23 (FChebKernel, order 7, height 5, the 20k
# fixture): potential and force errors against the direct sums
PUBLISHED_CHEB = {"pot": 1.12952e-08, "fx": 1.66854e-06, "fy": 1.75929e-06, "fz": 1.80712e-06}


def test_global_kernel_published_digits_within_the_degenerate_choice(oracle, fixture20k):
    """Why FChebKernel's published errors are not reproduced digit for digit.

    At order 7 the singular values 150, 151 and 152 of K_col and K_row are
    one triply degenerate value (the cube's axis permutations), well apart
    from 149 and 153, and the rank rule keeps 150 of them: the rank the
    reference printed, its M2L operator size 343 k^2 8 = 61,740,000 bytes
    (FChebM2LHandler.hpp:96-103).  Which vector of the degenerate plane is kept
    is the SVD routine's choice (LAPACK's dgesvd on the reference's machine).
    Operators from other choices inside the triplet (independently for U and
    B) span errors that contain every published value, so the published
    digits are one of the reference-conforming results, and the oracle's
    (Jacobi) choice another."""
    p, nn, eps = 7, 343, 1e-7
    X = _nodes(oracle, p)
    w = _weights(oracle, p)
    Ks, pos = [], []
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                if max(abs(i), abs(j), abs(k)) <= 1:
                    continue
                Y = X + 2.0 * np.array([i, j, k])
                Ks.append(w[:, None] * w[None, :] / np.linalg.norm(X[:, None, :] - Y[None, :, :], axis=2))
                pos.append((i + 3) * 49 + (j + 3) * 7 + (k + 3))
    _, S1, V1t = np.linalg.svd(np.concatenate(Ks, 0), full_matrices=False)  # K_col -> B
    U2, S2, _ = np.linalg.svd(np.concatenate(Ks, 1), full_matrices=False)   # K_row -> U
    for S in (S1, S2):
        tail = np.sqrt(np.cumsum((S ** 2)[::-1])[::-1] / np.sum(S ** 2))
        k = int(np.argmax(tail <= eps))  # FSvd::getRank: smallest k with the tail below eps
        assert k == 150 and 343 * k * k * 8 == 61_740_000
        assert abs(S[149] / S[150] - 1) < 1e-8 and abs(S[150] / S[151] - 1) < 1e-8  # sigma 150..152
        assert S[148] / S[149] > 1.02 and S[151] / S[152] > 1.1                      # isolated
    V1 = V1t.T
    P = [U2[:, :152].T @ K @ V1[:, :152] for K in Ks]  # weighted operators in the leading bases
    g = oracle.ChebGlobal(p, dense=True)  # an order-7 handle whose operators are replaced below
    d = fixture20k["data"]
    x, y, z, q = (d[:, c].copy() for c in range(4))

    def errors(a, b):
        Ua = np.concatenate([U2[:, :149], U2[:, 149:152] @ a[:, None]], 1)
        Bb = np.concatenate([V1[:, :149], V1[:, 149:152] @ b[:, None]], 1)
        Ea = np.zeros((152, 150))
        Ea[:149, :149] = np.eye(149)
        Eb = Ea.copy()
        Ea[149:152, 149] = a
        Eb[149:152, 149] = b
        C = np.zeros((343, 150, 150))
        for t, Pt in zip(pos, P):
            C[t] = Ea.T @ Pt @ Eb
        g.set_operators(Ua / w[:, None], Bb / w[:, None], C)
        got = oracle.fmm_global(g, x, y, z, q, fixture20k["center"], fixture20k["width"], 5)
        return [float(np.linalg.norm(got[c] - d[:, 4 + c]) / np.linalg.norm(d[:, 4 + c])) for c in range(4)]

    rng = np.random.default_rng(1)
    runs = [errors(np.eye(3)[c], np.eye(3)[c]) for c in range(3)]
    for _ in range(12):
        a, b = rng.normal(size=3), rng.normal(size=3)
        runs.append(errors(a / np.linalg.norm(a), b / np.linalg.norm(b)))
    runs = np.array(runs)
    for c, name in enumerate(("pot", "fx", "fy", "fz")):
        lo, hi = runs[:, c].min(), runs[:, c].max()
        assert lo <= PUBLISHED_CHEB[name] <= hi, (name, lo, PUBLISHED_CHEB[name], hi)
        assert hi - lo > 1e-3 * PUBLISHED_CHEB[name]  # the choice moves the fourth digit
