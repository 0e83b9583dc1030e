"""Oracle Chebyshev machinery: roots, interpolators, symmetries, QR/SVD,
compressed M2L operators, P2M/M2M/L2L/L2P.

Pins:
  - roots: the exact cosines (the reference table Src/Kernels/Chebyshev/FChebRoots.hpp:90-297
    is accurate to ~5e-16, so a wrong digit fails);
  - interpolators: partition of unity and exact reproduction of polynomials of
    degree < order;
  - symmetries: every far-field dense block equals the permuted canonical block
    (rule of Tests/Utils/testChebSymmetries.cpp:112-150, threshold 1e-15);
  - QR / SVD: numpy.linalg;
  - compressed operators: low-rank error against the dense weighted block;
  - P2M/M2M and L2L/L2P: the exact identities M2M(P2M(child)) = P2M(parent) and
    L2P(child, L2L(L)) = L2P(parent, L); gradient against central differences.
"""
import numpy as np
import pytest

ORDERS = [4, 5, 6, 7, 8]


@pytest.mark.parametrize("order", list(range(2, 15)))
def test_roots_table(oracle, order):
    r = oracle.roots(order)
    exact = np.sort(np.cos((2 * np.arange(order) + 1) * np.pi / (2 * order)))
    assert np.max(np.abs(r - exact)) < 1e-14
    assert np.all(np.diff(r) > 0)


def test_chebyshev_T(oracle):
    xs = np.linspace(-1, 1, 17)
    for n in range(0, 10):
        ref = np.polynomial.chebyshev.chebval(xs, [0] * n + [1])
        got = np.array([oracle.lib().orc_cheb_T(n, float(x)) for x in xs])
        assert np.max(np.abs(got - ref)) < 1e-13


def _interp_eval(order, x, roots):
    """Direct Lagrange basis at the Chebyshev nodes (independent of the oracle)."""
    out = np.ones(order)
    for j in range(order):
        for m in range(order):
            if m != j:
                out[j] *= (x - roots[m]) / (roots[j] - roots[m])
    return out


@pytest.mark.parametrize("order", ORDERS)
def test_child_interpolators(oracle, order):
    c = oracle.Cheb(order)
    S = c.child_interp()  # [child][axis][n][m] = S(x_child_m, xbar_n)
    r = oracle.roots(order)
    rel = np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1],
                    [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]])
    for ch in range(8):
        for d in range(3):
            xm = rel[ch, d] * 0.5 + 0.5 * r  # child nodes in the parent frame
            # Chebyshev interpolation weights equal the Lagrange basis at the nodes
            lag = np.array([_interp_eval(order, x, r) for x in xm])  # [m][n]
            assert np.max(np.abs(S[ch, d].T - lag)) < 1e-12
            assert np.max(np.abs(S[ch, d].sum(axis=0) - 1)) < 1e-13


@pytest.mark.parametrize("order", [4, 5, 6])
def test_symmetries(oracle, order):
    c = oracle.Cheb(order)
    pidx, pvec = c.symmetries()
    canon = {}
    for i in range(2, 4):
        for j in range(0, i + 1):
            for k in range(0, j + 1):
                canon[(i + 3) * 49 + (j + 3) * 7 + (k + 3)] = oracle.dense_block(order, i, j, k).T
    count = 0
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                if max(abs(i), abs(j), abs(k)) <= 1:
                    continue
                idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3)
                A = oracle.dense_block(order, i, j, k).T  # [target node][source node]
                P = pvec[idx].astype(np.int64)
                assert int(pidx[idx]) in canon
                B = np.empty_like(A)
                B[np.ix_(P, P)] = A
                assert np.max(np.abs(B - canon[int(pidx[idx])])) <= 1e-15
                count += 1
    assert count == 316


def test_qr_svd(oracle):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((60, 23))
    Q, R = oracle.qr(A)
    assert np.allclose(Q.T @ Q, np.eye(23), atol=1e-14)
    assert np.allclose(Q @ R, A, atol=1e-13)
    assert np.allclose(np.tril(R, -1), 0)
    B = rng.standard_normal((40, 40)) @ np.diag(10.0 ** -np.linspace(0, 12, 40)) @ rng.standard_normal((40, 40))
    U, S, VT = oracle.svd(B)
    S_np = np.linalg.svd(B, compute_uv=False)
    assert np.allclose(S, S_np, rtol=1e-10, atol=1e-15 * S_np[0])
    assert np.allclose((U * S) @ VT, B, atol=1e-13 * S_np[0])
    assert np.allclose(U[:, :20].T @ U[:, :20], np.eye(20), atol=1e-12)


def test_svd_rank_rule(oracle):
    S = 10.0 ** -np.arange(10.0)
    # keep the smallest k with sum_{i>=k} S_i^2 <= eps^2 sum S_i^2
    for eps, k in [(0.5, 1), (5e-4, 4), (5e-6, 6), (1e-20, 10)]:
        tail = np.array([np.sum(S[j:] ** 2) for j in range(11)])
        expect = int(np.argmax(tail <= eps ** 2 * np.sum(S ** 2)))
        assert oracle.svd_rank(S, eps) == expect == k


def _weights(order, r):
    w1 = np.pi / order * np.sqrt(1 - r ** 2)
    n = np.arange(order ** 3)
    return np.sqrt(w1[n % order] * w1[(n // order) % order] * w1[n // order ** 2])


@pytest.mark.parametrize("order", ORDERS)
def test_compressed_operators(oracle, order):
    c = oracle.Cheb(order)
    eps = 10.0 ** -order
    w = _weights(order, oracle.roots(order))
    for (i, j, k), rank in c.canonical_ranks().items():
        idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3)
        U, V = c.K(idx)
        assert U.shape[1] == rank and 0 < rank < order ** 3
        A = oracle.dense_block(order, i, j, k).T  # [target][source]
        Aw = w[:, None] * A * w[None, :]
        Ew = w[:, None] * (A - U @ V.T) * w[None, :]
        # The partially pivoted ACA stops when the last rank-1 update is small
        # (Utils/FAca.hpp:213-225); on the most symmetric far positions at
        # orders 4-5 that happens below the true eps-rank, which leaves errors
        # of a few 1e-3 relative.  Orders 6-8 stay within a few eps.
        bound = 10 * eps if order >= 6 else 1000 * eps
        assert np.linalg.norm(Ew) <= bound * np.linalg.norm(Aw)


def test_m2l_application(oracle):
    """apply_m2l of a far position = kernel action at cell width 2 (to ~eps)."""
    order = 6
    c = oracle.Cheb(order)
    rng = np.random.default_rng(3)
    for (i, j, k) in [(2, 0, 0), (-2, 3, 1), (0, -3, -3), (3, -2, 2), (-3, -3, -3), (1, 2, -1)]:
        idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3)
        A = oracle.dense_block(order, i, j, k).T
        M = rng.standard_normal(order ** 3)
        got = c.apply_m2l(idx, 1.0, M)
        ref = A @ M
        assert np.linalg.norm(got - ref) <= 1e-4 * np.linalg.norm(ref)
        # scale is linear and accumulation adds
        got2 = c.apply_m2l(idx, 3.0, M, L=got)
        assert np.allclose(got2, 4 * got, rtol=1e-13, atol=1e-13 * np.abs(got).max())


def _leaf_particles(rng, center, width, n):
    return [center[d] + (rng.random(n) - 0.5) * width for d in range(3)]


@pytest.mark.parametrize("order", [5, 7])
def test_m2m_p2m_identity(oracle, order):
    c = oracle.Cheb(order)
    rng = np.random.default_rng(order)
    pc = np.array([0.3, -0.2, 0.7])
    pw = 0.5
    Mp_direct = np.zeros(order ** 3)
    Mp = np.zeros(order ** 3)
    for child in range(8):
        off = np.array([(child >> 2) & 1, (child >> 1) & 1, child & 1]) - 0.5
        cc = pc + off * pw / 2
        x, y, z = _leaf_particles(rng, cc, pw / 2, 9)
        q = rng.standard_normal(9)
        Mc = c.p2m(cc, pw / 2, x, y, z, q)
        Mp = c.m2m(child, Mc, Mp)
        Mp_direct += c.p2m(pc, pw, x, y, z, q)
    assert np.max(np.abs(Mp - Mp_direct)) <= 1e-13 * np.abs(Mp_direct).max()


@pytest.mark.parametrize("order", [5, 7])
def test_l2l_l2p_identity(oracle, order):
    c = oracle.Cheb(order)
    rng = np.random.default_rng(10 + order)
    pc = np.array([0.3, -0.2, 0.7])
    pw = 0.5
    Lp = rng.standard_normal(order ** 3)
    for child in range(8):
        off = np.array([(child >> 2) & 1, (child >> 1) & 1, child & 1]) - 0.5
        cc = pc + off * pw / 2
        x, y, z = _leaf_particles(rng, cc, pw / 2, 7)
        q = rng.standard_normal(7)
        Lc = c.l2l(child, Lp)
        a = c.l2p(cc, pw / 2, Lc, x, y, z, q)
        b = c.l2p(pc, pw, Lp, x, y, z, q)
        for u, v in zip(a, b):
            assert np.max(np.abs(u - v)) <= 1e-12 * max(np.abs(v).max(), 1e-300)


def test_l2p_gradient_fd(oracle):
    order = 6
    c = oracle.Cheb(order)
    rng = np.random.default_rng(2)
    ctr = np.array([0.1, 0.2, 0.3])
    w = 0.25
    L = rng.standard_normal(order ** 3)
    x, y, z = _leaf_particles(rng, ctr, w * 0.8, 5)
    q = rng.standard_normal(5)
    pot, fx, fy, fz = c.l2p(ctr, w, L, x, y, z, q)
    h = 1e-6
    for d, f in enumerate((fx, fy, fz)):
        p = [x.copy(), y.copy(), z.copy()]
        m = [x.copy(), y.copy(), z.copy()]
        p[d] += h
        m[d] -= h
        fd = (c.l2p(ctr, w, L, *p, q)[0] - c.l2p(ctr, w, L, *m, q)[0]) / (2 * h)
        assert np.allclose(f, q * fd, rtol=1e-6, atol=1e-6 * np.abs(q * fd).max())


def test_p2m_l2p_interpolate_polynomial(oracle):
    """L2P reproduces a polynomial of degree < order sampled at the nodes."""
    order = 5
    c = oracle.Cheb(order)
    r = oracle.roots(order)
    ctr = np.array([0.0, 0.0, 0.0])
    w = 2.0
    n = np.arange(order ** 3)
    X = r[n % order]; Y = r[(n // order) % order]; Z = r[n // order ** 2]
    f = lambda a, b, cc: 1 + a ** 3 - 2 * a * b * cc ** 2 + b ** 4
    L = f(X, Y, Z)
    rng = np.random.default_rng(0)
    px, py, pz = (rng.random(10) * 2 - 1 for _ in range(3))
    pot, fx, _, _ = c.l2p(ctr, w, L, px, py, pz, np.ones(10))
    assert np.allclose(pot, f(px, py, pz), rtol=1e-12, atol=1e-12)
    assert np.allclose(fx, 3 * px ** 2 - 2 * py * pz ** 2, rtol=1e-11, atol=1e-11)
