"""GPU parity of the HIP engine (through the C-ABI) against the CPU oracle.

Bars:
  - Morton keys, sort permutation, cells, P2P and M2L lists: bit-exact;
  - compressed operators: same ranks, U V^T within 1e-13 (relative to max);
  - FP64 potentials / forces: relative L2 <= 1e-12 against the oracle (the
    FP64 form's tolerance), and the reference's published FMM errors on the
    20k fixture reproduced;
  - at the full 1M / height 7 / order 7 size: tree bit-exact against the
    oracle and FMM error against FP64 direct sums on sampled targets.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _engine(h, p, width=1.0, center=(0.5, 0.5, 0.5), eps=0.0):
    from scalfmm_amd import Engine
    return Engine(h, width, center, p, eps)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _clustered(n, seed):
    rng = np.random.default_rng(seed)
    pts = np.clip(rng.normal(0.35, 0.1, size=(n, 3)), 0.0, 1.0)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), rng.uniform(-1, 1, n)


def test_direct_matches_fixture_and_oracle(oracle, fixture20k):
    """The device FP64 direct sum against the fixture's stored direct results
    and the oracle's direct sum."""
    d = fixture20k["data"]
    x, y, z, q = (d[:, k].copy() for k in range(4))
    e = _engine(5, 5)
    e.set_particles(x, y, z, q)
    t = np.array([0, 1, 17, 5000, 12345, 19999, 7, 8, 9])
    got = e.direct(t)
    ref = oracle.direct(x, y, z, q, targets=t)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert _rel(a, b) <= 1e-13, k
        assert _rel(a, d[t, 4 + k]) <= 1e-12, k


def test_mfma_f64_layout():
    e = _engine(3, 5)
    D = e.selftest_mfma_f64()
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = 1.0 + 4 * i + k
    B = 0.5 + 16 * np.arange(4)[:, None] + np.arange(16)[None, :]
    assert np.array_equal(D, A @ B)


@pytest.mark.parametrize("order", [5, 6, 7, 8])
def test_operators_match_oracle(oracle, order):
    e = _engine(3, order)
    c = oracle.Cheb(order)
    assert e.canonical_ranks() == c.canonical_ranks()
    for (i, j, k) in c.canonical_ranks():
        idx = (i + 3) * 49 + (j + 3) * 7 + (k + 3)
        U, V = e.operator(idx)
        Uo, Vo = c.K(idx)
        A, Ao = U @ V.T, Uo @ Vo.T
        assert np.abs(A - Ao).max() <= 1e-13 * np.abs(Ao).max()


def _check_tree(oracle, e, x, y, z, h):
    t = oracle.Tree(x, y, z, [0.5, 0.5, 0.5], 1.0, h)
    keys, perm = e.sorted_keys()
    okeys, operm = t.sorted()
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)  # both stable: ties in input order
    for level in range(2, h):
        assert np.array_equal(e.cell_keys(level), t.cell_keys(level)), level
    s, cnt = e.leaf_ranges()
    os_, oc = t.leaf_ranges()
    assert np.array_equal(s, os_) and np.array_equal(cnt, oc)
    for a, b in zip(e.p2p_list(), t.p2p_list()):
        assert np.array_equal(a, b)
    for level in range(2, h):
        for a, b in zip(e.m2l_list(level), t.m2l_list(level)):
            assert np.array_equal(a, b), level


def test_tree_bit_exact_fixture(oracle, fixture20k):
    d = fixture20k["data"]
    x, y, z = (d[:, k].copy() for k in range(3))
    e = _engine(5, 7)
    e.set_particles(x, y, z, d[:, 3])
    e.build_tree()
    _check_tree(oracle, e, x, y, z, 5)


@pytest.mark.parametrize("h", [4, 6])
def test_tree_bit_exact_clustered(oracle, h):
    x, y, z, q = _clustered(30000, h)
    e = _engine(h, 5)
    e.set_particles(x, y, z, q)
    e.build_tree()
    _check_tree(oracle, e, x, y, z, h)


def test_tree_duplicate_keys_keep_input_order(oracle):
    # many particles in a few leaves: the sort must be stable
    rng = np.random.default_rng(0)
    n = 10000
    x = rng.choice([0.1, 0.6, 0.95], n)
    y = rng.choice([0.2, 0.7], n)
    z = np.full(n, 0.3)
    e = _engine(4, 5)
    e.set_particles(x, y, z, np.ones(n))
    e.build_tree()
    _check_tree(oracle, e, x, y, z, 4)


@pytest.mark.parametrize("ops_name", ["p2p", "far", "all"])
def test_fmm_fixture_order7(oracle, fixture20k, ops_name):
    from scalfmm_amd import FFmmFarField, FFmmNearAndFarFields, FFmmP2P
    ops = {"p2p": FFmmP2P, "far": FFmmFarField, "all": FFmmNearAndFarFields}[ops_name]
    f = fixture20k
    d = f["data"]
    x, y, z, q = (d[:, k].copy() for k in range(4))
    _say(f"fixture20k h5 p7 ops={ops}")
    e = _engine(5, 7)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(ops)
    got = e.results()
    ref = oracle.fmm(x, y, z, q, f["center"], f["width"], 5, 7, ops=ops)
    for a, b in zip(got, ref[:4]):
        assert _rel(a, b) <= FP64_TOL


def test_fmm_fixture_published_errors(oracle, fixture20k):
    from test_oracle_fmm import PUBLISHED
    f = fixture20k
    d = f["data"]
    e = _engine(5, 7)
    e.set_particles(*(d[:, k].copy() for k in range(4)))
    e.build_tree()
    e.evaluate()
    pot, fx, fy, fz = e.results()
    for name, col, got in (("pot", 4, pot), ("fx", 5, fx), ("fy", 6, fy), ("fz", 7, fz)):
        acc = oracle.accuracy(d[:, col], got)
        assert abs(acc["rel_l2"] / PUBLISHED[name][0] - 1) < 0.01, (name, acc)


@pytest.mark.parametrize("ops_name", ["p2p", "all"])
def test_fmm_plummer_big_leaves(oracle, ops_name):
    """Clustered input whose core leaves hold hundreds to thousands of particles
    (the LDS-staged P2P path) next to sparse outer leaves."""
    from scalfmm_amd import FFmmNearAndFarFields, FFmmP2P
    from scalfmm_amd.distributions import plummer
    ops = {"p2p": FFmmP2P, "all": FFmmNearAndFarFields}[ops_name]
    n, h, p = 60000, 5, 5
    x, y, z, q = plummer(n, 21)
    _say(f"plummer n={n} h={h} p={p} ops={ops}")
    e = _engine(h, p)
    e.set_particles(x, y, z, q)
    e.build_tree()
    counts = e.leaf_ranges()[1]
    _say(f"  largest leaf {counts.max()}, leaves above 64: {(counts > 64).sum()} of {len(counts)}")
    assert counts.max() > 8 * 64
    e.evaluate(ops)
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p, ops=ops)
    for a, b in zip(got, ref[:4]):
        assert _rel(a, b) <= FP64_TOL


@pytest.mark.parametrize("order,h,n", [(5, 5, 20000), (6, 4, 8000), (8, 4, 6000), (7, 6, 60000)])
def test_fmm_uniform_orders(oracle, order, h, n):
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(n, 100 + order)
    _say(f"uniform n={n} h={h} p={order}")
    e = _engine(h, order)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, order)
    for a, b in zip(got, ref[:4]):
        assert _rel(a, b) <= FP64_TOL


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_config_uniform20k_h5_p5(oracle, mode):
    """20,000 drand48 particles (seed 1), height 5, order 5: the FP64 form
    within 1e-12 of the oracle, both forms against the O(N^2) direct sum
    (order-5 truncation error, 8.1e-4 for the potential)."""
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED, Engine
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(20000, 1)
    e = Engine(5, 1.0, (0.5, 0.5, 0.5), 5, 0.0, 0, PRECISION_FP64 if mode == "fp64" else PRECISION_MIXED)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, 5, 5)
    direct = oracle.direct(x, y, z, q)
    vs_oracle = [_rel(a, b) for a, b in zip(got, ref[:4])]
    vs_direct = [_rel(a, b) for a, b in zip(got, direct)]
    _say(f"  {mode}: vs oracle {max(vs_oracle):.2e}, vs direct pot {vs_direct[0]:.2e} forces {max(vs_direct[1:]):.2e}")
    assert max(vs_oracle) <= (FP64_TOL if mode == "fp64" else 3e-6)
    assert vs_direct[0] <= 1e-3 and max(vs_direct[1:]) <= 1.2e-3


@pytest.mark.parametrize("levels", [(2, 5), (3, 5), (2, 4), (3, 4), (4, 5), (5, 5)])
def test_fmm_working_levels(oracle, levels):
    """hfmm_evaluate_levels against the oracle's FAbstractAlgorithm::execute(ops,
    upper, lower) semantics (FP64, 1e-12)."""
    from scalfmm_amd.distributions import uniform_cube
    x, y, z, q = uniform_cube(20000, 77)
    h, p = 5, 6
    _say(f"working levels {levels}: n=20000 h={h} p={p}")
    e = _engine(h, p)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate(levels=levels)
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p, levels=levels)
    for a, b in zip(got, ref[:4]):
        assert _rel(a, b) <= FP64_TOL
    with pytest.raises(Exception):
        e.evaluate(levels=(1, h))
    with pytest.raises(Exception):
        e.evaluate(levels=(4, 3))


def _canonical():
    return [(i + 3) * 49 + (j + 3) * 7 + (k + 3) for i in (2, 3) for j in range(i + 1) for k in range(j + 1)]


def _write_operator_file(path, ops):
    """ops: {index: (U, V)} written as the reference's cache file
    (FChebSymM2LHandler.hpp:471-507): int index, int rank, U, V column-major."""
    with open(path, "wb") as f:
        for idx in sorted(ops):
            U, V = ops[idx]
            f.write(np.array([idx, U.shape[1]], np.int32).tobytes())
            f.write(np.ascontiguousarray(U.T).tobytes())
            f.write(np.ascontiguousarray(V.T).tobytes())


def test_operator_file_round_trip(oracle, tmp_path):
    """hfmm_save_operators writes the reference's layout; operators written by
    the oracle in that layout and loaded into a context reproduce the oracle's
    FMM; a saved file loaded into a fresh context gives the same bits."""
    from scalfmm_amd.distributions import uniform_cube
    p, h = 6, 4
    e = _engine(h, p)
    path = tmp_path / "sym2l.bin"
    e.save_operators(path)
    raw = path.read_bytes()
    off, seen = 0, []
    while off < len(raw):
        idx, r = np.frombuffer(raw[off:off + 8], np.int32)
        off += 8
        K = np.frombuffer(raw[off:off + 16 * r * p ** 3], np.float64)
        off += 16 * r * p ** 3
        U, V = e.operator(int(idx))
        assert np.array_equal(K[: r * p ** 3].reshape(r, -1).T, U)
        assert np.array_equal(K[r * p ** 3:].reshape(r, -1).T, V)
        seen.append(int(idx))
    assert seen == sorted(_canonical())
    x, y, z, q = uniform_cube(8000, 91)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    a = e.results()
    f = _engine(h, p)
    f.load_operators(path)
    f.set_particles(x, y, z, q)
    f.build_tree()
    f.evaluate()
    assert all(np.array_equal(u, v) for u, v in zip(a, f.results()))
    # the oracle's own operators through the file
    cheb = oracle.Cheb(p)
    _write_operator_file(tmp_path / "orc.bin", {i: cheb.K(i) for i in _canonical()})
    g = _engine(h, p)
    g.load_operators(tmp_path / "orc.bin")
    g.set_particles(x, y, z, q)
    g.build_tree()
    g.evaluate()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p)
    for u, v in zip(g.results(), ref[:4]):
        assert _rel(u, v) <= FP64_TOL
    # a file without one canonical operator is refused, and the context keeps working
    ops = {i: cheb.K(i) for i in _canonical()[1:]}
    _write_operator_file(tmp_path / "bad.bin", ops)
    with pytest.raises(Exception):
        g.load_operators(tmp_path / "bad.bin")
    (tmp_path / "cut.bin").write_bytes(raw[: len(raw) // 2])
    with pytest.raises(Exception):
        g.load_operators(tmp_path / "cut.bin")


def test_fmm_clustered_expansions(oracle):
    x, y, z, q = _clustered(20000, 5)
    h, order = 6, 6
    _say("clustered expansions h6 p6")
    e = _engine(h, order)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    c = oracle.Cheb(order)
    run = oracle.Run(c, x, y, z, q, (0.5, 0.5, 0.5), 1.0, h)
    for level in range(2, h):
        M, L = e.expansions(level)
        Mo, Lo = run.expansions(level)
        assert np.abs(M - Mo).max() <= 1e-13 * np.abs(Mo).max(), level
        assert np.abs(L - Lo).max() <= 1e-12 * np.abs(Lo).max(), level
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, order, cheb=c)
    for a, b in zip(got, ref[:4]):
        assert _rel(a, b) <= FP64_TOL


def test_repeat_and_reuse(oracle):
    """Two evaluations and a second particle set on one context."""
    from scalfmm_amd.distributions import uniform_cube
    e = _engine(4, 5)
    for n, seed in ((7000, 1), (3000, 2)):
        x, y, z, q = uniform_cube(n, seed)
        e.set_particles(x, y, z, q)
        e.build_tree()
        e.evaluate()
        a = e.results()
        e.evaluate()
        b = e.results()
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
        ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, 4, 5)
        for u, v in zip(a, ref[:4]):
            assert _rel(u, v) <= FP64_TOL


def test_edge_cases(oracle):
    from scalfmm_amd import FmmError
    # a single particle: no interactions at all
    e = _engine(3, 5)
    e.set_particles([0.3], [0.3], [0.3], [1.0])
    e.build_tree()
    e.evaluate()
    assert all(np.all(r == 0) for r in e.results())
    # two particles in one leaf and on the upper faces of the box
    e.set_particles([1.0, 0.99], [1.0, 0.98], [1.0, 0.97], [1.0, -2.0])
    e.build_tree()
    e.evaluate()
    pot, fx, fy, fz = e.results()
    ref = oracle.direct(np.array([1.0, 0.99]), np.array([1.0, 0.98]), np.array([1.0, 0.97]), np.array([1.0, -2.0]))
    assert np.allclose(pot, ref[0], rtol=1e-14) and np.allclose(fz, ref[3], rtol=1e-14)
    # outside the box -> error, not a crash
    e.set_particles([0.5, 1.5], [0.5, 0.5], [0.5, 0.5], [1.0, 1.0])
    with pytest.raises(FmmError) as exc:
        e.build_tree()
    assert exc.value.code == 2
    with pytest.raises(FmmError):
        e.evaluate()  # tree not built


def test_full_size_tree_and_accuracy(oracle):
    """1,000,000 uniform particles, height 7, order 7 (the FP64 drop-in size)."""
    from scalfmm_amd.distributions import uniform_cube
    n, h, p = 1_000_000, 7, 7
    x, y, z, q = uniform_cube(n, 2)
    _say("full size 1M h7 p7: tree")
    e = _engine(h, p)
    e.set_particles(x, y, z, q)
    e.build_tree()
    _check_tree(oracle, e, x, y, z, h)
    _say("full size 1M h7 p7: evaluate")
    e.evaluate()
    pot, fx, fy, fz = e.results()
    # FMM truncation error at order 7 against FP64 direct sums on sampled
    # targets (charges of both signs: ~2e-6 on the potential)
    rng = np.random.default_rng(11)
    t = np.sort(rng.choice(n, 1000, replace=False))
    ref = oracle.direct(x, y, z, q, targets=t)
    assert _rel(pot[t], ref[0]) < 1e-5
    for a, b in zip((fx[t], fy[t], fz[t]), ref[1:]):
        assert _rel(a, b) < 1e-4
    # FP64 parity with the oracle's FMM on every particle
    _say("full size 1M h7 p7: oracle FMM")
    oref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p)
    for a, b in zip((pot, fx, fy, fz), oref[:4]):
        assert _rel(a, b) <= FP64_TOL
