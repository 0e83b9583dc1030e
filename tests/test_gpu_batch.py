"""Particle sets and batches (hfmm_set_particle_sets, hfmm_batch_*) through the C-ABI.

A forest of sets must give every set exactly what a context holding that set
alone gives: the set index only prefixes the keys, the lists never leave a
set and every pass sums a target's terms in the same order, so the bar is
bit-equality with the single-set engine (FP64 and mixed), plus the FP64
tolerance (1e-12 relative L2) against the oracle per set and per charge
vector.  Batches add the per-height grouping and the graph: every launch of
the captured graph, with other device work between launches, must repeat the
eager run's results bit for bit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _sets(sizes, nrhs, seed):
    """Uniform sets in the unit cube, charges U[-1, 1] per charge vector."""
    from scalfmm_amd.distributions import Drand48
    g = Drand48(seed)
    n = int(sum(sizes))
    pts = g.draw(3 * n).reshape(n, 3)
    q = 2.0 * g.draw(nrhs * n).reshape(nrhs, n) - 1.0
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return off, pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), q


def _single(h, p, precision, x, y, z, q, variant=None):
    from scalfmm_amd import Engine
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, precision)
    if variant is not None and precision:
        e.set_m2l_variant(variant)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    return e.results()


def _precision(mode):
    from scalfmm_amd import PRECISION_FP64, PRECISION_MIXED
    return PRECISION_FP64 if mode == "fp64" else PRECISION_MIXED


def test_forest_keys_prefix_the_set():
    """Sorted keys of a forest = (set << 3(h-1)) | the set's own keys, sets in order."""
    from scalfmm_amd import Engine
    h = 4
    sizes = [3000, 0, 5000, 800, 1]
    off, x, y, z, q = _sets(sizes, 1, 31)
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), 5)
    e.set_particle_sets(off, x, y, z, q)
    e.build_tree()
    keys, perm = e.sorted_keys()
    exp_k, exp_p = [], []
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        if a == b:
            continue
        f = Engine(h, 1.0, (0.5, 0.5, 0.5), 5)
        f.set_particles(x[a:b], y[a:b], z[a:b], q[0, a:b])
        f.build_tree()
        k, p = f.sorted_keys()
        exp_k.append((s << (3 * (h - 1))) | k)
        exp_p.append(p + a)
    assert np.array_equal(keys, np.concatenate(exp_k))
    assert np.array_equal(perm, np.concatenate(exp_p))
    assert e.tree_info().height == h + 1  # 5 sets fit one set level (8 >= 5)


@pytest.mark.parametrize("mode,variant", [("fp64", None), ("mixed", 0), ("mixed", 1), ("mixed", 2)])
def test_forest_matches_single_sets(mode, variant):
    from scalfmm_amd import Engine
    h, p, nrhs = 4, 5, 3
    sizes = [4000, 0, 2500, 9000, 700]
    off, x, y, z, q = _sets(sizes, nrhs, 32)
    _say(f"forest {mode} (M2L variant {variant}) sizes={sizes} nrhs={nrhs} h={h} p={p}")
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, _precision(mode))
    if variant is not None:
        e.set_m2l_variant(variant)
    e.set_particle_sets(off, x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    assert got[0].shape == (nrhs, off[-1])
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        if a == b:
            continue
        for r in range(nrhs):
            ref = _single(h, p, _precision(mode), x[a:b], y[a:b], z[a:b], q[r, a:b], variant)
            for k in range(4):
                assert np.array_equal(got[k][r, a:b], ref[k]), (s, r, k)


def test_forest_vs_oracle(oracle):
    from scalfmm_amd import Engine
    h, p, nrhs = 4, 6, 2
    sizes = [3000, 6000]
    off, x, y, z, q = _sets(sizes, nrhs, 33)
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p)
    e.set_particle_sets(off, x, y, z, q)
    e.build_tree()
    e.evaluate()
    got = e.results()
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        for r in range(nrhs):
            ref = oracle.fmm(x[a:b], y[a:b], z[a:b], q[r, a:b], (0.5, 0.5, 0.5), 1.0, h, p)
            errs = [_rel(got[k][r, a:b], ref[k]) for k in range(4)]
            _say(f"  set {s} rhs {r}: rel L2 vs oracle {max(errs):.2e}")
            assert max(errs) <= FP64_TOL, errs


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_batch_graph_replays(mode):
    """Sets of heights 3, 4 and 5 (default rule), two charge vectors; the
    eager run, three graph launches with other device work between them, and
    per-set single-engine runs all agree bit for bit."""
    from scalfmm_amd import BatchEngine, Engine
    from scalfmm_amd.distributions import uniform_cube
    sizes = [1200, 3000, 20000, 1500, 17000, 0, 2100]
    nrhs, p = 2, 5
    off, x, y, z, q = _sets(sizes, nrhs, 34)
    _say(f"batch {mode} sizes={sizes} nrhs={nrhs}")
    b = BatchEngine(p, sizes, nrhs=nrhs, precision=_precision(mode))
    assert list(b.heights) == [3, 4, 5, 3, 5, 3, 4]
    b.set_particles(x, y, z, q)
    b.build()
    assert b.info() == {"trees": 3, "graph_ready": False}
    b.evaluate()
    first = b.results()
    assert b.info()["graph_ready"]
    other = Engine(4, 1.0, (0.5, 0.5, 0.5), 5)
    ox, oy, oz, oq = uniform_cube(5000, 3)
    for rep in range(3):
        other.set_particles(ox, oy, oz, oq)
        other.build_tree()
        other.evaluate()
        other.results()
        b.evaluate()
        b.synchronize()
        again = b.results()
        for k in range(4):
            assert np.array_equal(again[k], first[k]), (rep, k)
    # the single-set engines use the M2L variant each tree's autotuner picked
    variant = {t.set_height: t.m2l_variant()[0] for t in b.trees()}
    for s in range(len(sizes)):
        lo, hi = off[s], off[s + 1]
        if lo == hi:
            continue
        for r in range(nrhs):
            ref = _single(b.heights[s], p, _precision(mode), x[lo:hi], y[lo:hi], z[lo:hi], q[r, lo:hi],
                          variant[b.heights[s]])
            for k in range(4):
                assert np.array_equal(first[k][r, lo:hi], ref[k]), (s, r, k)


def test_batch_rebuild_recaptures():
    """New particles after a graph exists: the batch rebuilds and recaptures."""
    from scalfmm_amd import BatchEngine
    sizes = [2500, 900]
    b = BatchEngine(5, sizes, nrhs=1)
    for seed in (40, 41):
        off, x, y, z, q = _sets(sizes, 1, seed)
        b.set_particles(x, y, z, q)
        b.build()
        b.evaluate()
        b.evaluate()
        got = b.results()
        for s in range(2):
            lo, hi = off[s], off[s + 1]
            ref = _single(b.heights[s], 5, 0, x[lo:hi], y[lo:hi], z[lo:hi], q[0, lo:hi])  # FP64: one M2L kernel
            assert np.array_equal(got[0][0, lo:hi], ref[0])


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_forest_working_levels(mode):
    """Working levels apply to every set's own tree: a forest evaluated over
    levels (3, 5) equals the single-set runs over the same levels, bit for bit."""
    from scalfmm_amd import Engine
    h, p = 5, 5
    sizes = [9000, 3000]
    off, x, y, z, q = _sets(sizes, 1, 35)
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, _precision(mode))
    if mode == "mixed":
        e.set_m2l_variant(0)
    e.set_particle_sets(off, x, y, z, q)
    e.build_tree()
    e.evaluate(levels=(3, 5))
    got = e.results()
    for s in range(2):
        a, b = off[s], off[s + 1]
        f = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, _precision(mode))
        if mode == "mixed":
            f.set_m2l_variant(0)
        f.set_particles(x[a:b], y[a:b], z[a:b], q[0, a:b])
        f.build_tree()
        f.evaluate(levels=(3, 5))
        for k, ref in enumerate(f.results()):
            assert np.array_equal(got[k][a:b], ref), (s, k)
