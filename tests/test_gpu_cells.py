"""Cells kept between evaluations (hfmm_set_cell_mode(HFMM_CELLS_KEPT)) on the
GPU: the reference's tree keeps its cells between executions and every pass
adds into them (FChebInterpolator.hpp:627-700, 560-582; FChebSymKernel.hpp:
193-206, 210-328, 331-350), so a sequence of execute(ops[, upper, lower])
calls composes.

Bars: split pass sequences equal one fresh evaluation bit for bit (each pass
runs once from zero cells: x + 0 = x); sequences with repeated passes, working
levels and resets within 1e-12 (relative L2) of the oracle's persistent run,
summing the engine's per-evaluation results as the reference's containers
sum them; several charge vectors keep their own cells."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H, P, C = 4, 5, (0.5, 0.5, 0.5)
P2P, P2M, M2M, M2L, L2L, L2P = 1, 2, 4, 8, 16, 32
FAR = P2M | M2M | M2L | L2L | L2P
ALL = FAR | P2P


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def parts():
    from scalfmm_amd.distributions import uniform_cube
    return uniform_cube(4000, 33)


def _engine(parts, precision=0, keep=True, q=None):
    from scalfmm_amd import _abi
    from scalfmm_amd.fmm import Engine
    x, y, z, q0 = parts
    e = Engine(H, 1.0, C, P, precision=precision)
    if q is None:
        e.set_particles(x, y, z, q0)
    else:
        e.set_particle_sets(np.array([0, len(x)]), x, y, z, q)
    e.build_tree()
    e.set_cell_mode(_abi.CELLS_KEPT if keep else _abi.CELLS_FRESH)
    if precision == _abi.PRECISION_MIXED:
        e.set_m2l_variant(0)  # the same tile shape in every engine: bitwise comparable sums
    return e


def _sum_results(e, seq):
    acc = None
    for item in seq:
        ops, levels = item if isinstance(item, tuple) else (item, None)
        e.evaluate(ops, levels)
        r = [a.copy() for a in e.results()]
        acc = r if acc is None else [a + b for a, b in zip(acc, r)]
    return acc


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("seq", [[P2M | M2M, M2L | L2L | L2P], [P2M | M2M | M2L, L2L | L2P],
                                 [P2M, M2M, M2L, L2L, L2P]], ids=["up-down", "transfer-down", "one-by-one"])
def test_split_sequences_equal_one_evaluation(parts, precision, seq):
    fresh = _engine(parts, precision, keep=False)
    fresh.evaluate(FAR)
    ref = fresh.results()
    e = _engine(parts, precision)
    got = _sum_results(e, seq)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    fresh.close()
    e.close()


@pytest.mark.parametrize("name,seq", [
    ("twice", [ALL, ALL]),
    ("levels", [(ALL, (2, 3)), (ALL, (3, H))]),
    ("far-then-all", [FAR, ALL]),
    ("up-twice", [P2M | M2M, P2M | M2M, M2L | L2L | L2P]),
])
def test_sequences_match_oracle(oracle, parts, name, seq):
    x, y, z, q = parts
    cheb = oracle.Cheb(P)
    run = oracle.Run(cheb, x, y, z, q, C, 1.0, H, ops=0)
    for item in seq:
        ops, levels = item if isinstance(item, tuple) else (item, (2, H))
        ref = run.execute(ops, *levels)
    e = _engine(parts)
    got = _sum_results(e, seq)
    for k in range(4):
        assert _rel(got[k], ref[k]) < 1e-12, (name, k)
    e.close()


def test_reset_and_mode_switch(parts):
    from scalfmm_amd import _abi
    fresh = _engine(parts, keep=False)
    fresh.evaluate(ALL)
    ref = [a.copy() for a in fresh.results()]
    fresh.evaluate(ALL)  # fresh cells: evaluating again gives the same results
    for a, b in zip(fresh.results(), ref):
        np.testing.assert_array_equal(a, b)
    e = _engine(parts)
    e.evaluate(ALL)
    for a, b in zip(e.results(), ref):
        np.testing.assert_array_equal(a, b)
    e.evaluate(ALL)  # kept cells: the second evaluation differs (old locals, doubled multipoles)
    assert _rel(e.results()[0], ref[0]) > 1e-3
    e.reset_cells()
    M, L = e.expansions(2)
    assert not M.any() and not L.any()
    e.evaluate(ALL)
    for a, b in zip(e.results(), ref):
        np.testing.assert_array_equal(a, b)
    e.evaluate(ALL)
    e.set_cell_mode(_abi.CELLS_FRESH)  # back to fresh cells
    e.evaluate(ALL)
    for a, b in zip(e.results(), ref):
        np.testing.assert_array_equal(a, b)
    e.set_cell_mode(_abi.CELLS_KEPT)  # a new kept state starts from zero cells
    e.evaluate(ALL)
    for a, b in zip(e.results(), ref):
        np.testing.assert_array_equal(a, b)
    e.build_tree()  # a new tree: zero cells again
    e.evaluate(ALL)
    for a, b in zip(e.results(), ref):
        np.testing.assert_array_equal(a, b)
    fresh.close()
    e.close()


def test_several_charge_vectors_keep_their_cells(parts):
    x, y, z, q = parts
    q2 = np.stack([q, np.sin(5 * x)])
    e = _engine(parts, q=q2)
    got = _sum_results(e, [P2M | M2M, M2L | L2L | L2P])
    for v in range(2):
        f = _engine(parts, keep=False, q=q2[v:v + 1])
        f.evaluate(FAR)
        for k, ref in enumerate(f.results()):
            np.testing.assert_array_equal(got[k][v], ref)
        f.close()
    e.close()


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["fp64", "mixed", "hilo"])
def test_stream_overlap_same_results(parts, precision):
    """The P2P on the side stream beside the far field gives the results of
    the one-stream order bit for bit (it writes the results first, the L2P
    waits for it and adds)."""
    got = []
    for overlap in (False, True):
        e = _engine(parts, precision, keep=False)
        if precision:
            e.set_m2l_variant(0)
        e.set_stream_overlap(overlap)
        e.evaluate(ALL)
        e.evaluate(ALL)  # a second evaluation: the side stream's P2P of the first has finished
        got.append(e.results())
        e.close()
    for a, b in zip(*got):
        np.testing.assert_array_equal(a, b)
