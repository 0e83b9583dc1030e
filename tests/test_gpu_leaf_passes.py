"""P2M and L2P at the boundaries of their work split.

P2M runs one wave per window of 64 sorted particles and walks the leaves that start in the window; a leaf
that runs past its window is summed in rounds of 64 from its own first particle. L2P runs one lane per
particle and reads its leaf's expansion directly. Leaves of chosen sizes in consecutive Morton order put
the leaf boundaries where that split has cases: a leaf that fills one window exactly, leaves of 65 to 200
particles spanning several windows, single-particle leaves at a window's first and last particle. The
multipoles and locals of the leaf level must equal the oracle's (1e-13 / 1e-12 of the largest entry), as
must the results, at an order whose coefficient count is a multiple of 64 (8) and one that is not (5).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 3  # leaf level 2: 4 x 4 x 4 leaves of width 1/4
SIZES = [1, 63, 64, 65, 62, 127, 128, 129, 2, 200, 3, 64, 1, 50, 14, 1]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _particles(seed):
    """SIZES[m] particles inside the leaf of Morton index m, charges U[-1, 1]."""
    from oracle import pyoracle
    from scalfmm_amd.distributions import Drand48
    g = Drand48(seed)
    pts, q = [], []
    for m, cnt in enumerate(SIZES):
        ix, iy, iz = pyoracle.demorton(m)
        u = g.draw(3 * cnt).reshape(cnt, 3)
        pts.append((np.array([ix, iy, iz]) + 0.05 + 0.9 * u) / 4.0)
        q.append(2.0 * g.draw(cnt) - 1.0)
    p = np.concatenate(pts)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.concatenate(q)


@pytest.mark.parametrize("order", [5, 8])
def test_leaf_passes_at_window_boundaries(oracle, order):
    from scalfmm_amd import Engine
    x, y, z, q = _particles(21)
    n = len(x)
    offsets = {int(s) % 64 for s in np.cumsum([0] + SIZES)[:-1]}
    assert n == 974 and {0, 1, 63} <= offsets  # leaves starting on a window's first, second and last particle
    _say(f"{n} particles in {len(SIZES)} leaves, order {order}")
    e = Engine(H, 1.0, (0.5, 0.5, 0.5), order)
    e.set_particles(x, y, z, q)
    e.build_tree()
    e.evaluate()
    c = oracle.Cheb(order)
    run = oracle.Run(c, x, y, z, q, (0.5, 0.5, 0.5), 1.0, H)
    M, L = e.expansions(H - 1)
    Mo, Lo = run.expansions(H - 1)
    assert M.shape == Mo.shape == (len(SIZES), order ** 3)
    assert np.abs(M - Mo).max() <= 1e-13 * np.abs(Mo).max()
    assert np.abs(L - Lo).max() <= 1e-12 * np.abs(Lo).max()
    got = e.results()
    ref = oracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, H, order, cheb=c)
    for a, b in zip(got, ref[:4]):
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
