"""The geometries of Files/FGenerateDistribution.hpp and the generator program
(Examples/generateDistributions.cpp), pinned by what the geometry fixes: points
on the surface they are drawn on (closed forms), the inverse-CDF identity of the
Plummer radii, the stream position after rejection loops, and the file the
program writes."""
import os

import numpy as np

from scalfmm_amd.distributions import (Drand48, _ellipsoid_rotation, generate, points_in_ball, points_in_cube,
                                       points_on_ellipsoid, points_on_hyperpara, points_on_prolate,
                                       points_on_sphere, points_plummer)


def test_cube_and_sphere():
    p = points_in_cube(Drand48(3), 5000, 1.0, 2.0, 3.0)
    assert np.all(p >= 0) and np.all(p < [1.0, 2.0, 3.0])
    assert np.array_equal(p / [1.0, 2.0, 3.0], Drand48(3).draw(15000).reshape(5000, 3))
    s = points_on_sphere(Drand48(4), 20000, 2.5)
    assert np.allclose(np.linalg.norm(s, axis=1), 2.5, rtol=0, atol=1e-13)
    for d in range(3):  # uniform on the sphere: each coordinate has mean 0 and variance R^2 / 3
        assert abs(s[:, d].mean()) < 4 * 2.5 / np.sqrt(3 * 20000)
        assert abs(s[:, d].var() - 2.5 ** 2 / 3) < 0.05 * 2.5 ** 2 / 3


def test_ball():
    b = points_in_ball(Drand48(5), 20000, 2.0)
    r = np.linalg.norm(b, axis=1)
    assert np.all(r < 2.0)
    frac = np.mean(r < 1.0)  # uniform in volume: 1/8 inside half the radius
    assert abs(frac - 0.125) < 4 * np.sqrt(0.125 * 0.875 / 20000)


def test_hyperpara_prolate_ellipsoid_surfaces():
    h = points_on_hyperpara(Drand48(6), 5000, 1.5, 2.0, 0.5)
    assert np.allclose(h[:, 2], 0.5 * ((h[:, 0] / 1.5) ** 2 - (h[:, 1] / 2.0) ** 2), atol=1e-14)
    pr = points_on_prolate(Drand48(7), 3000, 1.0, 2.0)
    assert np.allclose((pr[:, 0] ** 2 + pr[:, 1] ** 2) / 1.0 + pr[:, 2] ** 2 / 4.0, 1.0, atol=1e-13)
    rot = _ellipsoid_rotation()
    assert np.allclose(rot @ rot.T, np.eye(3), atol=1e-15)
    e = points_on_ellipsoid(Drand48(8), 3000, 2.0, 0.5, 7.0)
    loc = e @ rot  # back to the unrotated frame (rot is orthonormal)
    assert np.allclose((loc[:, 0] / 2.0) ** 2 + (loc[:, 1] ** 2 + loc[:, 2] ** 2) / 0.25, 1.0, atol=1e-12)
    assert np.mean(np.abs(loc[:, 0]) > 1.0) > 0.6  # density grows towards the tips


def test_plummer_radii_invert_the_mass_profile():
    """r = sqrt(1 / (m^(-2/3) - 1)) / r_max * R is the inverse of the Plummer mass
    fraction M(s) = s^3 / (1 + s^2)^(3/2) at s = r r_max / R."""
    n, R = 20000, 3.0
    pts, q = points_plummer(Drand48(9), n, R)
    assert np.all(q == 1.0 / n)
    s = np.linalg.norm(pts, axis=1) * np.sqrt(1.0 / (0.8 ** (-2.0 / 3.0) - 1.0)) / R
    m = s ** 3 / (1 + s ** 2) ** 1.5
    g = Drand48(9)
    g.draw(2 * n)
    assert np.allclose(m, g.draw(n), rtol=1e-10, atol=1e-14)


def test_rejection_leaves_the_stream_where_a_sequential_loop_would(monkeypatch):
    """After a rejection loop the stream continues right after the last value the
    loop used (the block-buffered draws it did not use are given back)."""
    import scalfmm_amd.distributions as dist
    used = [0]

    class CountingSeq(dist._Seq):
        def next(self):
            used[0] += 1
            return super().next()
    monkeypatch.setattr(dist, "_Seq", CountingSeq)
    for gen in (lambda g: dist.points_in_ball(g, 300, 1.0), lambda g: dist.points_on_prolate(g, 300, 1.0, 3.0),
                lambda g: dist.points_on_ellipsoid(g, 300, 1.0, 0.4)):
        used[0] = 0
        g = Drand48(11)
        gen(g)
        after = g.draw(4)
        ref = Drand48(11)
        ref.draw(used[0])
        assert used[0] > 900 and np.array_equal(after, ref.draw(4))


def test_generator_program(tmp_path):
    from scalfmm_amd.fma import FFmaGenericLoader
    from scalfmm_amd.generate import main
    out = tmp_path / "ball.bfma"
    assert main(["-N", "2000", "-ball", "-radius", "3", "-charge", "-zeromean", "-fout", str(out), "-seed", "5"]) == 0
    ld = FFmaGenericLoader(str(out))
    assert ld.getNumberOfParticles() == 2000 and ld.getBoxWidth() == 6.0 and ld.getCenterOfBox() == (0, 0, 0)
    assert np.all(np.linalg.norm(ld.data[:, :3], axis=1) < 3.0)
    assert abs(ld.data[:, 3].sum()) < 1e-10 and np.all(np.abs(ld.data[:, 3]) <= 2.0)
    data, center, width = generate("unitCube", 100, 5)
    assert center == (0.5, 0.5, 0.5) and width == 1.0
    assert np.array_equal(data[:, 3], Drand48(5).draw(100))      # values first from the stream
    assert np.array_equal(data[:, :3], Drand48(5).draw(300).reshape(100, 3))  # initRandom restarts it
    assert main(["-N", "10", "-fout", str(tmp_path / "x.fma")]) == 2  # no geometry
