"""Seeded particle distributions with the reference's generator.

The reference's loaders draw from POSIX drand48 (Src/Files/FRandomLoader.hpp:41,
93-95; Src/Files/FGenerateDistribution.hpp:27-44).  drand48 is the 48-bit LCG
x <- (0x5DEECE66D x + 0xB) mod 2^48, srand48(s) sets x = (s << 16) | 0x330E and
each draw returns x / 2^48; this module evaluates the same sequence in numpy
blocks by jumping ahead.

  uniform_cube  positions x,y,z = drand48() per particle in that order
                (FRandomLoader::fillParticle, FRandomLoader.hpp:85-90), then
                charges 2*drand48() - 1 in a second pass
  batch_sets    set sizes floor(nmin * ratio^u) from the first nsets draws,
                then positions of every particle (x, y, z per particle, sets
                concatenated, each in its own unit cube), then nrhs charge
                vectors 2*drand48() - 1 over all particles
  plummer       unit-sphere directions for all particles first, then Plummer
                radii with rejection above R (unifRandomPlummerLike,
                FGenerateDistribution.hpp:292-331, 395-410), mapped into the
                unit cube by x' = 0.5 + x / (2R)
"""
from __future__ import annotations

import numpy as np

_A = 0x5DEECE66D
_C = 0xB
_MASK = (1 << 48) - 1
_LO = (1 << 24) - 1


def _mulmod48(a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    al, ah = a & np.uint64(_LO), a >> np.uint64(24)
    bl, bh = b & np.uint64(_LO), b >> np.uint64(24)
    cross = (ah * bl + al * bh) & np.uint64(_LO)
    return (al * bl + (cross << np.uint64(24))) & np.uint64(_MASK)


class Drand48:
    """Vectorised drand48 stream (bit-identical to the C library's)."""

    BLOCK = 1 << 16

    def __init__(self, seed: int):
        self.state = ((int(seed) & 0xFFFFFFFF) << 16) | 0x330E
        # jump tables: state_j = A_j * s + C_j for j = 1..BLOCK
        Aj = np.empty(self.BLOCK, dtype=object)
        Cj = np.empty(self.BLOCK, dtype=object)
        a, c = 1, 0
        for j in range(self.BLOCK):
            a = (a * _A) & _MASK
            c = (c * _A + _C) & _MASK
            Aj[j] = a
            Cj[j] = c
        self._Aj = Aj.astype(np.uint64)
        self._Cj = Cj.astype(np.uint64)

    def draw(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float64)
        pos = 0
        while pos < n:
            m = min(self.BLOCK, n - pos)
            s = _mulmod48(self._Aj[:m], np.uint64(self.state))
            s = (s + self._Cj[:m]) & np.uint64(_MASK)
            out[pos:pos + m] = s.astype(np.float64) * (2.0 ** -48)
            self.state = int(s[-1])
            pos += m
        return out


def uniform_cube(n: int, seed: int, width: float = 1.0, center=(0.5, 0.5, 0.5)):
    """Positions in [c - w/2, c + w/2)^3 and charges U[-1, 1)."""
    g = Drand48(seed)
    xyz = g.draw(3 * n).reshape(n, 3)
    pos = [xyz[:, d] * width + center[d] - width / 2 for d in range(3)]
    q = 2.0 * g.draw(n) - 1.0
    return pos[0].copy(), pos[1].copy(), pos[2].copy(), q


def batch_sets(nsets: int, nrhs: int, seed: int, nmin: int = 1000, ratio: float = 64.0):
    """Sets of log-uniform sizes in [nmin, nmin * ratio], each uniform in the
    unit cube; returns (sizes, x, y, z, q) with q of shape (nrhs, n)."""
    g = Drand48(seed)
    sizes = np.floor(nmin * ratio ** g.draw(nsets)).astype(np.int64)
    n = int(sizes.sum())
    xyz = g.draw(3 * n).reshape(n, 3)
    q = 2.0 * g.draw(nrhs * n).reshape(nrhs, n) - 1.0
    return sizes, xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), q


def plummer(n: int, seed: int, R: float = 10.0):
    """Plummer sphere (scale 1) truncated at radius R, rescaled into [0, 1]^3."""
    g = Drand48(seed)
    uv = g.draw(2 * n).reshape(n, 2)
    theta = 2.0 * np.pi * uv[:, 0]
    phi = np.arccos(2 * uv[:, 1] - 1)
    sphi = np.sin(phi)
    dirs = np.stack([np.cos(theta) * sphi, np.sin(theta) * sphi, 2 * uv[:, 1] - 1], axis=1)
    radii = np.empty(n)
    filled = 0
    while filled < n:
        u = g.draw(max(1024, int((n - filled) * 1.02)))
        uu = u ** (2.0 / 3.0)
        r = np.sqrt(uu / (1.0 - uu))
        ok = r[r <= R]
        take = min(len(ok), n - filled)
        # rejection consumes draws in order; the stream position after the
        # last accepted draw is what the sequential loop would have reached
        if take < len(ok):
            last = np.nonzero(r <= R)[0][take - 1]
            unused = len(u) - (last + 1)
            g.state = _rewind(g, unused)
        radii[filled:filled + take] = ok[:take]
        filled += take
    pts = dirs * radii[:, None]
    x, y, z = (0.5 + pts[:, d] / (2 * R) for d in range(3))
    q = 2.0 * g.draw(n) - 1.0
    return x.copy(), y.copy(), z.copy(), q


def _rewind(g: Drand48, k: int) -> int:
    """State of the stream k draws earlier (inverse LCG step)."""
    a_inv = pow(_A, -1, 1 << 48)
    s = g.state
    for _ in range(k):
        s = ((s - _C) * a_inv) & _MASK
    return s


# ---- the geometries of Files/FGenerateDistribution.hpp ---------------------------
# Every generator draws from a drand48 stream in the reference's order; the
# rejection loops consume their draws one by one, and a stream handed in is left
# where the reference's sequential loop would leave it.

class _Seq:
    """Draws one value at a time from a Drand48 stream (block-buffered)."""

    def __init__(self, g: Drand48):
        self.g = g
        self.buf = np.empty(0)
        self.i = 0

    def next(self) -> float:
        if self.i >= len(self.buf):
            self.buf = self.g.draw(4096)
            self.i = 0
        v = float(self.buf[self.i])
        self.i += 1
        return v

    def close(self):
        unused = len(self.buf) - self.i
        if unused:
            self.g.state = _rewind(self.g, unused)
        self.buf = np.empty(0)
        self.i = 0


def points_in_cube(g: Drand48, n: int, lx: float = 1.0, ly: float = 1.0, lz: float = 1.0):
    """unifRandomPointsInCube (FGenerateDistribution.hpp:60-71): x, y, z per point."""
    p = g.draw(3 * n).reshape(n, 3)
    return p * np.array([lx, ly, lz])


def points_in_ball(g: Drand48, n: int, R: float):
    """unifRandomPointsInBall (:80-96): three draws per attempt, kept when |p| < R."""
    s = _Seq(g)
    out = np.empty((n, 3))
    for i in range(n):
        while True:
            p = [(s.next() - 0.5) * 2 * R for _ in range(3)]
            if p[0] * p[0] + p[1] * p[1] + p[2] * p[2] < R * R:
                break
        out[i] = p
    s.close()
    return out


def points_on_sphere(g: Drand48, n: int, R: float):
    """unifRandomPointsOnSphere (:291-308): theta = 2 pi u, z = (2 v - 1) R."""
    uv = g.draw(2 * n).reshape(n, 2)
    theta = 2.0 * np.pi * uv[:, 0]
    sphi = np.sin(np.arccos(2 * uv[:, 1] - 1))
    return np.stack([np.cos(theta) * sphi * R, np.sin(theta) * sphi * R, (2 * uv[:, 1] - 1) * R], axis=1)


def points_on_hyperpara(g: Drand48, n: int, a: float, b: float, c: float):
    """unifRandomPointsOnHyperPara (:267-279): (a u, b v, c (u^2 - v^2)), u, v in [-1, 1)."""
    uv = 2.0 * g.draw(2 * n).reshape(n, 2) - 1.0
    u, v = uv[:, 0], uv[:, 1]
    return np.stack([a * u, b * v, c * (u * u - v * v)], axis=1)


def _ellipsoid_rotation():
    """The fixed rotation of nonunifRandomPointsOnElipsoid (:117-143): Rz(pi/4) Ry(pi/8)."""
    al, om = np.pi / 8, np.pi / 4
    ry = np.array([[np.cos(al), 0.0, np.sin(al)], [0.0, 1.0, 0.0], [-np.sin(al), 0.0, np.cos(al)]])
    rz = np.array([[np.cos(om), -np.sin(om), 0.0], [np.sin(om), np.cos(om), 0.0], [0.0, 0.0, 1.0]])
    rot = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            s = 0.0
            for k in range(3):
                s += rz[i][k] * ry[k][j]
            rot[i][j] = s
    return rot


def points_on_ellipsoid(g: Drand48, n: int, a: float, b: float, c: float = 1.0):
    """nonunifRandomPointsOnElipsoid (:108-166, the default branch): x in [-a, a)
    accepted with density growing tenfold towards the tips, a circle of radius
    b sqrt(1 - x^2/a^2) around the x axis, then the fixed rotation.  As in the
    reference, c does not enter this branch: the surface is (x/a)^2 + (y^2 + z^2)/b^2 = 1
    before the rotation."""
    rot = _ellipsoid_rotation()
    s = _Seq(g)
    out = np.empty((n, 3))
    max_density = 10.0
    max_perimeter = 2.0 * np.pi * a
    for i in range(n):
        while True:
            px = (s.next() * a * 2.0) - a
            coef = abs(px) * max_density / a + 1.0
            subr = np.sqrt((1.0 - ((px * px) / (a * a))) * (b * b))
            if not (s.next() * max_perimeter > subr * coef):
                break
        omega = s.next() * 2.0 * np.pi
        py, pz = np.cos(omega) * subr, np.sin(omega) * subr
        out[i] = [px * rot[r][0] + py * rot[r][1] + pz * rot[r][2] for r in range(3)]
    s.close()
    return out


def points_on_prolate(g: Drand48, n: int, a: float, c: float):
    """unifRandomPointsOnProlate (:192-225): a point of the spheroid a:a:c from
    (u, v), then one more draw decides, as the reference writes the test, whether
    it is drawn again."""
    e = (a * a * a * a) / (c * c * c * c)
    s = _Seq(g)
    out = np.empty((n, 3))
    for i in range(n):
        while True:
            u = 2.0 * s.next() - 1.0
            v = 2.0 * np.pi * s.next()
            w = np.sqrt(1 - u * u)
            p = (a * w * np.cos(v), a * w * np.sin(v), c * u)
            ksi = s.next() / a
            if not ((p[0] * p[0] + p[1] * p[1] + e * p[2] * p[2]) < ksi * ksi):
                break
        out[i] = p
    s.close()
    return out


def points_plummer(g: Drand48, n: int, R: float):
    """unifRandomPlummer (:345-362): unit-sphere directions, then radius
    sqrt(1 / (m^(-2/3) - 1)) / r_max * R with r_max the radius of m = 0.8; each
    point's mass (physical value) 1/n."""
    dirs = points_on_sphere(g, n, 1.0)
    r_max = np.sqrt(1.0 / (0.8 ** (-2.0 / 3.0) - 1.0))
    m = g.draw(n)
    with np.errstate(divide="ignore"):
        r = np.sqrt(1.0 / (m ** (-2.0 / 3.0) - 1.0)) / r_max * R
    return dirs * r[:, None], np.full(n, 1.0 / n)


def generate(geometry: str, n: int, seed: int, charge: bool = False, zeromean: bool = False, radius: float = 2.0,
             size=(1.0, 1.0, 2.0), extra_length: float = 0.0):
    """Examples/generateDistributions.cpp: physical values first (U[0,1), or
    U[-1,1) with `charge`, optionally shifted to zero mean), then the points;
    the generators that call initRandom (cube, ball, sphere, Plummer) restart the
    stream from the seed, the others continue it.  Returns (data (n, 4): x y z q,
    box centre, box width) as the program writes them."""
    g = Drand48(seed)
    a, b = (2.0, -1.0) if charge else (1.0, 0.0)
    q = a * g.draw(n) + b
    if zeromean:
        q = q - np.sum(q) / n
    A, B, C = (float(v) for v in size)
    center = (0.0, 0.0, 0.0)
    if geometry == "unitCube":
        pts, width, center = points_in_cube(Drand48(seed), n), 1.0, (0.5, 0.5, 0.5)
    elif geometry == "ball":
        pts, width = points_in_ball(Drand48(seed), n, radius), 2.0 * radius
    elif geometry == "cuboid":
        width = max(A, B, C)
        pts, center = points_in_cube(Drand48(seed), n, A, B, C), (width / 2,) * 3
    elif geometry == "unitSphere":
        pts, width = points_on_sphere(Drand48(seed), n, 1.0), 2.0
    elif geometry == "sphere":
        pts, width = points_on_sphere(Drand48(seed), n, radius), 2.0 * radius
    elif geometry == "prolate":
        pts, width = points_on_prolate(g, n, A, C), 2.0 * C
    elif geometry == "hyperpara":
        pts, width = points_on_hyperpara(g, n, A, B, C), 2.0 * max(A, B, C)
    elif geometry == "ellipsoid":
        pts, width = points_on_ellipsoid(g, n, A, B, C), 2.0 * max(A, B, C)
    elif geometry == "plummer":
        pts, q = points_plummer(Drand48(seed), n, radius)
        width = 2.0 * radius
    else:
        raise ValueError(f"unknown geometry {geometry!r}")
    width += 2 * extra_length
    return np.column_stack([pts, q]), center, width
