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
