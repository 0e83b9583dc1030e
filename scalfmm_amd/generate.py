"""Particle distributions written as .fma / .bfma files, with the options of the
reference's Examples/generateDistributions.cpp:

    python -m scalfmm_amd.generate -N 100000 -ball -radius 2 -charge -zeromean -fout ball.bfma -seed 3

Geometries: -unitCube, -cuboid, -ball, -unitSphere, -sphere, -prolate,
-hyperpara, -ellipsoid, -plummer; -size a:b:c, -radius R, -charge,
-zeromean, -extraLength L.  -seed s (default 1) seeds the drand48 stream
(the reference seeds from the clock).
"""
from __future__ import annotations

import argparse
import sys

from .distributions import generate
from .fma import FFmaGenericWriter

GEOMETRIES = ("unitCube", "cuboid", "ball", "unitSphere", "sphere", "prolate", "hyperpara", "ellipsoid", "plummer")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m scalfmm_amd.generate", prefix_chars="-")
    ap.add_argument("-N", "-nb", "--number-of-particles", dest="n", type=int, default=20000)
    ap.add_argument("-fout", "--output-filename", dest="fout", default="unifPointDist.bfma")
    ap.add_argument("-size", default="1:1:2")
    ap.add_argument("-radius", type=float, default=None)
    ap.add_argument("-charge", action="store_true")
    ap.add_argument("-zeromean", action="store_true")
    ap.add_argument("-extraLength", type=float, default=0.0)
    ap.add_argument("-seed", type=int, default=1)
    for g in GEOMETRIES:
        ap.add_argument("-" + g, action="store_true")
    a = ap.parse_args(argv)
    chosen = [g for g in GEOMETRIES if getattr(a, g)]
    if len(chosen) != 1:
        print("Bad geometry option", file=sys.stderr)
        return 2
    geometry = chosen[0]
    radius = a.radius if a.radius is not None else 2.0  # the program's default for every geometry
    size = tuple(float(v) for v in a.size.split(":"))
    data, center, width = generate(geometry, a.n, a.seed, a.charge, a.zeromean, radius, size, a.extraLength)
    w = FFmaGenericWriter(a.fout)
    w.writeHeader(center, width, a.n, 8, 4)
    w.writeArrayOfReal(data, 4, a.n)
    print(f"Write {a.n} particles to '{a.fout}' ({geometry}, box width {width}, centre {center})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
