"""Benchmark of the FP64 Chebyshev 1/r FMM on one MI355X.

A step is one pass of the hot path over the resident particles: build the
octree on the device (keys, sort, leaves, levels, P2P and M2L lists) and
evaluate P2M/M2M/M2L/L2L/L2P/P2P, with the particles already in HBM.

Default workload: 1,000,000 particles uniform in the unit cube (drand48 seed
2), charges U[-1,1], height 7, order 7, epsilon 1e-7, FP64.

Prints one JSON line (rank 0).  `roofline` is the M2L kernel (the dominant
one): algorithmic flops 4 p^3 k_c - k_c - p^3 per interaction (the reference's
count, Src/Kernels/Chebyshev/FChebFlopsSymKernel.hpp:64-65) over the HIP-event
time of its launches in the timed region, against the FP64 MFMA peak.
`cpu_baseline` times the FP64 OpenMP oracle on a 100,000-particle sample with
the same particles per leaf (height 6) on all host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mparticles/s, 1/r Cheb FMM @ rel err 1e-6 (mixed vs FP64 GPU); %HBM/MFMA peak"
FP64_MFMA_PEAK_TFLOPS = 78.6


def _canonical(i, j, k):
    """Canonical operator index of relative position (i,j,k): sorted absolute values
    |i| >= |j| >= |k| (Src/Kernels/Interpolation/FInterpSymmetries.hpp:98-125)."""
    u = sorted((abs(i), abs(j), abs(k)), reverse=True)
    return (u[0] + 3) * 49 + (u[1] + 3) * 7 + (u[2] + 3)


def algorithmic_m2l_flops(engine, height, order):
    p3 = order ** 3
    rank_of = {}
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                if max(abs(i), abs(j), abs(k)) > 1:
                    c = _canonical(i, j, k)
                    if c not in rank_of:
                        U, _ = engine.operator(c)
                        rank_of[c] = U.shape[1]
                    rank_of[(i + 3) * 49 + (j + 3) * 7 + (k + 3)] = rank_of[c]
    flops_per_pos = np.zeros(343)
    for pos in range(343):
        if pos in rank_of:
            kc = rank_of[pos]
            flops_per_pos[pos] = 4 * p3 * kc - kc - p3
    total = 0.0
    count = 0
    for level in range(2, height):
        _, _, pos = engine.m2l_list(level)
        total += float(np.bincount(pos, minlength=343) @ flops_per_pos)
        count += len(pos)
    return total, count


def direct_sample(x, y, z, q, targets):
    """FP64 direct sum on a few targets (numpy, chunked over sources)."""
    pot = np.zeros(len(targets))
    fx = np.zeros(len(targets))
    for a, t in enumerate(targets):
        dx = x[t] - x
        dy = y[t] - y
        dz = z[t] - z
        r2 = dx * dx + dy * dy + dz * dz
        r2[t] = np.inf
        inv = 1.0 / np.sqrt(r2)
        pot[a] = np.dot(q, inv)
        fx[a] = -q[t] * np.dot(q, inv ** 3 * dx)
    return pot, fx


def cpu_baseline(seconds_budget=30.0):
    from oracle import pyoracle
    from scalfmm_amd.distributions import uniform_cube
    n, h, p = 100_000, 6, 7
    x, y, z, q = uniform_cube(n, 2)
    cheb = pyoracle.Cheb(p)
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    st = pyoracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p, cheb=cheb, nthreads=cores)[4]
    wall = time.perf_counter() - t0
    return {"value": n / st.t_total / 1e6, "unit": "Mparticles/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 OpenMP FMM (tree + all passes, operators prebuilt), {n} uniform particles, "
                      f"height {h}, order {p} (same particles per leaf as the 1M/height-7 workload); "
                      f"{st.t_total:.2f} s, M2L {st.t_m2l:.2f} s, P2P {st.t_p2p:.2f} s, wall {wall:.2f} s"}


def profiled_traffic(n, h, p):
    """HBM bytes per M2L launch from the committed rocprofv3 summary of this
    workload (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, see the summary), or None."""
    path = os.path.join(ROOT, "profiles", f"m2l_fp64_uniform{n // 1_000_000}M_h{h}_p{p}.md")
    if n % 1_000_000 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(mbytes=512, reps=3):
    """Elementwise copy of known size (read N bytes, write N bytes) for scaling
    the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch
    n = mbytes * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--height", type=int, default=7)
    ap.add_argument("--order", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 512 MiB elementwise copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only "
                                   "(see DESIGN.md), so --gpus must be 1"}))
        sys.exit(2)

    from scalfmm_amd import Engine
    from scalfmm_amd.distributions import uniform_cube

    n, h, p = args.n, args.height, args.order
    x, y, z, q = uniform_cube(n, 2)
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p)
    e.set_particles(x, y, z, q)  # host -> HBM once, outside the timed region
    for _ in range(args.warmup):
        e.build_tree()
        e.evaluate()
    e.synchronize()
    e.set_profiling(True)
    e.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e.build_tree()
        e.evaluate()
    e.synchronize()
    t1 = time.perf_counter()
    times = e.kernel_times()
    e.set_profiling(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step / 1e3) / 1e6

    flops, n_m2l = algorithmic_m2l_flops(e, h, p)
    m2l_ms, m2l_launches = times.get("m2l", (float("nan"), 0))
    m2l_ms_step = m2l_ms / args.steps
    achieved = flops / (m2l_ms_step / 1e3) / 1e12 if m2l_ms_step > 0 else float("nan")

    acc = None
    if not args.no_accuracy:
        pot, fx, _, _ = e.results()
        rng = np.random.default_rng(5)
        t = np.sort(rng.choice(n, 200, replace=False))
        rp, rfx = direct_sample(x, y, z, q, t)
        acc = {"targets": len(t), "pot_rel_l2": float(np.linalg.norm(pot[t] - rp) / np.linalg.norm(rp)),
               "fx_rel_l2": float(np.linalg.norm(fx[t] - rfx) / np.linalg.norm(rfx))}

    out = {
        "metric": METRIC, "value": value, "unit": "Mparticles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"uniform-{n}-h{h}-p{p}-fp64", "particles": n, "height": h, "order": p,
                   "epsilon": 10.0 ** -p, "precision": "fp64", "distribution": "uniform unit cube, drand48 seed 2"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": profiled_traffic(n, h, p),
                     "traffic_source": "profiles/ rocprofv3 FETCH_SIZE/WRITE_SIZE summary", "kernel": "k_m2l_f64",
                     "algorithmic_flops_per_step": flops, "m2l_interactions": n_m2l,
                     "m2l_ms_per_step": m2l_ms_step, "m2l_launches_per_step": m2l_launches / args.steps},
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in sorted(times.items())},
        "accuracy_vs_direct": acc,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
