"""Benchmark of the Chebyshev 1/r FMM on one MI355X.

A step is one pass of the hot path over the resident particles: build the
octree on the device (keys, sort, leaves, levels, P2P and M2L lists) and
evaluate P2M/M2M/M2L/L2L/L2P/P2P, with the particles already in HBM.

Default workload (the accuracy-targeted configuration): 10,000,000 particles
from a seeded Plummer sphere (scale 1, truncated at radius 10, rescaled into
the unit cube), charges U[-1,1], height 8, order 8.  `value` is the
mixed-precision form's throughput; the plain FP64 form runs on the same
particles beside it (`fp64_gpu`, `mixed_vs_fp64`), and both are checked
against an FP64 direct sum on 10,000 sampled targets (`accuracy_vs_direct`,
computed by the library's FP64 direct-sum kernel, outside the timed region).

`roofline` is the kernel with the largest share of the step, with its
algorithmic work over the HIP-event time of its launches in the timed region:
  - M2L: the reference's flop count 4 p^3 k - k - p^3 per interaction
    (Src/Kernels/Chebyshev/FChebFlopsSymKernel.hpp:62-63); the mixed form
    executes it as six bfloat16 partial products, so its algorithmic bf16 work
    is six times that count, against the bf16 dense MFMA peak;
  - P2P: 34 flops per target-source pair (FChebFlopsSymKernel.hpp:64-65)
    against the FP32 (mixed) or FP64 vector peak.
`cpu_baseline` times the FP64 OpenMP oracle on a bounded sample of the same
distribution with the same particles per leaf, on all host cores.
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
PEAK = {  # MI355X published peaks
    "fp64_mfma": 78.6, "bf16_mfma": 2500.0, "fp32_vector": 157.3, "fp64_vector": 78.6, "hbm_gbs": 8000.0,
}
WORKLOADS = {
    # name: (distribution, n, height, order, cpu sample (n, height)); the CPU
    # sample keeps the particles per leaf: n / 8^d particles d levels lower
    "plummer10M-h8-p8": ("plummer", 10_000_000, 8, 8, (1_250_000, 7)),
    "plummer10M-h8-p5": ("plummer", 10_000_000, 8, 5, (1_250_000, 7)),
    "uniform1M-h7-p7": ("uniform", 1_000_000, 7, 7, (1_000_000, 7)),
    # 4,096 sets of log-uniform sizes in [1e3, 6.4e4], heights 3..5 by size,
    # 8 charge vectors; the CPU sample is the first 512 sets with one vector
    "batch4096-p5-r8": ("batch", 4096, None, 5, (512, None)),
    # the reference's published kernel comparison (Data/ResultCompareKernels.txt:
    # Examples/compareAllPoissonKernels.cpp on unitCubeRef20kDouble.bfma, order 7,
    # height 5): kernel construction (operator set-up) plus one execution
    "compare20k-h5-p7": ("compare", 20_000, 5, 7, None),
}
# the reference's published times of that comparison (4 threads, 3 GHz Core i7)
PUBLISHED_COMPARE = {"symmetric": 1.50791, "global": 31.937, "uniform": 1.27096}
SEED = {"plummer": 3, "uniform": 2, "batch": 4}  # the configurations' documented seeds
BATCH_RHS = 8


def _canonical(i, j, k):
    """Canonical operator index of relative position (i,j,k): sorted absolute values
    |i| >= |j| >= |k| (Src/Kernels/Interpolation/FInterpSymmetries.hpp:98-125)."""
    u = sorted((abs(i), abs(j), abs(k)), reverse=True)
    return (u[0] + 3) * 49 + (u[1] + 3) * 7 + (u[2] + 3)


def algorithmic_m2l_flops(engine, height, order):
    """The reference's M2L flop count summed over this tree's interaction
    lists (interactions per relative position counted on the device)."""
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
        counts = engine.m2l_counts(level)
        total += float(counts @ flops_per_pos)
        count += int(counts.sum())
    return total, count


def p2p_pairs(engine):
    """Target-source pairs of the near field (self pairs excluded) by kernel:
    'small' = targets in leaves of at most 64 particles (k_p2p_*), 'big' = the
    other targets (the ordered big-leaf kernels); the mutual kernel's work is
    'big_big' ordered pairs between two big leaves (each unordered pair once)
    plus 'big_rest' = the big leaves' pairs with themselves and with small
    neighbours (ordered)."""
    _, cnt = engine.leaf_ranges()
    off, nb, _ = engine.p2p_list()
    owner = np.repeat(np.arange(len(cnt)), np.diff(off))
    cnt = cnt.astype(np.float64)
    big = cnt > 64
    nbr_sum = np.bincount(owner, weights=cnt[nb], minlength=len(cnt))
    nbr_big = np.bincount(owner, weights=cnt[nb] * big[nb], minlength=len(cnt))
    per_leaf = cnt * (cnt - 1) + cnt * nbr_sum
    big_big = (cnt * nbr_big)[big].sum()
    return {"small": float(per_leaf[~big].sum()), "big": float(per_leaf[big].sum()),
            "big_big": float(big_big), "big_rest": float(per_leaf[big].sum() - big_big)}


def make_particles(dist, n):
    from scalfmm_amd.distributions import plummer, uniform_cube
    return (plummer if dist == "plummer" else uniform_cube)(n, SEED[dist])


def cpu_baseline(dist, n, h, p):
    from oracle import pyoracle
    x, y, z, q = make_particles(dist, n)
    cheb = pyoracle.Cheb(p)
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    st = pyoracle.fmm(x, y, z, q, (0.5, 0.5, 0.5), 1.0, h, p, cheb=cheb, nthreads=cores)[4]
    wall = time.perf_counter() - t0
    return {"value": n / st.t_total / 1e6, "unit": "Mparticles/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 OpenMP FMM (tree + all passes, operators prebuilt), {n} {dist} particles, "
                      f"height {h}, order {p} (the same particles per leaf as the full workload); "
                      f"{st.t_total:.2f} s, M2L {st.t_m2l:.2f} s, P2P {st.t_p2p:.2f} s, wall {wall:.2f} s"}


# source file of each profiled kernel: a committed summary's traffic is used
# only while that file is unchanged (its sha256 is recorded in the summary)
KERNEL_SOURCE = {
    "k_p2p_mixed_big": "p2p_mixed.hip", "k_p2p_mixed_mutual": "p2p_mixed.hip", "k_m2l_mixed": "fmm_mixed.hip",
    "k_p2p_mixed": "p2p_mixed.hip", "k_m2l_f64": "fmm_device.hip", "k_p2p_f64_big": "fmm_device.hip",
    "k_p2p_f64": "fmm_device.hip", "k_unif_m2l": "fmm_global.hip",
}


def source_sha256(kernel):
    import hashlib
    src = KERNEL_SOURCE.get(kernel)
    if src is None:
        return None
    with open(os.path.join(ROOT, "scalfmm_amd", "csrc", src), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def profiled_traffic(kernel, workload, form=None):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy-kernel calibration) and a
    note; None when there is no summary or when the kernel's source changed
    after it was profiled.  A summary of one precision form is named
    ..._<workload>-<form>.md."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}-{form}.md")
    if not os.path.exists(path):
        path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None, "no profile summary"
    traffic, sha = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = float(line.split(":", 1)[1])
        if line.startswith("kernel_source_sha256:"):
            sha = line.split(":", 1)[1].strip()
    if sha is None or sha != source_sha256(kernel):
        return None, f"{os.path.basename(path)} was profiled on another version of the kernel source"
    return traffic, os.path.basename(path)


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


def run_form(parts, h, p, precision, steps, warmup):
    """Warm up, then time `steps` tree builds + evaluations.  The timed steps
    record HIP events on the engine's stream around the near-field and M2L
    kernels only (their rooflines); a separate profiled pass of the same
    steps gives the per-kernel breakdown."""
    from scalfmm_amd import Engine
    x, y, z, q = parts
    t0 = time.perf_counter()
    e = Engine(h, 1.0, (0.5, 0.5, 0.5), p, 0.0, 0, precision)  # operators: tables, device compression, packing
    e.setup_ms = (time.perf_counter() - t0) * 1e3
    e.set_particles(x, y, z, q)  # host -> HBM once, outside the timed region
    for _ in range(warmup):
        e.build_tree()
        e.evaluate()
    e.synchronize()
    e.set_profiling(True)
    e.set_profiling_filter(ROOFLINE_KERNELS)
    e.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.build_tree()
        e.evaluate()
    e.synchronize()
    t1 = time.perf_counter()
    live = e.kernel_times()
    e.set_profiling_filter(())
    e.reset_kernel_times()
    for _ in range(steps):
        e.build_tree()
        e.evaluate()
    times = e.kernel_times()
    e.set_profiling(False)
    ms = (t1 - t0) * 1e3 / steps
    times.update(live)  # the roofline kernels as timed in the timed steps
    return e, ms, {k: v[0] / steps for k, v in sorted(times.items())}, {k: v[1] / steps for k, v in times.items()}


# kernels whose HIP-event times the timed steps record (the roofline kernels)
ROOFLINE_KERNELS = ("p2p_mixed_mutual", "p2p_mixed_big", "p2p_big", "m2l_mixed", "m2l")


# M2L partial products per product of each form (bfloat16 MFMA in the mixed forms)
PRODUCTS = {"mixed": 6, "hilo": 3, "fp64": 1}
DTYPE = {"mixed": "bf16x3-mfma/f32 (M2L), f32 (P2P), f64 (rest)",
         "hilo": "bf16x2-mfma/f32 (M2L), f32 (P2P), f64 (rest)", "fp64": "f64"}


def precision_code(form):
    from scalfmm_amd import _abi
    return {"mixed": _abi.PRECISION_MIXED, "hilo": _abi.PRECISION_MIXED_HILO, "fp64": _abi.PRECISION_FP64}[form]


def roofline(engines, kernel_ms, launches, p, form, workload, nrhs=1, kernel=None):
    """Roofline of the kernel with the largest time share (or of `kernel`);
    `engines` are the trees the kernel times cover (one, or the per-height
    trees of a batch), each evaluated for `nrhs` charge vectors; `form` is
    mixed, hilo or fp64."""
    mixed = form != "fp64"
    m2l_name = "m2l_mixed" if mixed else "m2l"
    p2p_small, p2p_big = ("p2p_mixed", "p2p_mixed_big") if mixed else ("p2p", "p2p_big")
    top = kernel if kernel is not None else max(kernel_ms, key=kernel_ms.get)
    if top == m2l_name:
        flops = n_int = 0
        for e in engines:
            f, c = algorithmic_m2l_flops(e, e.height, p)
            flops += f * nrhs
            n_int += c * nrhs
        work = PRODUCTS[form] * flops
        peak = PEAK["bf16_mfma"] if mixed else PEAK["fp64_mfma"]
        kname, bound = ("k_m2l_mixed", "mfma") if mixed else ("k_m2l_f64", "mfma")
        extra = {"m2l_interactions": n_int, "fp64_equivalent_flops_per_step": flops,
                 "work": f"{PRODUCTS[form]} bf16 partial products per product" if mixed else "FP64 products"}
    elif top in (p2p_small, p2p_big, "p2p_mixed_mutual"):
        pairs = {"small": 0.0, "big": 0.0, "big_big": 0.0, "big_rest": 0.0}
        for e in engines:
            for k, v in p2p_pairs(e).items():
                pairs[k] += v * nrhs
        peak = PEAK["fp32_vector"] if mixed else PEAK["fp64_vector"]
        bound = "valu"
        if top == "p2p_mixed_mutual":
            # each unordered big-big pair once (39 flops, the reference's mutual
            # count), the other pairs of the big leaves one way (34 flops)
            work = 39.0 * pairs["big_big"] / 2 + 34.0 * pairs["big_rest"]
            kname = "k_p2p_mixed_mutual"
            extra = {"pairs": pairs["big"], "unordered_big_big_pairs": pairs["big_big"] / 2,
                     "ordered_other_pairs": pairs["big_rest"],
                     "work": "39 flops per unordered pair of two big leaves, 34 per other target-source pair "
                             "(FChebFlopsSymKernel.hpp:66-69)"}
        else:
            work = 34.0 * (pairs["big"] if top == p2p_big else pairs["small"])
            kname = {"p2p_mixed": "k_p2p_mixed", "p2p_mixed_big": "k_p2p_mixed_big", "p2p": "k_p2p_f64",
                     "p2p_big": "k_p2p_f64_big"}[top]
            extra = {"pairs": pairs["big"] if top == p2p_big else pairs["small"],
                     "work": "34 flops per target-source pair"}
    else:
        return {"kernel": top, "note": "no roofline model for this kernel"}
    ms = kernel_ms[top]
    achieved = work / (ms / 1e3) / 1e12
    traffic, traffic_src = profiled_traffic(kname, workload, form)
    out = {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
           "traffic": traffic, "traffic_source": traffic_src, "kernel": kname, "kernel_ms_per_step": ms,
           "launches_per_step": launches.get(top), "algorithmic_flops_per_launch": work / max(launches.get(top, 1), 1)}
    out.update(extra)
    return out


def run_batch_form(sets, p, precision, steps, warmup):
    """Build the per-height trees, run and capture the graph, then time
    `steps` graph launches; one more eager run with HIP events gives the
    kernel breakdown."""
    from scalfmm_amd import BatchEngine
    sizes, x, y, z, q = sets
    b = BatchEngine(p, sizes, nrhs=q.shape[0], precision=precision)
    b.set_particles(x, y, z, q)  # host -> HBM once
    t0 = time.perf_counter()
    b.build()
    b.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    b.evaluate()  # eager run + capture
    for _ in range(warmup):
        b.evaluate()
    b.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.evaluate()
    b.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    trees = b.trees()
    for t in trees:
        t.set_profiling(True)
        t.reset_kernel_times()
    b.evaluate()
    b.synchronize()
    kms, launches = {}, {}
    for t in trees:
        for k, (tot, cnt) in t.kernel_times().items():
            kms[k] = kms.get(k, 0.0) + tot
            launches[k] = launches.get(k, 0) + cnt
        t.set_profiling(False)
    return b, trees, ms, build_ms, dict(sorted(kms.items())), launches


def run_batch(args):
    """The batched configuration: every set and charge vector per graph launch."""
    from scalfmm_amd import Engine
    from scalfmm_amd.distributions import batch_sets
    dist, nsets, _, p, (csets, _) = WORKLOADS[args.workload]
    print(f"[bench] generating {nsets} sets x {BATCH_RHS} charge vectors", file=sys.stderr, flush=True)
    sets = batch_sets(nsets, BATCH_RHS, SEED[dist])
    sizes, x, y, z, q = sets
    n = int(sizes.sum())
    mixed = args.precision != "fp64"
    forms = [args.precision] + (["fp64"] if mixed and not args.no_fp64_compare else [])
    res = {}
    for form in forms:
        print(f"[bench] {form} batch: build, capture, {args.warmup} warmup + {args.steps} timed graph launches",
              file=sys.stderr, flush=True)
        b, trees, ms, build_ms, kms, launches = run_batch_form(
            sets, p, precision_code(form), args.steps, args.warmup)
        rl = roofline(trees, kms, launches, p, form, args.workload, nrhs=BATCH_RHS)
        acc = None
        if not args.no_accuracy:
            # FP64 direct sums on sampled sets, every charge vector
            print("[bench] direct sums on 16 sampled sets", file=sys.stderr, flush=True)
            pot = b.results(fields=("pot",))[0]
            rng = np.random.default_rng(5)
            errs = []
            for s in np.sort(rng.choice(nsets, 16, replace=False)):
                lo, hi = int(b.offsets[s]), int(b.offsets[s + 1])
                t = np.sort(rng.choice(hi - lo, min(200, hi - lo), replace=False))
                ref = Engine(3, 1.0, (0.5, 0.5, 0.5), p)
                for r in range(BATCH_RHS):
                    ref.set_particles(x[lo:hi], y[lo:hi], z[lo:hi], q[r, lo:hi])
                    d = ref.direct(t)[0]
                    errs.append(float(np.linalg.norm(pot[r, lo + t] - d) / np.linalg.norm(d)))
                ref.close()
            acc = {"sets": 16, "targets_per_set": 200, "pot_rel_l2_max": max(errs),
                   "pot_rel_l2_median": float(np.median(errs))}
        res[form] = {"value": n * BATCH_RHS / (ms / 1e3) / 1e6, "ms_per_step": ms, "build_ms": build_ms,
                     "kernel_ms_per_step": kms, "roofline": rl, "accuracy_vs_direct": acc,
                     "trees": [{"height": t.set_height, "sets": t.nsets, "particles": t.n,
                                "m2l_variant": t.m2l_variant()[0], "m2l_tune_ms": t.m2l_variant()[1]}
                               for t in trees]}
        for t in trees:
            t.close()
        b.close()
    main_form = res[args.precision]
    out = {
        "metric": "Mparticle-RHS/s, batched 1/r Cheb FMM (all sets and charge vectors per graph launch)",
        "value": main_form["value"], "unit": "Mparticle-RHS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": main_form["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": args.workload, "sets": nsets, "particles": n, "rhs": BATCH_RHS, "order": p,
                   "seed": SEED[dist], "heights": "smallest h in [3, 20] with n <= 32 * 8^(h-1)",
                   "precision": args.precision},
        "roofline": main_form["roofline"],
        "build_ms": main_form["build_ms"],
        "kernel_ms_per_step": main_form["kernel_ms_per_step"],
        "trees": main_form["trees"],
        "accuracy_vs_direct": main_form["accuracy_vs_direct"],
        "fp64_gpu": res.get("fp64") if mixed else None,
        "mixed_vs_fp64": (main_form["value"] / res["fp64"]["value"]) if "fp64" in res and mixed else None,
    }
    if not args.no_cpu_baseline:
        from oracle import pyoracle
        from scalfmm_amd import default_height
        print(f"[bench] cpu baseline (first {csets} sets, one charge vector)", file=sys.stderr, flush=True)
        cheb = pyoracle.Cheb(p)
        cores = os.cpu_count() or 1
        tot, m = 0.0, 0
        for s in range(csets):
            lo, hi = int(sets[0][:s].sum()), int(sets[0][:s + 1].sum())
            st = pyoracle.fmm(x[lo:hi], y[lo:hi], z[lo:hi], q[0, lo:hi], (0.5, 0.5, 0.5), 1.0,
                              default_height(hi - lo), p, cheb=cheb, nthreads=cores)[4]
            tot += st.t_total
            m += hi - lo
        out["cpu_baseline"] = {"value": m / tot / 1e6, "unit": "Mparticle-RHS/s", "cores": cores, "kind": "port",
                               "sample": f"oracle FP64 OpenMP FMM, the first {csets} sets ({m} particles), one "
                                         f"charge vector, one tree per set; {tot:.2f} s"}
    print(json.dumps(out))


def run_compare(args):
    """Kernel construction plus one FP64 execution on the reference's 20k
    fixture, for FChebSymKernel (the symmetric M2L, operators compressed on
    the device), FChebKernel (one global SVD, operators built on the host) and
    FUnifKernel (Lagrange basis, dense M2L between equispaced nodes), timed
    end to end from the host as the reference's comparison times them."""
    from scalfmm_amd import _abi
    from scalfmm_amd.fma import FFmaGenericLoader
    from scalfmm_amd.fmm import Engine
    _, n, h, p, _ = WORKLOADS[args.workload]
    f = FFmaGenericLoader(os.path.join(ROOT, "tests", "golden", "unitCubeRef20kDouble.bfma"))
    d = f.data
    x, y, z, q = (np.ascontiguousarray(d[:, i]) for i in range(4))
    w, c = f.getBoxWidth(), f.getCenterOfBox()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    res = {}
    for scheme in ("symmetric", "global", "uniform"):
        times = []
        for it in range(args.warmup + args.steps):
            print(f"[bench] {scheme} run {it}", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            e = Engine(h, w, c, p, m2l_scheme={"symmetric": _abi.M2L_SYMMETRIC, "global": _abi.M2L_GLOBAL_SVD,
                                                "uniform": _abi.M2L_UNIFORM}[scheme])
            e.set_particles(x, y, z, q)
            e.build_tree()
            e.evaluate()
            got = e.results()
            t1 = time.perf_counter()
            if it >= args.warmup:
                times.append(t1 - t0)
            if it == args.warmup + args.steps - 1:
                e.set_profiling(True)
                e.evaluate()
                kms = {k: v for k, v in e.kernel_times().items()}
                e.set_profiling(False)
            e.close()
        sec = float(np.mean(times))
        res[scheme] = {"seconds": sec, "published_seconds": PUBLISHED_COMPARE[scheme],
                       "speedup_vs_published": PUBLISHED_COMPARE[scheme] / sec,
                       "pot_rel_l2_vs_direct": rel(got[0], d[:, 4]),
                       "force_rel_l2_vs_direct": [rel(got[1 + k], d[:, 5 + k]) for k in range(3)],
                       "kernel_ms_one_evaluation": kms}
    main_t = res["symmetric"]["seconds"]
    out = {"metric": "s, FChebSymKernel construction + one execution, 20k uniform (unitCubeRef20kDouble), h5, p7",
           "value": main_t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": main_t * 1e3, "higher_is_better": False, "scaling": "strong",
           "vs_baseline": main_t / PUBLISHED_COMPARE["symmetric"], "dtype": "f64", "data": "reference fixture",
           "config": {"workload": args.workload, "particles": n, "height": h, "order": p,
                      "published": "Data/ResultCompareKernels.txt (4 threads, 3 GHz Core i7)"},
           "schemes": res,
           "roofline": None, "roofline_note": "the timed region is dominated by operator set-up, not by one kernel"}
    if not args.no_cpu_baseline:
        from oracle import pyoracle
        cores = os.cpu_count() or 1
        t0 = time.perf_counter()
        cheb = pyoracle.Cheb(p)
        st = pyoracle.fmm(x, y, z, q, c, w, h, p, cheb=cheb, nthreads=cores)[4]
        t1 = time.perf_counter()
        out["cpu_baseline"] = {"value": t1 - t0, "unit": "s", "cores": cores, "kind": "port",
                               "sample": f"the oracle's FP64 OpenMP FMM with its operator set-up on the same 20k "
                                         f"particles, h{h}, p{p} (execute {st.t_total:.3f} s)"}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plummer10M-h8-p8", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "hilo", "fp64"],
                    help="mixed: three-part bf16 splits in the M2L; hilo: two-part (hi/lo) splits; fp64")
    ap.add_argument("--no-fp64-compare", action="store_true", help="skip the FP64 form beside the mixed one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--targets", type=int, default=10_000, help="sampled targets of the direct-sum check")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 512 MiB elementwise copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only, "
                                   "so --gpus must be 1"}))
        sys.exit(2)

    if WORKLOADS[args.workload][0] == "batch":
        run_batch(args)
        return
    if WORKLOADS[args.workload][0] == "compare":
        run_compare(args)
        return
    dist, n, h, p, (cn, ch) = WORKLOADS[args.workload]
    print(f"[bench] generating {n} {dist} particles", file=sys.stderr, flush=True)
    parts = make_particles(dist, n)
    mixed = args.precision != "fp64"
    print(f"[bench] {args.precision} form: {args.warmup} warmup + {args.steps} timed steps", file=sys.stderr,
          flush=True)
    e, ms, kms, launches = run_form(parts, h, p, precision_code(args.precision), args.steps, args.warmup)
    m2l_variant = e.m2l_variant()
    setup_ms = e.setup_ms
    value = n / (ms / 1e3) / 1e6
    rl = roofline([e], kms, launches, p, args.precision, args.workload)
    m2l_key = "m2l_mixed" if mixed else "m2l"
    # the M2L's own roofline when another kernel dominates
    rl_m2l = (roofline([e], kms, launches, p, args.precision, args.workload, kernel=m2l_key)
              if m2l_key in kms and max(kms, key=kms.get) != m2l_key else None)
    got = e.results()
    direct = None
    if not args.no_accuracy:
        print(f"[bench] FP64 direct sum on {args.targets} sampled targets", file=sys.stderr, flush=True)
        t = np.sort(np.random.default_rng(5).choice(n, min(args.targets, n), replace=False))
        direct = (t,) + e.direct(t)[:2]
    e.close()

    fp64 = None
    got64 = None
    if mixed and not args.no_fp64_compare:
        print("[bench] fp64 form on the same particles", file=sys.stderr, flush=True)
        e64, ms64, kms64, launches64 = run_form(parts, h, p, precision_code("fp64"), args.steps, args.warmup)
        fp64 = {"value": n / (ms64 / 1e3) / 1e6, "ms_per_step": ms64, "kernel_ms_per_step": kms64,
                "roofline": roofline([e64], kms64, launches64, p, "fp64", args.workload)}
        got64 = e64.results()
        e64.close()

    acc = None
    if direct is not None:
        t, rp, rfx = direct

        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        acc = {"targets": len(t), args.precision: {"pot_rel_l2": rel(got[0][t], rp), "fx_rel_l2": rel(got[1][t], rfx)}}
        if got64 is not None:
            acc["fp64"] = {"pot_rel_l2": rel(got64[0][t], rp), "fx_rel_l2": rel(got64[1][t], rfx)}
            acc["mixed_vs_fp64_pot_rel_l2"] = rel(got[0], got64[0])
            acc["mixed_vs_fp64_fx_rel_l2"] = rel(got[1], got64[1])

    out = {
        "metric": METRIC, "value": value, "unit": "Mparticles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "distribution": dist,
                   "seed": SEED[dist], "height": h, "order": p, "epsilon": 10.0 ** -p, "precision": args.precision},
        "roofline": rl,
        "roofline_m2l": rl_m2l,
        "kernel_ms_per_step": kms,
        "m2l_variant": {"chosen": m2l_variant[0], "tune_ms": m2l_variant[1]} if mixed else None,
        "operator_setup_ms": setup_ms,
        "fp64_gpu": fp64,
        "mixed_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "accuracy_vs_direct": acc,
    }
    if not args.no_cpu_baseline:
        print(f"[bench] cpu baseline ({cn} particles, height {ch})", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(dist, cn, ch, p)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
