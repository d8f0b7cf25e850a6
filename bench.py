#!/usr/bin/env python3
"""Factor + solve throughput of the RS path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config wavy262k]

A step is one recursive-skeletonization factorization plus one solve
(SkelFactorization::skeletonize + linear_solve, timed as in
bie_time_trial, ie_solver/linear_solve.cpp:156-185): tree build, geometry
and data generation happen before the timed region, and the geometry and
right-hand side are resident in HBM when it starts.  Prints one JSON line.

The CPU baseline runs the oracle (oracle/, a from-scratch restatement of
the reference using LAPACK from scipy's OpenBLAS) on the same workload on
the host cores.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oracle"))  # checker + CPU baseline only
sys.path.insert(0, os.path.join(ROOT, "ie-solver_amd"))

import geometry as G  # noqa: E402

METRIC = "factor+solve unknowns/s, mixed+learned vs plain FP64 GPU, at 1e-10 rel err"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3

CONFIGS = {
    # Laplace N=262144 on the seeded wavy boundary (the headline workload).
    "wavy262k": dict(shape="wavy", n=262144, seed=0, tol=1e-10, leaf=64),
    # Laplace N=65536 star r(t)=0.3(1+0.2cos 7t).
    "star65k": dict(shape="star", n=65536, seed=0, tol=1e-10, leaf=64),
    # Laplace N=4096 circle r=0.5 (the reference's CPU-runnable case).
    "circle4k": dict(shape="circle", n=4096, seed=0, tol=1e-10, leaf=64),
}


def make_workload(cfg):
    if cfg["shape"] == "wavy":
        g = G.wavy(cfg["n"], seed=cfg["seed"])
        ys, q = G.exterior_sources(cfg["seed"])
        f = G.harmonic_data(g["points"], ys, q)
    elif cfg["shape"] == "star":
        g = G.star(cfg["n"])
        ys, q = G.exterior_sources(cfg["seed"])
        f = G.harmonic_data(g["points"], ys, q)
    else:
        g = G.circle(cfg["n"])
        ys, q = np.array([[2.0, 2.0]]), np.array([1.0])
        f = G.harmonic_data(g["points"], ys, q)
    return g, f, ys, q


def stage_work(info, stage):
    """Algorithmic work of one stage summed over all boxes: (amount, unit).

    assemble: bytes of ID matrices written, 8 m n per box;
    qrcp: flops of the reference's ID, 2 m n^2 - 2/3 n^3 (full pivoted QR)
          + k^2 r (triangular solve);
    decouple: 6 s^2 r + 8 r^2 s + 2/3 r^3 flops (s = rank, r = n - s);
    sweep_up / sweep_down: factor bytes read, 8 (3 s r + r^2) per box.
    """
    m = info["m_rows"].astype(np.float64)
    n = info["n_active"].astype(np.float64)
    k = info["rank"].astype(np.float64)
    comp = k > 0
    r = np.where(comp, n - k, 0.0)
    if stage == "assemble":
        return float(np.sum(8.0 * m * n)), "bytes"
    if stage == "qrcp":
        has = m > 0
        return float(np.sum(np.where(has, 2 * m * n * n - (2.0 / 3.0) * n ** 3, 0) +
                            np.where(comp, k * k * r, 0))), "flops"
    if stage == "decouple":
        return float(np.sum(np.where(comp, 6 * k * k * r + 8 * r * r * k + (2.0 / 3.0) * r ** 3, 0))), "flops"
    if stage in ("sweep_up", "sweep_down"):
        return float(np.sum(np.where(comp, 8.0 * (3 * k * r + r * r), 0))), "bytes"
    return None, None


def cpu_baseline(g, f, cfg, threads, reps):
    import oracle
    o = oracle.OracleProblem(g, leaf_cap=cfg["leaf"])
    times = []
    mu = None
    for i in range(reps + 1):
        o = oracle.OracleProblem(g, leaf_cap=cfg["leaf"])  # fresh tree, untimed
        t0 = time.perf_counter()
        o.factor(cfg["tol"], threads)
        mu, _ = o.solve(f, threads)
        t1 = time.perf_counter()
        if i > 0:
            times.append(t1 - t0)
    return statistics.median(times), mu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="wavy262k", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-profile", action="store_true",
                    help="skip per-stage HIP events (roofline fields become null)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the RS factorization here runs on a single "
              "MI355X (multi-GPU sharding is design-only); rerun with --gpus 1.",
              file=sys.stderr)
        return 2
    cfg = CONFIGS[args.config]
    g, f, ys, q = make_workload(cfg)
    N = g["points"].shape[0]

    import torch
    import iesolver

    dev = torch.device("cuda:0")
    s = iesolver.RSSolver(0)
    s.set_geometry(g, iesolver.PDE_LAPLACE)
    s.build_tree(cfg["leaf"])
    f_dev = torch.from_numpy(f).to(dev)
    mu_dev = torch.empty_like(f_dev)
    torch.cuda.synchronize()

    def step():
        s.factor(cfg["tol"], iesolver.FP64)
        s.solve_device(f_dev.data_ptr(), mu_dev.data_ptr(), None, None)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if not args.no_profile:
        s.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    prof = s.profile_table() if not args.no_profile else {}
    s.set_profiling(False)
    mu = mu_dev.cpu().numpy()
    info = s.box_info()
    nl, nb, nd, _ = s.tree_info()

    # Accuracy: interior field vs the analytic harmonic function.
    import oracle
    tg = G.interior_points(1, 100, radius=0.1)
    u = oracle.evaluate(g, tg, mu)
    ue = G.harmonic_data(tg, ys, q)
    interior_err = float(np.linalg.norm(u - ue) / np.linalg.norm(ue))

    roofline = None
    stages_ms = {}
    if prof:
        for name, (ms, launches) in prof.items():
            if launches:
                stages_ms[name] = round(ms / args.steps, 4)
        cands = []
        for name in ("assemble", "qrcp", "decouple", "sweep_up", "sweep_down"):
            ms, launches = prof[name]
            work, unit = stage_work(info, name)
            if launches and work:
                cands.append((ms, name, launches, work, unit))
        if cands:
            ms, name, launches, work, unit = max(cands)
            per_launch = work * args.steps / launches   # algorithmic work per launch
            avg_s = ms / launches * 1e-3
            if unit == "bytes":
                ach = per_launch / avg_s / 1e9
                roofline = dict(kernel=name, bound="hbm", achieved=round(ach, 3),
                                peak=PEAK_HBM_GBS, unit="GB/s", frac=round(ach / PEAK_HBM_GBS, 6),
                                traffic=None, launches=launches,
                                avg_launch_us=round(avg_s * 1e6, 3),
                                work_per_launch=per_launch)
            else:
                ach = per_launch / avg_s / 1e12
                roofline = dict(kernel=name, bound="mfma", achieved=round(ach, 4),
                                peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                                frac=round(ach / PEAK_FP64_TFLOPS, 6), traffic=None,
                                launches=launches, avg_launch_us=round(avg_s * 1e6, 3),
                                work_per_launch=per_launch)

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        t_cpu, mu_cpu = cpu_baseline(g, f, cfg, threads, args.cpu_reps)
        cpu = dict(value=round(N / t_cpu, 1), unit="unknowns/s", cores=threads, kind="port",
                   sample=f"full workload ({args.config}, N={N}), median of {args.cpu_reps} "
                          f"factor+solve runs after 1 warm-up, OpenMP over boxes, BLAS threads=1",
                   seconds=round(t_cpu, 4))
        parity = float(np.linalg.norm(mu - mu_cpu) / np.linalg.norm(mu_cpu))

    out = {
        "metric": METRIC,
        "value": round(N / (ms_per_step * 1e-3), 1),
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"laplace_{args.config}_plain_fp64", "pde": "laplace",
                   "n_unknowns": N, "id_tol": cfg["tol"], "leaf_dofs": cfg["leaf"],
                   "levels": nl, "boxes": nb, "form": "plain_fp64",
                   "parallelism": "single_gpu"},
        "rel_err_interior": interior_err,
        "rel_err_vs_cpu_oracle": parity,
        "stages_ms": stages_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
