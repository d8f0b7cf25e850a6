#!/usr/bin/env python3
"""Factor + solve throughput of the RS path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config wavy262k]
                    [--form learned|randomized|plain]

A step is one recursive-skeletonization factorization plus one solve
(SkelFactorization::skeletonize + linear_solve, timed as in
bie_time_trial, ie_solver/linear_solve.cpp:156-185): tree build, geometry
and data generation happen before the timed region, and the geometry and
right-hand side are resident in HBM when it starts.  Prints one JSON line.

Forms: "plain" is the reference's algorithm in FP64 (full pivoted-QR IDs);
"randomized" takes each box's ID from a 32-row Gaussian sketch on the
matrix cores (full-QRCP fallback); "learned" (default) sizes each box's
sketch from the rank model, trained on the GPU before the timed region on
other boundaries (ie-solver_amd/rank_model.py).  The plain form is timed
beside the chosen one as its comparator.

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

# HBM bytes per launch of the dominant kernels of the learned form at
# wavy262k, from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE)
# calibrated on a copy kernel of known size: profiles/*_hbm_traffic.md.
try:
    with open(os.path.join(ROOT, "profiles", "hbm_traffic_wavy262k_learned.json")) as _f:
        TRAFFIC = {k: v["bytes_per_launch"] for k, v in json.load(_f).items()}
except (OSError, ValueError, KeyError):
    TRAFFIC = {}

CONFIGS = {
    # Laplace N=262144 on the seeded wavy boundary (the headline workload).
    "wavy262k": dict(shape="wavy", n=262144, seed=0, tol=1e-10, leaf=64),
    # Laplace N=65536 star r(t)=0.3(1+0.2cos 7t).
    "star65k": dict(shape="star", n=65536, seed=0, tol=1e-10, leaf=64),
    # Laplace N=4096 circle r=0.5 (the reference's CPU-runnable case).
    "circle4k": dict(shape="circle", n=4096, seed=0, tol=1e-10, leaf=64),
    # 256 independent Laplace problems, star boundaries with sizes
    # log-uniform in [2048, 32768], k in {3..9}, amplitude U[0.05, 0.25],
    # factored and solved together (one launch per stage per level); the
    # FP32 factorization + FP64 refinement form is timed beside them.
    "batch256": dict(shape="batch", n_problems=256, nmin=2048, nmax=32768, seed=0,
                     tol=1e-10, leaf=64, fp32_tol=1e-6, cpu_problems=24),
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


def roofline_from(prof, info, id_stats, steps, traffic_table):
    """Per-stage ms per step and the roofline of the stage with the most
    device time (algorithmic work per launch / average launch duration)."""
    roofline = None
    stages_ms = {}
    if prof:
        for name, (ms, launches) in prof.items():
            if launches:
                stages_ms[name] = round(ms / steps, 4)
        cands = []
        for name in ("assemble", "rid", "qrcp", "decouple", "sweep_up", "sweep_down"):
            ms, launches = prof[name]
            if name == "rid":
                # k_rid's work: the sketches on the matrix cores.
                work, unit = (id_stats["sketch_gflop"] * 1e9, "flops") if id_stats else (None, None)
            elif name == "qrcp" and id_stats is not None:
                # What is left to the full QRCP: its boxes' flops are not
                # separated from the sketched boxes', so it is not a candidate.
                work, unit = None, None
            else:
                work, unit = stage_work(info, name)
            if launches and work:
                cands.append((ms, name, launches, work, unit))
        defs = dict(assemble="8 m n bytes per box", decouple="6 s^2 r + 8 r^2 s + 2/3 r^3 flops",
                    sweep_up="8 (3 s r + r^2) factor bytes", sweep_down="8 (3 s r + r^2) factor bytes",
                    qrcp="2 m n^2 - 2/3 n^3 + k^2 r flops (full pivoted QR)",
                    rid="2 l m n16 sketch flops on the matrix cores (k_rid)")
        if cands:
            ms, name, launches, work, unit = max(cands)
            per_launch = work * steps / launches   # algorithmic work per launch
            avg_s = ms / launches * 1e-3
            traffic = traffic_table.get(name)
            if unit == "bytes":
                ach = per_launch / avg_s / 1e9
                roofline = dict(kernel=name, bound="hbm", achieved=round(ach, 3),
                                peak=PEAK_HBM_GBS, unit="GB/s", frac=round(ach / PEAK_HBM_GBS, 6),
                                traffic=traffic, launches=launches,
                                avg_launch_us=round(avg_s * 1e6, 3),
                                work_per_launch=per_launch, work_def=defs[name])
            else:
                ach = per_launch / avg_s / 1e12
                roofline = dict(kernel=name, bound="mfma", achieved=round(ach, 4),
                                peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                                frac=round(ach / PEAK_FP64_TFLOPS, 6), traffic=traffic,
                                launches=launches, avg_launch_us=round(avg_s * 1e6, 3),
                                work_per_launch=per_launch, work_def=defs[name])
    return stages_ms, roofline


def time_steps(step, s, args, profile):
    """ms per step of K steps after W warm-up steps, timed without HIP
    events; with profile, K more steps with per-stage HIP events around
    every launch give the per-stage device times (the events add gaps
    between launches, so they stay out of the timed steps)."""
    import torch
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = {}
    if profile:
        s.set_profiling(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        prof = s.profile_table()
        s.set_profiling(False)
    return (t1 - t0) * 1e3 / args.steps, prof


def tune(s, args, tol, precision):
    """With --autotune, tune the launch caps for the current form (untimed);
    returns the caps in use."""
    if args.autotune:
        s.autotune(tol, precision, reps=2)
    return [int(c) for c in s.launch_caps()]


def set_form(s, args):
    """Select the ID of the timed form; returns the rank model's record."""
    import iesolver
    model_info = None
    if args.form == "learned":
        import rank_model
        t0 = time.perf_counter()
        if args.rank_model:
            model = np.load(args.rank_model)
            model_info = dict(loaded=args.rank_model, margin=3)
        else:
            X, y = rank_model.collect(rank_model.training_cases())
            model, loss = rank_model.train(X, y)
            model_info = dict(training_boxes=int(len(y)), final_loss=round(float(loss), 5),
                              margin=3, seconds_untimed=round(time.perf_counter() - t0, 2),
                              trained_on="stars k=3,5,9 N=16384, wavy seeds 1,2 N=65536, "
                                         "Stokes holes N=16384 and star N=8192, "
                                         "tol 1e-6/8/10")
        s.set_rank_model(model, margin=3)
        s.set_id_method(iesolver.ID_RANDOMIZED, 0)
    elif args.form == "randomized":
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    else:
        s.set_id_method(iesolver.ID_QRCP)
    return model_info


def host_cpu():
    """The host's CPU model and core counts (the CPU baseline's cores)."""
    model, phys = None, set()
    try:
        with open("/proc/cpuinfo") as fh:
            cur = {}
            for line in fh:
                if ":" in line:
                    k, v = [x.strip() for x in line.split(":", 1)]
                    cur[k] = v
                    if k == "model name" and model is None:
                        model = v
                elif cur:
                    phys.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
            if cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(model=model, logical=os.cpu_count(), physical=len(phys) or None, usable=usable)


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


def time_update(s, cfg, args):
    """Incremental re-factorization after local deformations of the wavy
    boundary (untimed in the headline): per frame, a bump of 3e-4 r0 over
    an arc of 0.4 rad moves to a new place (larger bumps at this N sweep
    nodes into large empty leaves past twice the leaf capacity, and the
    tree is rebuilt); iegpu_update_geometry re-inserts the nodes and marks
    boxes, the next factor re-factors only the marked ones.  A full
    factorization of the same frame on the same tree gives the comparison
    time and solution."""
    import iesolver
    frames = []
    f_of = lambda g: G.harmonic_data(g["points"], *G.exterior_sources(cfg["seed"]))
    for i in range(max(args.steps, 2)):
        g1 = G.bump_wavy(cfg["n"], 3e-4 * (1 if i % 2 == 0 else -1), 1.0 + 0.5 * i,
                         seed=cfg["seed"])
        t0 = time.perf_counter()
        kept = s.update_geometry(g1)
        t_host = (time.perf_counter() - t0) * 1e3
        s.factor(cfg["tol"], iesolver.FP64)
        inc_ms = s.timings()["factor_ms"]
        n_dirty, inc, marks = s.update_info()
        f1 = f_of(g1)
        x_inc, _ = s.solve(f1)
        s.discard_update()  # the next factor covers every box
        s.factor(cfg["tol"], iesolver.FP64)
        full_ms = s.timings()["factor_ms"]
        x_full, _ = s.solve(f1)
        frames.append(dict(tree_kept=kept, incremental=inc, boxes_refactored=int(n_dirty),
                           boxes=int(marks.size), update_host_ms=round(t_host, 2),
                           incremental_factor_ms=round(inc_ms, 3), full_factor_ms=round(full_ms, 3),
                           rel_diff_vs_full=float(np.linalg.norm(x_inc - x_full) /
                                                  np.linalg.norm(x_full))))
    inc = [fr["incremental_factor_ms"] for fr in frames[1:]]
    full = [fr["full_factor_ms"] for fr in frames[1:]]
    return dict(frames=frames, median_incremental_factor_ms=float(np.median(inc)),
                median_full_factor_ms=float(np.median(full)),
                speedup=float(np.median(full) / np.median(inc)))


def mixed_subrecord(s, f, mu_exact, cfg, args):
    """The FP32 factorization (learned-rank randomized ID at tol 1e-6) as
    the preconditioner of FP64 iterative refinement to a relative residual
    of 1e-10, timed over the same steps as the headline (untimed in it).
    The residual's operator is the FP64 fast multipole apply (order 48).
    iegpu_solve_refined takes host vectors, so b and x cross PCIe in the
    step."""
    import torch
    import iesolver
    tol32 = 1e-6
    s.set_matvec(iesolver.MATVEC_FMM)
    for _ in range(max(2, args.warmup)):
        s.factor(tol32, iesolver.FP32)
        s.solve_refined(f, 1e-10, 10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t_ref = 0.0
    for _ in range(args.steps):
        s.factor(tol32, iesolver.FP32)
        t1 = time.perf_counter()
        x, _, its, res = s.solve_refined(f, 1e-10, 10)
        t_ref += time.perf_counter() - t1
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    info = s.matvec_info()
    N = f.size
    s.set_matvec(iesolver.MATVEC_AUTO)
    return dict(form="fp32_learned_factor+fp64_refinement", ms_per_step=round(ms, 4),
                unknowns_per_s=round(N / (ms * 1e-3), 1), factor_id_tol=tol32,
                refine_ms=round(t_ref * 1e3 / args.steps, 4),
                corrections=its, rel_residual=res,
                rel_err_vs_exact=float(np.linalg.norm(x - mu_exact) / np.linalg.norm(mu_exact)),
                residual_operator=("FP64 fast multipole apply, order 48, leaf level %d, %d leaves"
                                   % (info[1], info[3])) if info[0] == iesolver.MATVEC_FMM
                else "FP64 direct sum",
                note="b and x cross PCIe inside the step")


def batch_subrecord(args):
    """The 256-problem batch (learned and plain forms), a few steps each."""
    import torch
    import iesolver
    cfg = CONFIGS["batch256"]
    geoms, meta = G.star_batch(cfg["n_problems"], seed=cfg["seed"], nmin=cfg["nmin"],
                               nmax=cfg["nmax"])
    f = G.batch_data(geoms, seed=cfg["seed"])
    dev = torch.device("cuda:0")
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    s.build_tree(cfg["leaf"])
    fd = torch.from_numpy(f).to(dev)
    xd = torch.empty_like(fd)
    N = s.ndof

    def run(nsteps):
        for _ in range(nsteps):
            s.factor_solve_device(cfg["tol"], iesolver.FP64, fd.data_ptr(), xd.data_ptr())

    out = dict(workload="laplace_batch256", n_problems=cfg["n_problems"], n_unknowns=N,
               problem_sizes=[int(meta["sizes"].min()), int(meta["sizes"].max())])
    steps = max(2, args.steps // 4)
    for form in ("plain", "learned"):
        if form == "learned":
            s.set_rank_model(np.load(os.path.join(ROOT, "ie-solver_amd",
                                                  "rank_model_default.npy")), margin=3)
            s.set_id_method(iesolver.ID_RANDOMIZED, 0)
        else:
            s.set_id_method(iesolver.ID_QRCP)
        run(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out[form] = dict(ms_per_step=round(ms, 4), unknowns_per_s=round(N / (ms * 1e-3), 1),
                         steps=steps)
    y, _ = s.apply(xd.cpu().numpy())
    off = s.dof_offsets
    out["learned"]["worst_backward_error"] = float(max(
        np.linalg.norm(y[off[p]:off[p + 1]] - f[off[p]:off[p + 1]]) /
        np.linalg.norm(f[off[p]:off[p + 1]]) for p in range(cfg["n_problems"])))
    out["speedup_learned_vs_plain"] = round(out["plain"]["ms_per_step"] /
                                            out["learned"]["ms_per_step"], 4)
    s.close()
    return out


def run_batch(args, cfg):
    """The batch workload: every problem's boxes share each stage launch."""
    import torch
    import iesolver
    import oracle

    geoms, meta = G.star_batch(cfg["n_problems"], seed=cfg["seed"], nmin=cfg["nmin"],
                               nmax=cfg["nmax"])
    f = G.batch_data(geoms, seed=cfg["seed"])
    dev = torch.device("cuda:0")
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    s.build_tree(cfg["leaf"])
    s.set_graphs(args.graphs)
    s.set_overlap(not args.no_overlap)
    N = s.ndof
    off = s.dof_offsets
    f_dev = torch.from_numpy(f).to(dev)
    mu_dev = torch.empty_like(f_dev)
    torch.cuda.synchronize()

    def step():
        if args.no_fused:
            s.factor(cfg["tol"], iesolver.FP64)
            s.solve_device(f_dev.data_ptr(), mu_dev.data_ptr(), None, None)
        else:  # the solve's upward sweeps beside the factorization
            s.factor_solve_device(cfg["tol"], iesolver.FP64, f_dev.data_ptr(), mu_dev.data_ptr())

    plain_ms, plain_prof, mu_plain, caps_plain = None, {}, None, None
    if args.form == "plain" or not args.no_plain:
        s.set_id_method(iesolver.ID_QRCP)
        caps_plain = tune(s, args, cfg["tol"], iesolver.FP64)
        plain_ms, plain_prof = time_steps(step, s, args, args.form == "plain" and not args.no_profile)
        mu_plain = mu_dev.cpu().numpy()
    model_info = None
    if args.form == "plain":
        ms_per_step, prof, caps = plain_ms, plain_prof, caps_plain
    else:
        model_info = set_form(s, args)
        caps = tune(s, args, cfg["tol"], iesolver.FP64)
        ms_per_step, prof = time_steps(step, s, args, not args.no_profile)
    mu = mu_dev.cpu().numpy()
    s.set_matvec(iesolver.MATVEC_DIRECT)  # the exact O(N^2) sum for the reference solution
    mu_exact, _, ref_its, ref_res = s.solve_refined(f, 1e-15, 4)
    s.set_matvec(iesolver.MATVEC_AUTO)

    def err(x):
        return float(np.linalg.norm(x - mu_exact) / np.linalg.norm(mu_exact))

    id_stats = None
    if args.form != "plain":
        ns, nf, sf = s.id_stats(with_flops=True)
        id_stats = dict(sketched=ns, fallback=nf, sketch_gflop=round(sf / 1e9, 3))
    info = s.box_info()
    nl, nb, nd, _ = s.tree_info()
    # Comparator of the learned form: fixed 32-row sketches (the randomized
    # form without the rank model), same steps, untimed in the headline.
    fixed32 = None
    if args.form == "learned" and not args.no_fixed32:
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
        f32_ms, _ = time_steps(step, s, args, False)
        mu32 = mu_dev.cpu().numpy()
        ns32, nf32, sf32 = s.id_stats(with_flops=True)
        fixed32 = dict(ms_per_step=round(f32_ms, 4), unknowns_per_s=round(N / (f32_ms * 1e-3), 1),
                       learned_speedup=round(f32_ms / ms_per_step, 4),
                       rel_err_vs_exact=float(np.linalg.norm(mu32 - mu_exact) /
                                              np.linalg.norm(mu_exact)),
                       id_stats=dict(sketched=ns32, fallback=nf32,
                                     sketch_gflop=round(sf32 / 1e9, 3)))
        s.set_id_method(iesolver.ID_RANDOMIZED, 0)  # the learned form again
    stages_ms, roofline = roofline_from(prof, info, id_stats, args.steps, {})

    # FP32 factorization as the preconditioner of FP64 refinement (host
    # vectors in and out: the call uploads b and downloads x).
    ms32, its32, res32, err32 = None, None, None, None
    if not args.no_fp32:
        s.set_id_method(iesolver.ID_QRCP)
        for _ in range(args.warmup):
            s.factor(cfg["fp32_tol"], iesolver.FP32)
            s.solve_refined(f, 1e-11, 10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.factor(cfg["fp32_tol"], iesolver.FP32)
            x32, _, its32, res32 = s.solve_refined(f, 1e-11, 10)
        t1 = time.perf_counter()
        ms32 = (t1 - t0) * 1e3 / args.steps
        err32 = err(x32)

    # Interior field of problem 0 vs its analytic harmonic function.
    tg = G.interior_points(1, 100, radius=0.1)
    ys, q = G.exterior_sources(cfg["seed"])
    u = oracle.evaluate(geoms[0], tg, mu[off[0]:off[1]])
    ue = G.harmonic_data(tg, ys, q)
    interior_err = float(np.linalg.norm(u - ue) / np.linalg.norm(ue))

    cpu, parity = None, None
    if not args.no_cpu_baseline:
        info_cpu = host_cpu()
        threads = info_cpu["usable"] or 1
        k = cfg["cpu_problems"]
        times, parts = [], []
        for p in range(k):
            t_p, mu_p = cpu_baseline(geoms[p], f[off[p]:off[p + 1]], cfg, threads, 1)
            times.append(t_p)
            parts.append(np.linalg.norm(mu[off[p]:off[p + 1]] - mu_p) / np.linalg.norm(mu_p))
        n_s = int(off[k])
        cpu = dict(value=round(n_s / sum(times), 1), unit="unknowns/s", cores=threads, kind="port",
                   sample=f"the batch's first {k} problems ({n_s} unknowns), each factored and "
                          f"solved once after 1 warm-up, OpenMP over boxes, BLAS threads=1",
                   seconds=round(sum(times), 4), cpu_model=info_cpu["model"],
                   cpu_logical=info_cpu["logical"], cpu_physical_cores=info_cpu["physical"])
        parity = float(max(parts))

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
        "config": {"workload": f"laplace_{args.config}_{args.form}_fp64", "pde": "laplace",
                   "n_problems": cfg["n_problems"], "n_unknowns": N,
                   "problem_sizes": [int(meta["sizes"].min()), int(meta["sizes"].max())],
                   "id_tol": cfg["tol"], "leaf_dofs": cfg["leaf"], "levels": nl, "boxes": nb,
                   "form": args.form + "_fp64", "parallelism": "single_gpu"},
        "plain_fp64_ms_per_step": round(plain_ms, 4) if plain_ms else None,
        "speedup_vs_plain_fp64": round(plain_ms / ms_per_step, 4) if plain_ms else None,
        "rel_err_vs_exact": err(mu),
        "randomized_fixed32": fixed32,
        "plain_fp64_rel_err_vs_exact": err(mu_plain) if mu_plain is not None else None,
        "exact": f"FP64 refinement to relative residual {ref_res:.1e} ({ref_its} corrections)",
        "fp32_refined": (dict(ms_per_step=round(ms32, 4), id_tol=cfg["fp32_tol"],
                              corrections=its32, rel_residual=res32, rel_err_vs_exact=err32,
                              unknowns_per_s=round(N / (ms32 * 1e-3), 1),
                              note="FP32 factor + FP64 refinement to relative residual 1e-11; "
                                   "b and x cross PCIe inside the step")
                         if ms32 else None),
        "hip_graph": args.graphs,
        "overlap_id_decouple": not args.no_overlap,
        "fused_factor_solve": not args.no_fused,
        "launch_caps": {"form": caps, "plain_fp64": caps_plain,
                        "order": "assemble, qrcp_id, rowcomp4, rowcomp8, qrcp_id8, qrcp, decouple_split",
                        "autotuned": args.autotune},
        "id_stats": id_stats,
        "rank_model": model_info,
        "rel_err_interior": interior_err,
        "rel_err_vs_cpu_oracle": parity,
        "stages_ms": stages_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    s.close()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="wavy262k", choices=sorted(CONFIGS))
    ap.add_argument("--form", default="learned", choices=["learned", "randomized", "plain"])
    ap.add_argument("--rank-model", default=None,
                    help="load a trained model (.npy from rank_model.py) instead of training")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the factorization from a captured HIP graph (the replay "
                         "runs the two overlapped chains on shared queues, which measured "
                         "slower than launching the two streams directly)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 comparator (profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fused", action="store_true",
                    help="iegpu_factor then iegpu_solve_device instead of iegpu_factor_solve_device")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run each level's decouple on the main stream instead of "
                         "overlapping it with the next level's ID")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch caps (untimed) per form before timing it")
    ap.add_argument("--no-fp32", action="store_true",
                    help="batch workload: skip the FP32 + refinement form")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-fixed32", action="store_true",
                    help="skip timing the fixed 32-row randomized form beside the learned one")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the FP32 + FP64-refinement sub-record")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-problem batch sub-record")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip per-stage HIP events (roofline fields become null)")
    ap.add_argument("--no-update", action="store_true",
                    help="skip the timing of incremental re-factorizations after local "
                         "boundary deformations (iegpu_update_geometry; untimed in the "
                         "headline)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the RS factorization here runs on a single "
              "MI355X (multi-GPU sharding is design-only); rerun with --gpus 1.",
              file=sys.stderr)
        return 2
    cfg = CONFIGS[args.config]
    if cfg["shape"] == "batch":
        return run_batch(args, cfg)
    g, f, ys, q = make_workload(cfg)
    N = g["points"].shape[0]

    import torch
    import iesolver

    dev = torch.device("cuda:0")
    s = iesolver.RSSolver(0)
    s.set_geometry(g, iesolver.PDE_LAPLACE)
    s.build_tree(cfg["leaf"])
    s.set_graphs(args.graphs)
    s.set_overlap(not args.no_overlap)
    f_dev = torch.from_numpy(f).to(dev)
    mu_dev = torch.empty_like(f_dev)
    torch.cuda.synchronize()

    def step():
        if args.no_fused:
            s.factor(cfg["tol"], iesolver.FP64)
            s.solve_device(f_dev.data_ptr(), mu_dev.data_ptr(), None, None)
        else:  # the solve's upward sweeps beside the factorization
            s.factor_solve_device(cfg["tol"], iesolver.FP64, f_dev.data_ptr(), mu_dev.data_ptr())

    def timed(profile):
        return time_steps(step, s, args, profile)

    # Comparator: the plain FP64 form on the same workload.
    plain_ms, plain_prof, mu_plain, caps_plain = None, {}, None, None
    if args.form == "plain" or not args.no_plain:
        s.set_id_method(iesolver.ID_QRCP)
        caps_plain = tune(s, args, cfg["tol"], iesolver.FP64)
        plain_ms, plain_prof = timed(args.form == "plain" and not args.no_profile)
        mu_plain = mu_dev.cpu().numpy()
    model_info = None
    if args.form == "plain":
        ms_per_step, prof, caps = plain_ms, plain_prof, caps_plain
    else:
        model_info = set_form(s, args)
        caps = tune(s, args, cfg["tol"], iesolver.FP64)
        ms_per_step, prof = timed(not args.no_profile)
    mu = mu_dev.cpu().numpy()
    # Accuracy against the exact discrete solution (untimed): this form's
    # factorization as the preconditioner of FP64 iterative refinement with
    # the matrix-free FP64 operator, to a relative residual of 1e-15.
    s.set_matvec(iesolver.MATVEC_DIRECT)  # the exact O(N^2) sum for the reference solution
    mu_exact, _, ref_its, ref_res = s.solve_refined(f, 1e-15, 4)
    s.set_matvec(iesolver.MATVEC_AUTO)
    err_exact = float(np.linalg.norm(mu - mu_exact) / np.linalg.norm(mu_exact))
    err_exact_plain = (float(np.linalg.norm(mu_plain - mu_exact) / np.linalg.norm(mu_exact))
                       if mu_plain is not None else None)
    id_stats = None
    if args.form != "plain":
        ns, nf, sf = s.id_stats(with_flops=True)
        id_stats = dict(sketched=ns, fallback=nf, sketch_gflop=round(sf / 1e9, 3))
    info = s.box_info()
    nl, nb, nd, _ = s.tree_info()
    # Comparator of the learned form: fixed 32-row sketches (the randomized
    # form without the rank model), same steps, untimed in the headline.
    fixed32 = None
    if args.form == "learned" and not args.no_fixed32:
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
        f32_ms, _ = timed(False)
        mu32 = mu_dev.cpu().numpy()
        ns32, nf32, sf32 = s.id_stats(with_flops=True)
        fixed32 = dict(ms_per_step=round(f32_ms, 4), unknowns_per_s=round(N / (f32_ms * 1e-3), 1),
                       learned_speedup=round(f32_ms / ms_per_step, 4),
                       rel_err_vs_exact=float(np.linalg.norm(mu32 - mu_exact) /
                                              np.linalg.norm(mu_exact)),
                       id_stats=dict(sketched=ns32, fallback=nf32,
                                     sketch_gflop=round(sf32 / 1e9, 3)))
        s.set_id_method(iesolver.ID_RANDOMIZED, 0)  # the learned form again

    # Accuracy: interior field vs the analytic harmonic function.
    import oracle
    tg = G.interior_points(1, 100, radius=0.1)
    u = oracle.evaluate(g, tg, mu)
    ue = G.harmonic_data(tg, ys, q)
    interior_err = float(np.linalg.norm(u - ue) / np.linalg.norm(ue))

    stages_ms, roofline = roofline_from(prof, info, id_stats, args.steps,
                                        TRAFFIC if args.form == "learned" and args.config == "wavy262k" else {})

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        # Every core this process may run on (no cap), and the reference's
        # default of 4 threads (ie_solver/io/io.h:14).
        info_cpu = host_cpu()
        threads = info_cpu["usable"] or 1
        t_cpu, mu_cpu = cpu_baseline(g, f, cfg, threads, args.cpu_reps)
        t_cpu4, _ = cpu_baseline(g, f, cfg, 4, max(1, args.cpu_reps // 2))
        cpu = dict(value=round(N / t_cpu, 1), unit="unknowns/s", cores=threads, kind="port",
                   sample=f"full workload ({args.config}, N={N}), median of {args.cpu_reps} "
                          f"factor+solve runs after 1 warm-up, OpenMP over boxes, BLAS threads=1",
                   seconds=round(t_cpu, 4), cpu_model=info_cpu["model"],
                   cpu_logical=info_cpu["logical"], cpu_physical_cores=info_cpu["physical"],
                   value_4_threads=round(N / t_cpu4, 1), seconds_4_threads=round(t_cpu4, 4))
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
        "config": {"workload": f"laplace_{args.config}_{args.form}_fp64", "pde": "laplace",
                   "n_unknowns": N, "id_tol": cfg["tol"], "leaf_dofs": cfg["leaf"],
                   "levels": nl, "boxes": nb, "form": args.form + "_fp64",
                   "parallelism": "single_gpu"},
        "plain_fp64_ms_per_step": round(plain_ms, 4) if plain_ms else None,
        "speedup_vs_plain_fp64": round(plain_ms / ms_per_step, 4) if plain_ms else None,
        "rel_err_vs_exact": err_exact,
        "randomized_fixed32": fixed32,
        "plain_fp64_rel_err_vs_exact": err_exact_plain,
        "exact": f"FP64 refinement to relative residual {ref_res:.1e} ({ref_its} corrections)",
        "rel_err_vs_plain_fp64": (float(np.linalg.norm(mu - mu_plain) / np.linalg.norm(mu_plain))
                                  if mu_plain is not None else None),
        "hip_graph": args.graphs,
        "overlap_id_decouple": not args.no_overlap,
        "fused_factor_solve": not args.no_fused,
        "launch_caps": {"form": caps, "plain_fp64": caps_plain,
                        "order": "assemble, qrcp_id, rowcomp4, rowcomp8, qrcp_id8, qrcp, decouple_split",
                        "autotuned": args.autotune},
        "id_stats": id_stats,
        "rank_model": model_info,
        "rel_err_interior": interior_err,
        "rel_err_vs_cpu_oracle": parity,
        "stages_ms": stages_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not args.no_mixed and cfg["shape"] == "wavy":
        # the headline form's ID method, in FP32 (the model predicts per tol)
        out["mixed_fp32_refined"] = mixed_subrecord(s, f, mu_exact, cfg, args)
    if not args.no_update and cfg["shape"] == "wavy":
        out["geometry_update"] = time_update(s, cfg, args)
    if not args.no_batch and cfg["shape"] == "wavy":
        out["batch256"] = batch_subrecord(args)
    print(json.dumps(out), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
