"""Benchmark of the CVO registration hot path on one MI355X.

Headline workload (the JSON line's `value`): plain two-view CVO on the
60k-point street scene (5 color features, 19-class semantics, 5 deg / 1 m
motion), registrations/s with both clouds resident in HBM before the timed
region (source sort, grid build, the whole iteration loop and the pose
readback are inside it).

Further legs, reported as nested objects of the same line:
  ba_window   IRLS multi-frame bundle adjustment, 12 frames of 5k-40k points,
              every frame pair whose exact cos angle exceeds 0.05 is an edge;
              s/iter = median over outer iterations 2..end
  multistart  global registration of two 20k-point clouds (uniform random
              rotation, 2 m) from 96 starts on a super-Fibonacci SO(3) grid in
              one batched launch sequence; starts/s and success
  learned     the learned proposer: inference of 256 pairs x 2048 points
              (MFMA roofline of its forward), and global registration from its
              8 proposals (grid fallback) against the plain 96-start grid
  knn         the 60k street registration with is_using_kdtree = 1 (K nearest)
  adaptive    the adaptive-lengthscale registration of the 5k planes pair
  aniso       the non-isotropic association (3x3 kernel) of the 60k street pair
              in camera axes

Each leg carries the roofline of its dominant kernel (HIP events on the
library's stream) and the CPU oracle timed on the host cores on a bounded
sample of the same workload.
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

METRIC = "registrations/s @60k pts + BA-window s/iter; batched/learned vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
# HBM bytes per association launch on the 60k street workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes scaled by a 256 MiB copy probe
# (profiles/assoc_street60k_pmc_traffic.md)
ASSOC_TRAFFIC_BYTES = 25.78e6
# HBM bytes per BA pair-pass launch on the 12-frame window, the same way
# (profiles/ba_window12_pmc_traffic.md)
BA_PAIR_TRAFFIC_BYTES = 912.5e6
# per-launch work of the association kernel on the 60k street registration
# with its autotuned launch shape, from rocprofv3 counters (SQ_INSTS_VALU_TRANS_F32
# x 64 lanes; SQ_INSTS_VALU_FLOPS_FP32 x 64): the transcendental and FP32
# roofs are checked against the live launch time
ASSOC_EXP_PER_LAUNCH = 3.96e6
ASSOC_FP32_FLOP_PER_LAUNCH = 2.92e8
EXP_PEAK_PER_S = 256 * 16 * 2.4e9   # quarter-rate v_exp_f32: 16 lanes per CU per clock
FP32_PEAK_TFLOPS = 157.3
CONFIGS = os.path.join(ROOT, "rkhs_ba_amd", "configs")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pose_error(T, T_gt):
    E = np.asarray(T, np.float64) @ np.asarray(T_gt, np.float64)
    ang = float(np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1))))
    return ang, float(np.linalg.norm(E[:3, 3]))


def cpu_threads():
    return min(16, os.cpu_count() or 1)


def cpu_host(threads):
    """The host the CPU baselines ran on: threads used, CPU model, logical
    CPUs, the oracle's compiler and flags."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    from oracle import pyoracle
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"threads": threads, "cpu_model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "compiler": pyoracle.build_info()}


def rot_trans_diff(A, B):
    """Rotation angle (rad) and translation distance (m) between two poses."""
    E = np.linalg.inv(np.asarray(A, np.float64)) @ np.asarray(B, np.float64)
    return (float(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1))), float(np.linalg.norm(E[:3, 3])))


# ---------------------------------------------------------------------------
# two-view (headline)
# ---------------------------------------------------------------------------

def cpu_twoview(params, src, tgt, T_gt, trace, threads, budget_s=12.0):
    """Times the oracle's iteration (FP32-mirror algorithm, OpenMP over rows,
    grid radius search) at each distinct lengthscale the GPU run visited and
    weights by the GPU's iteration count at that lengthscale: an estimate of
    one full registration on the host cores."""
    from oracle import pyoracle
    ells = {}
    for r in trace:
        key = round(float(r.ell), 7)
        e = ells.setdefault(key, [0, 0])
        e[0] += 1
        e[1] = max(e[1], int(r.num_neighbors))
    total_est = 0.0
    sampled = 0
    t_start = time.time()
    per_iter = {}
    order = sorted(ells.items(), reverse=True)
    probe = pyoracle.time_iterations(params, src, tgt, T_gt, order[0][0], max(order[0][1][1], 1), 1, threads)
    n_it = int(min(20, max(2, budget_s / max(1e-3, probe * len(order)))))
    for ell, (count, K) in order:
        t = pyoracle.time_iterations(params, src, tgt, T_gt, ell, max(K, 1), n_it, threads)
        per_iter[ell] = t
        sampled += n_it
        total_est += count * t
    return total_est, sampled, per_iter, time.time() - t_start


def tune(cvo, ds, dt, args, params=None):
    """Association launch shape for this workload: the autotuner's choice
    (cached per workload key), or the library default with --tune default."""
    if args.tuning:
        v = [float(x) for x in args.tuning.split(",")]
        small = (int(v[3]), v[4], v[5]) if len(v) >= 6 else (int(v[0]), 0.0, v[2])
        cvo.set_tuning(int(v[0]), int(v[1]), v[2], lanes_small_ell=small[0], split_ell=small[1],
                       cell_scale_small_ell=small[2])
        return cvo.get_tuning()
    if args.tune == "default":
        return cvo.get_tuning()
    if params is None:
        return cvo.autotune(ds, dt, np.eye(4), cache="default")
    # a workload with its own lengthscale range (BA): tune on a context that
    # carries that range, then apply the choice
    from rkhs_ba_amd import CvoGPU
    tc = CvoGPU(params=params)
    try:
        d1, d2 = tc.upload(ds), tc.upload(dt)
        t = tc.autotune(d1, d2, np.eye(4), cache="default")
    finally:
        tc.close()
    cvo.apply_tuning(t)
    return t


def leg_twoview(args):
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "street_60k.yaml"))
    log(f"[bench] two-view: street scene n={args.points} seed={args.seed}")
    src, tgt, T_gt = synth.street_pair(args.seed, n=args.points)
    cvo = CvoGPU(params=params)
    dsrc, dtgt = cvo.upload(src), cvo.upload(tgt)
    I = np.eye(4, dtype=np.float32)
    tuning = tune(cvo, dsrc, dtgt, args)
    for w in range(args.warmup):
        log(f"[bench] two-view warmup {w}")
        cvo.align(dsrc, dtgt, I)
    log(f"[bench] two-view timed: {args.steps} registrations")
    t0 = time.perf_counter()
    results = [cvo.align(dsrc, dtgt, I) for _ in range(args.steps)]
    wall = time.perf_counter() - t0
    ret, T, dev_s, iters, _ = results[-1]
    rot_e, tr_e = pose_error(T, T_gt)
    log(f"[bench] two-view {wall / args.steps * 1e3:.1f} ms/registration, {iters} iters, "
        f"err {rot_e:.4f} deg {tr_e * 100:.2f} cm")
    out = {"value": args.steps / wall, "ms_per_step": wall / args.steps * 1e3, "iterations": iters,
           "device_loop_ms": dev_s * 1e3, "pose_error": {"deg": rot_e, "m": tr_e}, "params": params,
           "tuning": tuning}
    if args.profile_only:
        return out
    # host to host: host clouds in (upload, source sort, grid build, the loop)
    # and the pose out, the median over at least five registrations
    n_h2h = max(5, args.steps)
    log(f"[bench] two-view host-to-host: {n_h2h} registrations from host memory")
    h2h = []
    for _ in range(n_h2h):
        t0 = time.perf_counter()
        cvo.align(src, tgt, I)
        h2h.append(time.perf_counter() - t0)
    med = float(np.median(h2h))
    out["host_to_host"] = {"ms_median": med * 1e3, "registrations_per_s": 1.0 / med, "runs": n_h2h,
                           "ms_min": float(np.min(h2h)) * 1e3, "ms_max": float(np.max(h2h)) * 1e3,
                           "note": "host clouds in, pose out: upload, sort, grid build, the loop, readback"}
    log("[bench] two-view profiled registration")
    cvo.reset_kernel_stats()
    cvo.set_profiling(True)
    cvo.set_trace(params.MAX_ITER)
    cvo.align(dsrc, dtgt, I)
    cvo.set_profiling(False)
    trace = cvo.get_trace()
    st = cvo.kernel_stats()
    n_assoc = max(1, st["launches"]["assoc"])
    assoc_ms = st["ms"]["assoc"] / n_assoc
    assoc_bytes = st["bytes"]["assoc"] / n_assoc
    achieved = assoc_bytes / (assoc_ms * 1e-3) / 1e9
    exp_rate = ASSOC_EXP_PER_LAUNCH / (assoc_ms * 1e-3)
    flop_rate = ASSOC_FP32_FLOP_PER_LAUNCH / (assoc_ms * 1e-3) / 1e12
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": args.traffic,
                       "kernel": "assoc_kernel", "avg_launch_ms": assoc_ms,
                       "algorithmic_bytes_per_launch": assoc_bytes, "launches": n_assoc,
                       "transcendental": {"achieved_per_s": exp_rate, "peak_per_s": EXP_PEAK_PER_S,
                                          "frac": exp_rate / EXP_PEAK_PER_S, "per_launch": ASSOC_EXP_PER_LAUNCH},
                       "fp32": {"achieved_tflops": flop_rate, "peak_tflops": FP32_PEAK_TFLOPS,
                                "frac": flop_rate / FP32_PEAK_TFLOPS, "per_launch": ASSOC_FP32_FLOP_PER_LAUNCH},
                       "timing": "HIP events at the kernel's own start and end (hipExtLaunchKernelGGL), "
                                 "the library's stream, every launch of one registration",
                       "note": "latency-bound (dependent gathers); per-launch counts from rocprofv3 counters"}
    out["kernels_ms_per_launch"] = {
        "assoc": assoc_ms,
        "overflow": st["ms"]["overflow"] / max(1, st["launches"]["overflow"]),
        "step": st["ms"]["step"] / max(1, st["launches"]["step"])}
    if not args.no_cpu and args.cpu_sample_only:
        threads = cpu_threads()
        log(f"[bench] two-view cpu baseline on {threads} threads")
        est_s, sampled, per_iter, spent = cpu_twoview(params, src, tgt, T_gt, trace, threads)
        out["cpu_baseline"] = {
            "value": 1.0 / est_s if est_s > 0 else None, "unit": "registrations/s", "cores": threads, "kind": "port",
            "sample": (f"oracle FP32-mirror iteration (grid radius search, OpenMP) timed {sampled} times over the "
                       f"{len(per_iter)} lengthscales of the GPU run, weighted by the GPU's iterations per "
                       f"lengthscale ({len(trace)} iterations); {spent:.1f} s of CPU work"),
            "seconds_per_registration": est_s, "host": cpu_host(threads)}
    if not args.no_cpu and not args.cpu_sample_only:
        from oracle import pyoracle
        log(f"[bench] two-view: one whole oracle registration in FP64 (the library's default pose arithmetic), "
            f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}")
        t0 = time.perf_counter()
        To, it_o, ret_o, _ = pyoracle.align(params, src, tgt, I, precision=pyoracle.POLICY)
        spent = time.perf_counter() - t0
        dr, dt_ = rot_trans_diff(T, To)
        eo = pose_error(To, T_gt)
        log(f"[bench] two-view oracle: {spent:.1f} s, {it_o} iterations; GPU vs oracle {dr:.2e} rad {dt_:.2e} m")
        omp = int(os.environ.get("OMP_NUM_THREADS") or (os.cpu_count() or 1))
        out["cpu_full_registration"] = {
            "seconds": spent, "registrations_per_s": 1.0 / spent, "iterations": it_o, "ret": ret_o,
            "precision": "FP64 pose arithmetic and sums (the library's default)",
            "gpu_vs_oracle": {"rad": dr, "m": dt_}, "oracle_pose_error": {"deg": eo[0], "m": eo[1]},
            "host": cpu_host(omp)}
        out["cpu_baseline"] = {
            "value": 1.0 / spent, "unit": "registrations/s", "cores": omp, "kind": "port",
            "sample": (f"one whole oracle registration of the same pair from the same start ({it_o} iterations; "
                       f"FP64 pose and sums over the FP32 mirror's association, grid radius search, OpenMP), "
                       f"{spent:.1f} s of CPU work"),
            "seconds_per_registration": spent, "host": cpu_host(omp)}
    cvo.close()
    return out


# ---------------------------------------------------------------------------
# IRLS multi-frame BA window
# ---------------------------------------------------------------------------

def window_edges(cvo, clouds, poses, ell, thr=0.05):
    """Frame pairs whose exact cos angle (function_angle, approx=False) at the
    initial poses exceeds thr."""
    edges = []
    for i in range(len(clouds)):
        for j in range(i + 1, len(clouds)):
            T = (np.linalg.inv(poses[j]) @ poses[i]).astype(np.float32)
            if cvo.function_angle(clouds[i], clouds[j], T, ell, is_approximate=False) > thr:
                edges.append((i, j))
    return edges


def rel_pose_errors(P, gts):
    out = []
    for f in range(1, len(P)):
        E = np.linalg.inv(np.linalg.inv(gts[0]) @ gts[f]) @ (np.linalg.inv(P[0]) @ P[f])
        out.append((float(np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1)))),
                    float(np.linalg.norm(E[:3, 3]))))
    return np.array(out)


def leg_ba(args):
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "ba_window.yaml"))
    log("[bench] BA window: 12 frames, 5k-40k points")
    clouds, gts, inits = synth.window_frames(args.seed, n_frames=12, n_min=5000, n_max=40000)
    cvo = CvoGPU(params=params)
    edges = window_edges(cvo, clouds, inits, params.multiframe_ell_init)
    hold = [1] + [0] * 11
    # tuned on the window's largest edge at the BA lengthscale range
    bp = params.copy()
    bp.ell_init, bp.ell_min = params.multiframe_ell_init, params.multiframe_ell_min
    bp.nearest_neighbors_max = params.multiframe_num_neighbors
    i0, j0 = max(edges, key=lambda e: clouds[e[0]].n * clouds[e[1]].n)
    tuning = tune(cvo, clouds[i0], clouds[j0], args, params=bp)
    dev = [cvo.upload(c) for c in clouds]
    log(f"[bench] BA window: {len(edges)} edges, {sum(c.n for c in clouds)} points; warmup")
    cvo.align_multiframe(dev, inits, hold, edges)
    log("[bench] BA window timed run")
    t0 = time.perf_counter()
    r = cvo.align_multiframe(dev, inits, hold, edges)
    wall = time.perf_counter() - t0
    per = r["per_iter_seconds"][:r["iters"]]
    s_iter = float(np.median(per[1:])) if len(per) > 1 else float(per[0])
    e0, e1 = rel_pose_errors(inits, gts), rel_pose_errors(r["poses"], gts)
    log(f"[bench] BA window {r['iters']} iterations, {s_iter * 1e3:.2f} ms/iter (median 2..end), "
        f"first {per[0] * 1e3:.2f} ms; max error {e0[:, 0].max():.3f}->{e1[:, 0].max():.3f} deg")
    out = {"s_per_iter": s_iter, "first_iter_s": float(per[0]), "iterations": r["iters"], "wall_s": wall,
           "frames": len(clouds), "edges": len(edges), "points": int(sum(c.n for c in clouds)),
           "max_error_before": {"deg": float(e0[:, 0].max()), "m": float(e0[:, 1].max())},
           "max_error_after": {"deg": float(e1[:, 0].max()), "m": float(e1[:, 1].max())}, "tuning": tuning}
    if args.profile_only:
        return out
    log("[bench] BA window profiled run")
    cvo.reset_kernel_stats()
    cvo.set_profiling(True)
    cvo.align_multiframe(dev, inits, hold, edges)
    cvo.set_profiling(False)
    st = cvo.kernel_stats()
    n_pair = max(1, st["launches"]["ba_pair"])
    pair_ms = st["ms"]["ba_pair"] / n_pair
    pair_bytes = st["bytes"]["ba_pair"] / n_pair
    assoc_ms = st["ms"]["assoc"] / max(1, st["launches"]["assoc"])
    assoc_bytes = st["bytes"]["assoc"] / max(1, st["launches"]["assoc"])
    out["roofline"] = {"bound": "hbm", "achieved": pair_bytes / (pair_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": pair_bytes / (pair_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "kernel": "ba_pair_kernel", "avg_launch_ms": pair_ms,
                       "algorithmic_bytes_per_launch": pair_bytes, "launches": n_pair,
                       "traffic": BA_PAIR_TRAFFIC_BYTES}
    out["assoc_roofline"] = {"achieved": assoc_bytes / (assoc_ms * 1e-3) / 1e9, "avg_launch_ms": assoc_ms,
                             "algorithmic_bytes_per_launch": assoc_bytes}
    out["kernels_ms_per_iteration"] = {
        k: st["ms"][k] / max(1, st["launches"]["assoc"]) for k in ("assoc", "overflow", "ba_pair", "ba_solve")}
    if not args.no_cpu:
        from oracle import pyoracle
        threads = cpu_threads()
        p1 = read_params_yaml(os.path.join(CONFIGS, "ba_window.yaml"))
        p1.multiframe_max_iters = 1
        log(f"[bench] BA window cpu baseline: one outer iteration on {threads} threads")
        t0 = time.perf_counter()
        P1, it, secs = pyoracle.align_multiframe(p1, clouds, inits, hold, edges, threads=threads)
        spent = time.perf_counter() - t0
        # the GPU's poses after the same single outer iteration
        cvo.write_params(p1)
        g1 = cvo.align_multiframe(dev, inits, hold, edges)
        out["gpu_vs_fp64_oracle_1iter"] = {
            "max_abs_pose_entry": float(np.abs(g1["poses"][:, :3, :4] - P1[:, :3, :4]).max())}
        out["cpu_baseline"] = {"value": float(secs[0]), "unit": "s/iter", "cores": threads, "kind": "port",
                               "host": cpu_host(threads),
                               "sample": f"oracle IRLS outer iteration 1 (association of all {len(edges)} edges "
                                         f"+ up to {p1.multiframe_iterations_per_solve} LM candidates), "
                                         f"{spent:.1f} s of CPU work",
                               "gpu_first_iter_s": float(per[0]),
                               "speedup_first_iter": float(secs[0]) / float(per[0])}
        log(f"[bench] BA window: the reference's CPU edge states, one outer iteration on {threads} threads")
        t0 = time.perf_counter()
        _, it, secs = pyoracle.align_multiframe_cpu_states(p1, clouds, inits, hold, edges, threads=threads)
        spent = time.perf_counter() - t0
        out["cpu_states_baseline"] = {
            "value": float(secs[0]), "unit": "s/iter", "cores": threads, "kind": "port", "host": cpu_host(threads),
            "sample": ("oracle IRLS outer iteration 1 with BinaryStateCPU edges (IRLS_State_CPU.cpp: fixed-ell "
                       f"radius search of frame 2 over frame 1, weight ck k, no cap), {spent:.1f} s of CPU work")}
    cvo.close()
    return out


# ---------------------------------------------------------------------------
# the CPU-runnable configuration: 5k/5k planes and spheres, 10 deg / 0.3 m
# ---------------------------------------------------------------------------

def leg_c1(args):
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "planes_spheres_5k.yaml"))
    log("[bench] c1: planes and spheres 5k/5k, 10 deg / 0.3 m")
    src, tgt, T_gt = synth.planes_spheres_pair(args.seed)
    cvo = CvoGPU(params=params)
    ds, dt = cvo.upload(src), cvo.upload(tgt)
    I = np.eye(4, dtype=np.float32)
    tuning = tune(cvo, ds, dt, args)
    cvo.align(ds, dt, I)
    t0 = time.perf_counter()
    runs = [cvo.align(ds, dt, I) for _ in range(args.steps)]
    wall = time.perf_counter() - t0
    ret, T, dev_s, iters, _ = runs[-1]
    e = pose_error(T, T_gt)
    cvo.close()
    log(f"[bench] c1 {wall / args.steps * 1e3:.1f} ms/registration, {iters} iters, err {e[0]:.4f} deg {e[1] * 100:.2f} cm")
    out = {"registrations_per_s": args.steps / wall, "ms_per_registration": wall / args.steps * 1e3,
           "iterations": iters, "pose_error": {"deg": e[0], "m": e[1]}, "tuning": tuning}
    if args.no_cpu or args.profile_only:
        return out
    from oracle import pyoracle
    threads = cpu_threads()
    log(f"[bench] c1 cpu_ref: oracle FP32-mirror registration on {threads} threads")
    t0 = time.perf_counter()
    To, it_o, ret_o, _ = pyoracle.align(params, src, tgt, I)
    spent = time.perf_counter() - t0
    eo = pose_error(To, T_gt)
    out["cpu_baseline"] = {"value": 1.0 / spent, "unit": "registrations/s", "cores": threads, "kind": "port",
                           "sample": f"one whole oracle FP32-mirror registration ({it_o} iterations)",
                           "iterations": it_o, "pose_error": {"deg": eo[0], "m": eo[1]}, "host": cpu_host(threads)}
    # the GPU pose against the FP64 oracle (the library's default arithmetic)
    Tp, it_p, _, _ = pyoracle.align(params, src, tgt, I, precision=pyoracle.POLICY)
    dr, dt_ = rot_trans_diff(T, Tp)
    out["gpu_vs_fp64_oracle"] = {"rad": dr, "m": dt_, "oracle_iterations": it_p, "gpu_iterations": iters}
    log(f"[bench] c1 cpu_cvo_legacy: the reference's CPU CVO on {threads} threads")
    lp = pyoracle.legacy_params(params)
    Tl, it_l, ret_l, secs_l, _ = pyoracle.cvo_legacy_align(lp, src, tgt, I, threads=threads)
    el = pose_error(Tl, T_gt)
    per_it = secs_l / max(it_l, 1)
    out["cpu_cvo_legacy"] = {
        "value": 1.0 / secs_l if secs_l > 0 else None, "unit": "registrations/s", "cores": threads, "kind": "port",
        "host": cpu_host(threads),
        "sample": ("one whole run of the restated cvo::cvo::align (Cvo.cpp:885-1089) with the same YAML: "
                   "unnormalised flow, float steps; it stops on its own criteria"),
        "iterations": it_l, "seconds_per_iteration": per_it, "pose_error": {"deg": el[0], "m": el[1]},
        "seconds_at_gpu_iterations": per_it * iters}
    return out


# ---------------------------------------------------------------------------
# batched multi-start global registration
# ---------------------------------------------------------------------------

def leg_multistart(args):
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "multistart_20k.yaml"))
    params.MAX_ITER = 1000
    log("[bench] multi-start: 20k/20k, uniform rotation + 2 m, 96 starts")
    src, tgt, T_gt = synth.planes_spheres_pair(100 + args.seed, 20000, 20000, uniform_rot=True, trans=2.0)
    cx = src.xyz.mean(0).astype(np.float64)
    cy = tgt.xyz.mean(0).astype(np.float64)
    Ts = np.array([synth.rigid(R, cy - R @ cx) for R in synth.so3_grid(96)], np.float32)
    cvo = CvoGPU(params=params)
    ds, dt = cvo.upload(src), cvo.upload(tgt)
    tuning = tune(cvo, ds, dt, args)
    cvo.align_multistart(ds, dt, Ts, 0.3, mode=0)
    t0 = time.perf_counter()
    r = cvo.align_multistart(ds, dt, Ts, 0.3, mode=0)
    wall = time.perf_counter() - t0
    e = pose_error(r["T_best"], T_gt)
    per_ok = sum(1 for T in r["per_start_T"] if pose_error(T, T_gt)[0] < np.degrees(0.01) and
                 pose_error(T, T_gt)[1] < 0.01)
    success = bool(e[0] < np.degrees(0.01) and e[1] < 0.01)
    log(f"[bench] multi-start {wall:.2f} s, best err {e[0]:.4f} deg {e[1] * 100:.2f} cm, "
        f"{per_ok}/96 starts converge")
    out = {"seconds": wall, "starts_per_s": 96 / wall, "success": success,
           "best_error": {"deg": e[0], "m": e[1]}, "starts_converged": per_ok, "tuning": tuning,
           "iterations_total": int(r["per_start_iters"].sum())}
    if not args.no_cpu and not args.profile_only:
        from oracle import pyoracle
        threads = cpu_threads()
        t1 = pyoracle.time_iterations(params, src, tgt, np.linalg.inv(T_gt).astype(np.float32), params.ell_init,
                                      params.nearest_neighbors_max, 3, threads)
        est = t1 * out["iterations_total"]
        out["cpu_baseline"] = {"value": 96 / est, "unit": "starts/s", "cores": threads, "kind": "port",
                               "host": cpu_host(threads),
                               "sample": "oracle iteration at ell_init timed 3 times x the GPU run's total "
                                         "iterations over all 96 starts",
                               "seconds_estimate": est}
        # the winning start registered by the FP64 oracle from the same start
        best = int(np.argmax(r["per_start_cos"]))
        log(f"[bench] multi-start: FP64 oracle from start {best}")
        t1 = time.perf_counter()
        To, it_o, _, _ = pyoracle.align(params, src, tgt, Ts[best], precision=pyoracle.FP64)
        dr, dt_ = rot_trans_diff(r["per_start_T"][best], To)
        out["gpu_vs_fp64_oracle_best_start"] = {"rad": dr, "m": dt_, "oracle_iterations": int(it_o),
                                                "gpu_iterations": int(r["per_start_iters"][best]),
                                                "oracle_seconds": time.perf_counter() - t1}
    cvo.close()
    return out


# ---------------------------------------------------------------------------
# learned proposals + CVO vs the plain grid
# ---------------------------------------------------------------------------

MFMA_BF16_PEAK_TFLOPS = 2500.0


def leg_learned(args):
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    from rkhs_ba_amd import proposer as P
    params = read_params_yaml(os.path.join(CONFIGS, "multistart_20k.yaml"))
    params.MAX_ITER = 1000
    cvo = CvoGPU(params=params)
    prop = P.PoseProposer.load(cvo)
    log("[bench] learned: inference batch of 256 pairs x 2048 points")
    a, b, t = P.make_batch(range(8_000_000, 8_000_256))
    from rkhs_ba_amd import DeviceArray
    a_dev, b_dev = DeviceArray(a), DeviceArray(b)
    prop.propose(a_dev, b_dev)
    reps = 3
    # batches resident in device memory (the timed form), then from host arrays
    t0 = time.perf_counter()
    for _ in range(reps):
        prop.propose(a_dev, b_dev)
    wall = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        prop.propose(a, b)
    wall_host = (time.perf_counter() - t0) / reps
    cvo.reset_kernel_stats()
    cvo.set_profiling(True)
    prop.propose(a_dev, b_dev)
    cvo.set_profiling(False)
    st = cvo.kernel_stats()
    fwd_ms = st["ms"]["mlp"] / max(1, st["launches"]["mlp"])
    flops = st["bytes"]["mlp"] / max(1, st["launches"]["mlp"])
    achieved = flops / (fwd_ms * 1e-3) / 1e12
    out = {"inference_pairs_per_s": 256 / wall, "inference_batch_ms": wall * 1e3, "forward_ms": fwd_ms,
           "inference_batch_ms_host_inputs": wall_host * 1e3,
           "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "kernel": "encoder_fused_kernel + head gemm_bf16_kernel",
                        "flops_per_launch": flops, "traffic": None}}
    n = args.learned_pairs
    ok_g = ok_l = 0
    t_g = t_l = 0.0
    fallbacks = 0
    for s in range(n):
        src, tgt, Tgt = synth.planes_spheres_pair(7100 + s, 20000, 20000, uniform_rot=True, trans=2.0)
        g = P.register_global(cvo, src, tgt, None)
        l = P.register_global(cvo, src, tgt, prop, min_cos=P.DEFAULT_MIN_COS)
        eg, el = pose_error(g["T_best"], Tgt), pose_error(l["T_best"], Tgt)
        ok_g += eg[0] <= np.degrees(0.01) and eg[1] <= 0.01
        ok_l += el[0] <= np.degrees(0.01) and el[1] <= 0.01
        t_g += g["seconds"]
        t_l += l["seconds"]
        fallbacks += int(l["fallback"])
    log(f"[bench] learned: grid {ok_g}/{n} in {t_g / n:.3f} s/pair, learned {ok_l}/{n} in {t_l / n:.3f} s/pair")
    out.update({"pairs": n, "grid_success": ok_g / n, "learned_success": ok_l / n, "grid_s_per_pair": t_g / n,
                "learned_s_per_pair": t_l / n, "throughput_vs_grid": t_g / max(t_l, 1e-9), "fallbacks": fallbacks})
    if not args.no_cpu and not args.profile_only:
        from oracle import mlp_oracle
        w = prop.get_weights().astype(np.float64)
        t0 = time.perf_counter()
        mlp_oracle.forward(w, a[:32].astype(np.float64), b[:32].astype(np.float64), quantize=False)
        cpu_s = (time.perf_counter() - t0) * 256 / 32
        out["cpu_baseline"] = {"value": 256 / cpu_s, "unit": "pairs/s (inference)", "cores": cpu_threads(),
                               "kind": "port", "sample": "numpy FP64 forward of the same network on 32 of the "
                                                         "256 pairs, scaled to the batch",
                               "host": cpu_host(cpu_threads())}
    prop.close()
    cvo.close()
    return out


# ---------------------------------------------------------------------------
# the paths beside the plain association: K nearest, adaptive lengthscale,
# non-isotropic kernel
# ---------------------------------------------------------------------------

def cpu_by_lengthscale(trace, time_at, budget_s=12.0):
    """Estimate of a whole registration on the host: the GPU run's iterations
    grouped by (lengthscale, row cap); time_at(ell, K) times one oracle pass
    for a spread of the groups within the budget, every other group takes the
    time of the sampled group nearest in log(ell)."""
    groups = {}
    for r in trace:
        key = (round(float(r.ell), 7), max(int(r.num_neighbors), 1))
        groups[key] = groups.get(key, 0) + 1
    keys = sorted(groups, reverse=True)
    t0 = time.time()
    sampled = {}
    first = time_at(*keys[0])
    sampled[keys[0]] = first
    n_fit = int(max(1, min(len(keys), budget_s / max(first, 1e-3))))
    for i in np.linspace(0, len(keys) - 1, n_fit).round().astype(int)[1:]:
        k = keys[i]
        if k not in sampled:
            sampled[k] = time_at(*k)
    total = 0.0
    for k, n in groups.items():
        near = min(sampled, key=lambda s: abs(np.log(s[0]) - np.log(k[0])))
        total += n * sampled[near]
    return total, len(sampled), len(keys), time.time() - t0


def leg_knn(args):
    """The 60k street registration with is_using_kdtree = 1: each row's
    K nearest targets inside the fixed-lengthscale gate (K = 512 on the first
    iteration, then ceil(stored pairs / N))."""
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "street_60k.yaml"))
    params.is_using_kdtree = 1
    log(f"[bench] knn: street scene n={args.points}, is_using_kdtree 1")
    src, tgt, T_gt = synth.street_pair(args.seed, n=args.points)
    cvo = CvoGPU(params=params)
    ds, dt = cvo.upload(src), cvo.upload(tgt)
    I = np.eye(4, dtype=np.float32)
    tuning = tune(cvo, ds, dt, args)
    cvo.align(ds, dt, I)
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    runs = [cvo.align(ds, dt, I) for _ in range(steps)]
    wall = (time.perf_counter() - t0) / steps
    ret, T, dev_s, iters, _ = runs[-1]
    e = pose_error(T, T_gt)
    log(f"[bench] knn {wall * 1e3:.1f} ms/registration, {iters} iters, err {e[0]:.4f} deg {e[1] * 100:.2f} cm")
    out = {"registrations_per_s": 1.0 / wall, "ms_per_registration": wall * 1e3, "iterations": iters,
           "pose_error": {"deg": e[0], "m": e[1]}, "tuning": tuning}
    if args.profile_only:
        cvo.close()
        return out
    cvo.reset_kernel_stats()
    cvo.set_profiling(True)
    cvo.set_trace(params.MAX_ITER)
    cvo.align(ds, dt, I)
    cvo.set_profiling(False)
    trace = cvo.get_trace()
    cvo.set_trace(0)
    st = cvo.kernel_stats()
    n = max(1, st["launches"]["assoc"])
    ms, by = st["ms"]["assoc"] / n, st["bytes"]["assoc"] / n
    out["roofline"] = {"bound": "hbm", "achieved": by / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": "knn_assoc_kernel",
                       "avg_launch_ms": ms, "algorithmic_bytes_per_launch": by, "launches": n, "traffic": None,
                       "note": "latency-bound: the ball is scanned 2 + select passes times per row"}
    out["kernels_ms_per_launch"] = {k: st["ms"][k] / max(1, st["launches"][k]) for k in ("assoc", "overflow", "step")}
    if not args.no_cpu:
        from oracle import pyoracle
        threads = cpu_threads()
        log(f"[bench] knn cpu baseline ({threads} threads)")

        def time_at(ell, K):
            t = time.perf_counter()
            pyoracle.assoc_knn(params, src, tgt, T_gt, ell, K, precision=pyoracle.MIRROR)
            return time.perf_counter() - t

        est, n_s, n_g, spent = cpu_by_lengthscale(trace, time_at)
        out["cpu_baseline"] = {"value": 1.0 / est, "unit": "registrations/s", "cores": threads, "kind": "port",
                               "host": cpu_host(threads),
                               "sample": (f"oracle K-nearest association (FP32 mirror, OpenMP) timed at {n_s} of the "
                                          f"{n_g} (lengthscale, K) settings of the GPU run, weighted by its "
                                          f"iterations; step terms not counted; {spent:.1f} s of CPU work"),
                               "seconds_per_registration": est}
    cvo.close()
    return out


def leg_adaptive(args):
    """AdaptiveCvoGPU::align on the 5k/5k planes and spheres pair: three
    association problems per iteration (x-y, x-x, y-y) and the lengthscale
    gradient."""
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "planes_spheres_5k.yaml"))
    log("[bench] adaptive: planes and spheres 5k/5k, 10 deg / 0.3 m")
    src, tgt, T_gt = synth.planes_spheres_pair(args.seed)
    cvo = CvoGPU(params=params)
    ds, dt = cvo.upload(src), cvo.upload(tgt)
    I = np.eye(4, dtype=np.float32)
    cvo.align_adaptive(ds, dt, I)
    steps = max(1, min(args.steps, 5))
    t0 = time.perf_counter()
    runs = [cvo.align_adaptive(ds, dt, I) for _ in range(steps)]
    wall = (time.perf_counter() - t0) / steps
    ret, T, dev_s, iters = runs[-1]
    e = pose_error(T, T_gt)
    log(f"[bench] adaptive {wall * 1e3:.1f} ms/registration, {iters} iters, err {e[0]:.4f} deg {e[1] * 100:.2f} cm")
    out = {"registrations_per_s": 1.0 / wall, "ms_per_registration": wall * 1e3, "iterations": iters,
           "pose_error": {"deg": e[0], "m": e[1]}}
    if args.profile_only:
        cvo.close()
        return out
    cvo.reset_kernel_stats()
    cvo.set_profiling(True)
    cvo.align_adaptive(ds, dt, I)
    cvo.set_profiling(False)
    st = cvo.kernel_stats()
    n = max(1, st["launches"]["assoc"])
    ms, by = st["ms"]["assoc"] / n, st["bytes"]["assoc"] / n
    out["roofline"] = {"bound": "hbm", "achieved": by / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": "assoc_kernel (three problems)",
                       "avg_launch_ms": ms, "algorithmic_bytes_per_launch": by, "launches": n, "traffic": None,
                       "note": "bytes count every point three times and the registration problem's pairs only"}
    out["kernels_ms_per_launch"] = {k: st["ms"][k] / max(1, st["launches"][k]) for k in ("assoc", "overflow", "step")}
    if not args.no_cpu:
        from oracle import pyoracle
        log(f"[bench] adaptive cpu baseline: one oracle registration")
        t0 = time.perf_counter()
        To, it_o, ret_o, _ = pyoracle.align_adaptive(params, src, tgt, I, precision=pyoracle.MIRROR)
        spent = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / spent, "unit": "registrations/s", "cores": cpu_threads(),
                               "kind": "port", "sample": f"one whole oracle registration (FP32 mirror, OpenMP "
                                                         f"associations), {it_o} iterations",
                               "iterations": it_o, "host": cpu_host(cpu_threads())}
        # the GPU pose against the FP64 oracle (the library's default arithmetic)
        Tp, it_p, _, _ = pyoracle.align_adaptive(params, src, tgt, I, precision=pyoracle.POLICY)
        dr, dt_ = rot_trans_diff(T, Tp)
        out["gpu_vs_fp64_oracle"] = {"rad": dr, "m": dt_, "oracle_iterations": it_p, "gpu_iterations": iters}
    cvo.close()
    return out


ANISO_KERNEL = np.diag([1.0, 0.2, 3.0]).astype(np.float32)
# world (x forward, y left, z up) -> camera (x right, y down, z forward)
CAMERA_FROM_WORLD = np.array([[0, -1, 0, 0], [0, 0, -1, 0], [1, 0, 0, 0], [0, 0, 0, 1]], np.float64)


def leg_aniso(args):
    """compute_association_gpu with a 3x3 kernel on the 60k street pair (19
    classes, colour), kernel diag(1, 0.2, 3) (the third axis along the
    viewing ray), the clouds in camera axes: the reference builds each
    point's frame from its ray around the camera's z axis (degenerate for
    rays in the z = 0 plane), so the workload is the scene as a forward
    camera frames it.  Each row's ellipsoid is searched on the target grid."""
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "street_60k.yaml"))
    log(f"[bench] non-isotropic association: street scene n={args.points}, camera axes")
    src, tgt, T_gt = synth.street_pair(args.seed, n=args.points)
    C4 = CAMERA_FROM_WORLD
    src, tgt = src.transformed(C4), tgt.transformed(C4)
    T_gt = (C4 @ np.asarray(T_gt, np.float64) @ C4.T).astype(np.float32)
    cvo = CvoGPU(params=params)
    a = cvo.compute_association_gpu(src, tgt, T_gt, ANISO_KERNEL)
    steps = max(1, min(args.steps, 5))
    t0 = time.perf_counter()
    for _ in range(steps):
        a = cvo.compute_association_gpu(src, tgt, T_gt, ANISO_KERNEL)
    wall = (time.perf_counter() - t0) / steps
    cvo.reset_kernel_stats()
    cvo.set_profiling(True)
    for _ in range(steps):
        cvo.compute_association_gpu(src, tgt, T_gt, ANISO_KERNEL)
    cvo.set_profiling(False)
    st = cvo.kernel_stats()
    n = max(1, st["launches"]["aniso"])
    ms, by = st["ms"]["aniso"] / n, st["bytes"]["aniso"] / n
    log(f"[bench] non-isotropic {wall * 1e3:.2f} ms per call, kernel {ms:.3f} ms, {a.nnz} pairs")
    out = {"calls_per_s": 1.0 / wall, "ms_per_call": wall * 1e3, "kernel_ms": ms, "pairs": int(a.nnz),
           "note": "a call uploads both clouds, builds the target grid, associates, compacts and copies the "
                   "CSR back; kernel_ms is the association kernel alone (HIP events)",
           "roofline": {"bound": "hbm", "achieved": by / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": "aniso_assoc_kernel",
                        "avg_launch_ms": ms, "algorithmic_bytes_per_launch": by, "launches": n,
                        "traffic": None}}
    if not args.no_cpu and not args.profile_only:
        from oracle import pyoracle
        log("[bench] non-isotropic cpu baseline: one whole oracle association")
        t0 = time.perf_counter()
        pyoracle.assoc_nonisotropic(params, src, tgt, T_gt, ANISO_KERNEL)
        spent = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / spent, "unit": "associations/s", "cores": cpu_threads(),
                               "host": cpu_host(cpu_threads()),
                               "kind": "port", "sample": ("one whole oracle association (every target per row, "
                                                          f"OpenMP), {spent:.1f} s of CPU work"),
                               "seconds_per_association": spent}
    cvo.close()
    return out


def assoc_probe(args):
    """Association passes of the 60k street pair at two lengthscales for each
    (lanes, XCD order) at cell scale 1, six passes each, for per-dispatch
    counters; prints one line per setting."""
    from rkhs_ba_amd import CvoGPU, read_params_yaml, synth
    params = read_params_yaml(os.path.join(CONFIGS, "street_60k.yaml"))
    src, tgt, T = synth.street_pair(0, n=args.points)
    cvo = CvoGPU(params=params)
    ds, dt = cvo.upload(src), cvo.upload(tgt)
    Ti = np.linalg.inv(T)
    for ell in (0.5, 0.05):
        for lanes, xcd in ((16, 0), (16, 1), (8, 0), (8, 1)):
            cvo.set_tuning(lanes, xcd, 1.0)
            for _ in range(6):
                cvo.inner_product_gpu(ds, dt, Ti, ell)
            log(f"[probe] ell {ell} lanes {lanes} xcd_remap {xcd}")
    cvo.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=60000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--legs", default="twoview,c1,ba,multistart,learned,knn,adaptive,aniso")
    ap.add_argument("--learned-pairs", type=int, default=4, help="pairs for learned vs grid registration")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-sample-only", action="store_true",
                    help="estimate the two-view CPU baseline from timed oracle iterations at each lengthscale "
                         "instead of timing one whole oracle registration of the 60k street pair (~27 s)")
    ap.add_argument("--tune", choices=("auto", "default"), default="auto",
                    help="association launch shape: autotuned (cached per workload) or the library default")
    ap.add_argument("--tuning", default="",
                    help="fixed launch shape LANES,XCD,SCALE[,LANES_SMALL,SPLIT_ELL,SCALE_SMALL] (profiling runs, "
                         "where a profiler would distort the autotuner's timings)")
    ap.add_argument("--profile-only", action="store_true", help="timed loops only (for rocprofv3)")
    ap.add_argument("--traffic", type=float, default=ASSOC_TRAFFIC_BYTES,
                    help="HBM bytes per association launch from rocprofv3 counters")
    ap.add_argument("--assoc-probe", action="store_true",
                    help="only association passes of the 60k street pair at fixed launch shapes (for counters)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device 5 times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.assoc_probe:
        assoc_probe(args)
        return
    if args.copy_probe:
        import torch
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 256 << 20, "copies": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    legs = set(args.legs.split(","))
    tv = leg_twoview(args)
    params = tv.pop("params")
    extra = {}
    if "c1" in legs:
        extra["c1_planes5k"] = leg_c1(args)
    if "ba" in legs:
        extra["ba_window"] = leg_ba(args)
    if "multistart" in legs:
        extra["multistart"] = leg_multistart(args)
    if "learned" in legs:
        extra["learned"] = leg_learned(args)
    if "knn" in legs:
        extra["knn_street60k"] = leg_knn(args)
    if "adaptive" in legs:
        extra["adaptive_planes5k"] = leg_adaptive(args)
    if "aniso" in legs:
        extra["nonisotropic_street60k"] = leg_aniso(args)
    cpu = tv.get("cpu_baseline")
    out = {
        "metric": METRIC,
        "value": tv["value"],
        "unit": "registrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tv["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulation and pose)",
        "data": "synthetic (seeded street scene, clouds resident in HBM)",
        "config": {"workload": "two-view CVO, 60k/60k street scene, 5 features + 19 classes, 5deg/1m",
                   "points": args.points, "iterations": tv["iterations"], "max_iter": params.MAX_ITER,
                   "parallelism": "single GPU"},
        "iterations_per_registration": tv["iterations"],
        "device_loop_ms": tv["device_loop_ms"],
        "pose_error": tv["pose_error"],
        "roofline": tv.get("roofline"),
        "tuning": tv.get("tuning"),
        "kernels_ms_per_launch": tv.get("kernels_ms_per_launch"),
        "host_to_host": tv.get("host_to_host"),
        "cpu_baseline": cpu,
        "cpu_full_registration": tv.get("cpu_full_registration"),
        "speedup_vs_cpu": tv["value"] / cpu["value"] if cpu and cpu["value"] else None,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
