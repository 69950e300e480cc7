#!/usr/bin/env python3
"""Benchmark: simulated steps/s and tet projections/s of the XPBD solver on one MI355X.

Default workload (the JSON line's `value`): the heterogeneous tissue block
(56^3 cells x 6 tets = 1,053,696 tets, 185,193 vertices, base pinned),
colour-ordered Gauss-Seidel with 20 iterations, dt = 1e-3, a sphere
(r = 0.05 m) pressed down at 0.1 m/s with contacts re-detected every step.
Inputs are synthetic (seeded) and resident in HBM before timing; a step is one
full time step (detect, predict, 20 iterations, pin reset, velocity update)
replayed from a HIP graph.

Extra keys of the same line, each timed the same way with its own roofline:
`beam_gs` and `beam_jacobi` (200x20x20-cell beam, 480,000 tets, 10
iterations), `batch` (256 cubes of 4..24 cells per axis, 10 GS iterations,
a box collider each) and `block_large` (the tissue block at 120^3 cells,
10,368,000 tets, 20 GS iterations, the same sphere).

`roofline` covers the element-projection kernel: algorithmic bytes per solver
iteration (80 B/tet + 28 B/vertex) divided by the summed projection-kernel
time per iteration (HIP events on the solver's stream).  `traffic` is the
counter-measured HBM bytes per iteration of the same kernel shape, read from
profiles/traffic_<workload>_<solver>.json when that file was recorded for the
launch configuration the autotuner picked here (else null).  `cpu_baseline`
times the FP64 oracle (single thread, same loop order as the reference CPU
solver) on a bounded sample of the same workload, `cpu_baseline_parallel` the
same oracle on all host cores.  `learned` compares the learned warm start with
the plain solver on the same workload.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xpbd-tissue-sim_amd"))

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {"cluster": "k_gs_cluster x colours", "quad": "k_gs_cluster_quad x colours",
                "simple": "k_gs_color x colours / k_jacobi_project", "patch": "k_jacobi_patch + k_jacobi_shared"}
METRIC = "sim steps/sec & tet projections/sec at equal residual, learned vs plain, 1 GPU"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="block", choices=["block", "block_large", "beam", "cube", "batch"])
    ap.add_argument("--solver", default=None, choices=["gs", "jacobi"])
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--variant", default="auto", choices=["auto", "simple", "cluster", "patch", "quad"])
    ap.add_argument("--nondeterministic", action="store_true", help="Jacobi: LDS float atomics")
    ap.add_argument("--timing-steps", type=int, default=5)
    ap.add_argument("--extra-steps", type=int, default=100, help="timed steps of each extra workload (0: none)")
    ap.add_argument("--cpu-steps", type=int, default=None, help="oracle steps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--learned-steps", type=int, default=50, help="steps of the learned-vs-plain comparison (0: off)")
    ap.add_argument("--learned-iters", type=int, default=3, help="XPBD iterations of a learned step")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step eagerly (kernel timing mode) instead of replaying HIP graphs")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file (shared with xpbd_sim --tune-cache for counter runs)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "xpbd-tissue-sim_amd", "weights", "warmstart.npz"))
    return ap.parse_args()


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"host_cpu": model or platform.processor() or platform.machine(), "host_nproc": os.cpu_count()}


def build_workload(name, solver, iters, variant="auto", det=1):
    import workloads as W
    import xpbd_amd as X
    kw = dict(variant=variant, deterministic=det)
    if name == "block":
        mesh = W.block(56)
        s = X.Solver(mesh, dt=1e-3, iters=iters or 20, solver=solver or "gs", collision_every=1, **kw)
        kind, pos, quat, prm, vel = W.block_sphere(56)
        s.set_sdf(kind, pos, quat, prm)
        s.set_sdf_velocity(vel)
        cfg = {"workload": "tissue block 56^3 cells x6 tets, heterogeneous E/nu, sphere contact",
               "tets": mesh.nt, "vertices": mesh.nv, "solver": solver or "gs", "iterations": iters or 20,
               "dt": 1e-3}
        return s, [mesh], cfg, (kind, pos, quat, prm, vel)
    if name == "block_large":
        n = 120
        mesh = W.block(n)
        s = X.Solver(mesh, dt=1e-3, iters=iters or 20, solver=solver or "gs", collision_every=1, **kw)
        kind, pos, quat, prm, vel = W.block_sphere(n)
        s.set_sdf(kind, pos, quat, prm)
        s.set_sdf_velocity(vel)
        cfg = {"workload": "tissue block 120^3 cells x6 tets (ten times the 56^3 block), heterogeneous E/nu, "
                           "sphere contact", "tets": mesh.nt, "vertices": mesh.nv, "solver": solver or "gs",
               "iterations": iters or 20, "dt": 1e-3}
        return s, [mesh], cfg, (kind, pos, quat, prm, vel)
    if name == "beam":
        mesh = W.beam(200, 20, 20)
        s = X.Solver(mesh, dt=1e-3, iters=iters or 10, solver=solver or "gs", collision_every=0, **kw)
        cfg = {"workload": "beam 200x20x20 cells x6 tets, E=1e4, nu=0.49", "tets": mesh.nt, "vertices": mesh.nv,
               "solver": solver or "gs", "iterations": iters or 10, "dt": 1e-3}
        return s, [mesh], cfg, None
    if name == "cube":
        mesh = W.cube(10)
        s = X.Solver(mesh, dt=1e-3, iters=iters or 5, solver=solver or "gs", collision_every=0, **kw)
        cfg = {"workload": "cube 10^3 cells x6 tets, E=3e3, nu=0.45", "tets": mesh.nt, "vertices": mesh.nv,
               "solver": solver or "gs", "iterations": iters or 5, "dt": 1e-3}
        return s, [mesh], cfg, None
    meshes, sdfs = W.batch(256)
    s = X.Solver(meshes, dt=1e-3, iters=iters or 10, solver=solver or "gs", collision_every=1, **kw)
    for i, (kind, pos, quat, prm) in enumerate(sdfs):
        s.set_sdf(kind, pos, quat, prm, instance=i)
    cfg = {"workload": "256 cubes of 4..24 cells per axis x6 tets, box contacts", "instances": 256,
           "tets": s.nt, "vertices": s.nv, "solver": solver or "gs", "iterations": iters or 10, "dt": 1e-3}
    return s, meshes, cfg, list(sdfs)


def traffic_for(workload, solver, launch):
    """Counter-measured bytes per iteration, if recorded for this launch configuration."""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{workload}_{solver}.json")
    if not os.path.exists(tfile):
        return None, None, "no counter profile for this workload"
    with open(tfile) as f:
        tj = json.load(f)
    for ent in tj.get("by_launch_config", []):
        if list(ent["launch_config"]) == list(launch):
            return ent["bytes_per_iteration"], tj.get("kernel"), None
    return None, tj.get("kernel"), f"no counter profile recorded for launch config {list(launch)}"


def measure(s, cfg, workload, steps, warmup, timing_steps, no_graph=False, tune_cache=None, variant="auto"):
    """Autotune, time `steps` graph-replayed steps, then HIP-event kernel timing."""
    if variant == "auto":
        s.autotune(tune_cache)
        cfg["autotuned"] = list(s.launch_config())
    cfg["colors"] = s.num_colors()
    T, V, iters = s.nt, s.nv, cfg["iterations"]
    if no_graph:
        s.set_kernel_timing(True)
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    el = time.perf_counter() - t0
    steps_per_s = steps / el
    # projection-kernel timing with HIP events on the solver stream (no graph)
    s.set_kernel_timing(True)
    s.step(timing_steps)
    ms, nit, _ = s.kernel_stats()
    s.set_kernel_timing(False)
    bytes_iter = 80.0 * T + 28.0 * V
    us_per_iter = 1e3 * ms / max(nit, 1)
    achieved = bytes_iter / (us_per_iter * 1e-6) / 1e9
    launch = list(s.launch_config())
    traffic, traffic_kernel, note = traffic_for(workload, cfg["solver"], launch)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": KERNEL_NAMES.get(launch[0], launch[0]) + " (one solver iteration)",
            "launch_config": launch, "us_per_iteration": us_per_iter,
            "algorithmic_bytes_per_iteration": bytes_iter, "iterations_timed": nit, "graph_replay": not no_graph}
    if traffic_kernel:
        roof["traffic_kernel"] = traffic_kernel
    if note:
        roof["traffic_note"] = note
    return {"steps_per_s": steps_per_s, "ms_per_step": 1e3 / steps_per_s,
            "tet_projections_per_s": steps_per_s * iters * T, "steps": steps, "roofline": roof}


def _oracle_step_loop(o, sdf, cfg, nsteps, step0=0):
    """nsteps oracle steps; the kinematic collider is where the device has it at step step0 + n."""
    t0 = time.perf_counter()
    for n in range(step0, step0 + nsteps):
        if sdf is not None:
            kind, pos, quat, prm, vel = sdf
            o.set_sdf(kind, [pos[k] + vel[k] * n * cfg["dt"] for k in range(3)], quat, prm)
        o.step(1, detect=sdf is not None)
    return time.perf_counter() - t0


def cpu_baseline_batch(meshes, cfg, sdfs, nsteps, stride=16):
    """FP64 oracle, one thread, on every stride-th instance of the batch; the
    reference steps objects one after another (Simulation.cpp:246-255), so the
    whole batch costs the sum over instances, extrapolated by element count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    idx = list(range(0, len(meshes), stride))
    el = 0.0
    for i in idx:
        o = orc.Sim(meshes[i], dt=cfg["dt"], iters=cfg["iterations"], mode=cfg["solver"])
        kind, pos, quat, prm = sdfs[i]
        o.set_sdf(kind, pos, quat, prm)
        t0 = time.perf_counter()
        for _ in range(nsteps):
            o.step(1, detect=True)
        el += time.perf_counter() - t0
    t_sample = sum(meshes[i].nt for i in idx)
    t_all = sum(m.nt for m in meshes)
    el_all = el * t_all / t_sample
    out = {"value": nsteps / el_all, "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": f"{nsteps} steps of {len(idx)} of the {len(meshes)} instances (every {stride}th, "
                     f"{t_sample} of {t_all} tets), {cfg['iterations']} {cfg['solver']} iterations, box contacts, "
                     f"one host core, FP64 oracle; whole-batch time extrapolated by element count",
           "tet_proj_per_s": nsteps * cfg["iterations"] * t_sample / el}
    out.update(host_info())
    return out


def _start_text(start):
    return "from rest" if start is None else f"from the GPU run's state after its {start[2]} steps"


def cpu_baseline(meshes, cfg, sdf, nsteps, start=None):
    """FP64 oracle, one thread, on the first mesh of the workload (bounded
    sample), from `start` = (positions, velocities, step count) of the GPU run
    when given (the collider in contact, as in the timed steps), else from rest."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    if isinstance(sdf, list):
        return cpu_baseline_batch(meshes, cfg, sdf, nsteps)
    mesh = meshes[0]
    o = orc.Sim(mesh, dt=cfg["dt"], iters=cfg["iterations"], mode=cfg["solver"])
    if start is not None:
        o.set_state(start[0], start[1])
    el = _oracle_step_loop(o, sdf, cfg, nsteps, 0 if start is None else start[2])
    out = {"value": nsteps / el, "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": f"{nsteps} steps of the same workload {_start_text(start)} ({mesh.nt} tets, {cfg['iterations']} "
                     f"{cfg['solver']} iterations) on one host core, FP64 oracle",
           "tet_proj_per_s": nsteps * cfg["iterations"] * mesh.nt / el}
    out.update(host_info())
    return out


def cpu_baseline_parallel(meshes, cfg, sdf, nsteps, start=None):
    """FP64 oracle on all host cores (OpenMP): colour-ordered Gauss-Seidel over
    clusters of consecutive elements (one hex cell each, 8 colours on the block;
    an order of this build, not the reference's element order), or Jacobi with
    the projections in parallel.  Bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    mesh = meshes[0]
    o = orc.Sim(mesh, dt=cfg["dt"], iters=cfg["iterations"], mode=cfg["solver"])
    ncol = o.color_elements(cluster=True) if cfg["solver"] == "gs" else 0
    o.set_threads(threads)
    if start is not None:
        o.set_state(start[0], start[1])
    el = _oracle_step_loop(o, sdf, cfg, nsteps, 0 if start is None else start[2])
    how = (f"colour-ordered GS, {ncol} colours of element clusters" if cfg["solver"] == "gs"
           else "Jacobi, projections in parallel")
    out = {"value": nsteps / el, "unit": "steps/s", "cores": threads, "kind": "port",
           "sample": f"{nsteps} steps of the same workload {_start_text(start)} ({mesh.nt} tets, {cfg['iterations']} "
                     f"{cfg['solver']} iterations), FP64 oracle, {how}, {threads} OpenMP threads",
           "tet_proj_per_s": nsteps * cfg["iterations"] * mesh.nt / el}
    out.update(host_info())
    return out


MLP_FLOP_PER_VERTEX = 2 * (24 * 128 + 128 * 128 + 128 * 3)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def learned_vs_plain(a, net_path):
    """Learned warm start against the plain solver on the same workload and
    steps.  The plain run reads its end-of-solve residual back every step (the
    reference for the fallback test, ratio 1.1); the learned runs take
    a.learned_iters iterations per step:
      with_fallback  network correction, multiplier start (the default);
      no_correction  the same step with a network that outputs zero;
      positions_start  network correction, multipliers from zero;
      positions_start_no_fallback  the same, never falling back (its residual
                       is reported, not judged);
      consistent_start  network correction, multipliers -C / alpha~ there and
                       positions that keep the momentum balance;
      consistent_start_no_fallback  the same, never falling back;
    and plain_same_iterations is the plain solver with a.learned_iters
    iterations (no network, no fallback), for scale.  Each run reports its
    trajectory error against the plain run after the last step, as
    max |x - x_plain| / max |x_plain| and in units of the plain run's largest
    displacement from rest.  primary_residual repeats the plain run and four
    of the others untimed with the primary residual (the momentum balance
    (x - xt) / w - sum grad C^T lambda) recorded every step, and reports its
    norm against the plain run's, step by step.  The learned runs replay captured steps; the
    network's inference time comes from a separate run of a few steps in
    kernel timing mode."""
    import numpy as np
    import train_warmstart as TW
    import xpbd_amd as X
    det = 0 if a.nondeterministic else 1

    def make(iters):
        """A solver of the workload on the launch shape the other legs use (shared tuning cache)."""
        s_, meshes_, cfg_, _ = build_workload(a.workload, a.solver, iters, a.variant, det)
        if a.variant == "auto":
            s_.autotune(a.tune_cache)
        return s_, meshes_, cfg_

    s_plain, meshes, cfg = make(a.iters)
    launch_plain = list(s_plain.launch_config())
    rest = np.concatenate([m.x0 for m in meshes])
    s_plain.set_residual_policy(True)
    r_plain, t_plain = [], 0.0
    for _ in range(a.learned_steps):
        t0 = time.perf_counter()
        s_plain.step(1)
        r_plain.append(s_plain.last_step_residual())
        t_plain += time.perf_counter() - t0
    x_plain = s_plain.positions()
    s_plain.close()
    disp = float(np.max(np.abs(x_plain - rest)))
    net = TW.load(net_path)
    zero = X.Network(np.zeros(X.MLP_PARAMS, np.float32))
    zero.set_scaling(np.zeros(24), np.ones(24), 1.0)
    out = {"iterations": a.learned_iters, "plain_iterations": cfg["iterations"], "fallback_ratio": 1.1,
           "launch_config": launch_plain,
           "steps": a.learned_steps, "weights": os.path.relpath(net_path, ROOT),
           "plain_steps_per_s_with_residual_readback": a.learned_steps / t_plain,
           "plain_residual_median": float(np.median(r_plain))}
    for mode, network, start, judged in (("with_fallback", net, "multipliers", True),
                                         ("no_correction", zero, "multipliers", True),
                                         ("positions_start", net, "positions", True),
                                         ("positions_start_no_fallback", net, "positions", False),
                                         ("consistent_start", net, "consistent", True),
                                         ("consistent_start_no_fallback", net, "consistent", False)):
        s, _, _ = make(a.iters)
        s.attach(network)
        s.set_learned_start(start)
        ratios, fb, t, diverged = [], 0, 0.0, None
        for n in range(a.learned_steps):
            t0 = time.perf_counter()
            f, r = s.step_learned(a.learned_iters, r_plain[n] if judged else 0.0, 1.1)
            t += time.perf_counter() - t0
            if not np.isfinite(r):   # the run left the finite numbers: report where, stop
                diverged = n
                break
            fb += int(f)
            ratios.append(r / max(r_plain[n], 1e-300))
        taken = len(ratios)
        ratios = np.array(ratios) if taken else np.full(1, np.nan)
        out[mode] = {"start": start, "steps_per_s": max(taken, 1) / t, "fallback_rate": fb / max(taken, 1),
                     "residual_ratio_median": float(np.median(ratios)), "residual_ratio_max": float(ratios.max())}
        if diverged is None:
            dx = float(np.max(np.abs(s.positions() - x_plain)))
            out[mode]["rel_pos_err_vs_plain_end"] = dx / float(np.max(np.abs(x_plain)))
            out[mode]["traj_err_vs_plain_end_over_displacement"] = dx / disp
        else:
            out[mode]["diverged_at_step"] = diverged
        if mode == "with_fallback":
            # inference time: a few more steps, launched eagerly with HIP events around the network
            s.set_kernel_timing(True)
            for _ in range(5):
                s.step_learned(a.learned_iters, 0.0, 1.1)
            st = s.learned_stats()
            ms_inf = st["mlp_ms"] / max(st["mlp_calls"], 1)
            flop = MLP_FLOP_PER_VERTEX * s.nv
            out["mlp_ms_per_inference"] = ms_inf
            out["mlp_tflops"] = flop / (ms_inf * 1e-3) / 1e12
            out["mlp_mfma_util"] = out["mlp_tflops"] / MFMA_BF16_PEAK_TFLOPS
        s.close()
    # the plain solver at the learned step's iteration count, for scale
    s, _, _ = make(a.learned_iters)
    s.set_residual_policy(True)
    ratios, t = [], 0.0
    for n in range(a.learned_steps):
        t0 = time.perf_counter()
        s.step(1)
        ratios.append(s.last_step_residual() / max(r_plain[n], 1e-300))
        t += time.perf_counter() - t0
    dx = float(np.max(np.abs(s.positions() - x_plain)))
    out["plain_same_iterations"] = {"iterations": a.learned_iters, "steps_per_s": a.learned_steps / t,
                                    "residual_ratio_median": float(np.median(ratios)),
                                    "residual_ratio_max": float(np.max(ratios)),
                                    "rel_pos_err_vs_plain_end": dx / float(np.max(np.abs(x_plain))),
                                    "traj_err_vs_plain_end_over_displacement": dx / disp}
    s.close()
    out["learned_over_plain_at_equal_residual"] = (out["with_fallback"]["steps_per_s"] /
                                                   out["plain_steps_per_s_with_residual_readback"])

    def primary_norms(iters, network=None, start=None, judged=False):
        """Primary residual norm (momentum balance) of every step of a run,
        recorded on the device in an untimed repeat of the run."""
        s_, _, _ = make(iters)
        s_.set_primary_residual(True)
        if network is not None:
            s_.attach(network)
            s_.set_learned_start(start)
        p = []
        for n in range(a.learned_steps):
            if network is None:
                s_.step(1)
            else:
                s_.step_learned(a.learned_iters, r_plain[n] if judged else 0.0, 1.1)
            v = s_.last_step_primary_residual()
            if not np.isfinite(v):
                break
            p.append(v)
        s_.close()
        return np.array(p)

    p_plain = primary_norms(a.iters)
    pr = {"plain_median": float(np.median(p_plain))}
    for key, args in (("with_fallback", (a.iters, net, "multipliers", True)),
                      ("positions_start_no_fallback", (a.iters, net, "positions", False)),
                      ("consistent_start_no_fallback", (a.iters, net, "consistent", False)),
                      ("plain_same_iterations", (a.learned_iters,))):
        q = primary_norms(*args)
        q = q / np.maximum(p_plain[:len(q)], 1e-300)
        if len(q) < len(p_plain):
            pr[key + "_diverged_at_step"] = len(q)
        pr[key + "_over_plain_median"] = float(np.median(q)) if len(q) else None
        pr[key + "_over_plain_max"] = float(np.max(q)) if len(q) else None
    out["primary_residual"] = pr
    return out


def _finite(o):
    """The record with every non-finite number replaced by null (strict JSON)."""
    if isinstance(o, dict):
        return {k: _finite(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_finite(v) for v in o]
    if isinstance(o, float) and not math.isfinite(o):
        return None
    return o


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md but not built)"}))
        sys.exit(2)
    import xpbd_amd as X  # noqa: F401  (fails loudly if the HIP library is missing)
    if a.tune_cache is None:
        # one tuning per workload shape, shared by every leg of this run
        import atexit
        import shutil
        import tempfile
        tmp = tempfile.mkdtemp(prefix="xpbd_tune_")
        atexit.register(shutil.rmtree, tmp, True)
        a.tune_cache = os.path.join(tmp, "tune.txt")
    s, meshes, cfg, sdf = build_workload(a.workload, a.solver, a.iters, a.variant, 0 if a.nondeterministic else 1)
    cfg["variant"] = a.variant
    cfg["deterministic"] = not a.nondeterministic
    m = measure(s, cfg, a.workload, a.steps, a.warmup, a.timing_steps, a.no_graph, a.tune_cache, a.variant)
    res = s.residual()
    # the CPU baselines resume from where the GPU run ended (one object; the batch samples from rest)
    start = None if a.workload == "batch" else (s.positions(), s.velocities(), a.warmup + a.steps + a.timing_steps)
    s.close()
    out = {
        "metric": METRIC,
        "value": m["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "tet_projections_per_s": m["tet_projections_per_s"],
        "residual_after_run": res,
        "roofline": m["roofline"],
    }
    if a.extra_steps > 0 and a.workload == "block":
        for key, wl, solver in (("beam_gs", "beam", "gs"), ("beam_jacobi", "beam", "jacobi"), ("batch", "batch", "gs"),
                                ("block_large", "block_large", "gs")):
            se, _, ce, _ = build_workload(wl, solver, None, "auto", 1)
            me = measure(se, ce, wl, a.extra_steps, 10, a.timing_steps, tune_cache=a.tune_cache)
            se.close()
            me["config"] = ce
            out[key] = me
    if a.learned_steps > 0 and a.workload != "batch":
        out["learned"] = learned_vs_plain(a, a.weights)
    if not a.no_cpu:
        n = a.cpu_steps if a.cpu_steps is not None else {"block": 10, "batch": 40}.get(a.workload, 20)
        out["cpu_baseline"] = cpu_baseline(meshes, cfg, sdf, n, start)
        if a.workload != "batch":
            # second reported baseline: the same oracle on all host cores
            out["cpu_baseline_parallel"] = cpu_baseline_parallel(meshes, cfg, sdf, 4 * n, start)
    print(json.dumps(_finite(out), allow_nan=False))


if __name__ == "__main__":
    main()
