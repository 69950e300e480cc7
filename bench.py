#!/usr/bin/env python3
"""Benchmark: simulated steps/s and tet projections/s of the XPBD solver on one MI355X.

Default workload: the heterogeneous tissue block (56^3 cells x 6 tets =
1,053,696 tets, 185,193 vertices, base pinned), colour-ordered Gauss-Seidel
with 20 iterations, dt = 1e-3, a sphere (r = 0.05 m) pressed down at 0.1 m/s
with contacts re-detected every step.  Inputs are synthetic (seeded) and
resident in HBM before timing; a step is one full time step (detect, predict,
20 iterations, pin reset, velocity update) replayed from a HIP graph.

Prints one JSON line.  `roofline` covers the element-projection kernel
(algorithmic bytes per solver iteration, 80 B/tet + 28 B/vertex, divided by
the summed projection-kernel time per iteration, HIP events on the solver's
stream); `cpu_baseline` times the FP64 oracle (single thread, same loop order
as the reference CPU solver) on a bounded sample of the same workload, and
`cpu_baseline_parallel` the same oracle on all host cores (OpenMP colour-ordered
Gauss-Seidel over element clusters, or parallel Jacobi projection).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xpbd-tissue-sim_amd"))

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {"cluster": "k_gs_cluster x colours", "quad": "k_gs_cluster_quad x colours",
                "simple": "k_gs_color x colours / k_jacobi_project", "patch": "k_jacobi_patch + k_jacobi_shared"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="block", choices=["block", "beam", "cube", "batch"])
    ap.add_argument("--solver", default=None, choices=["gs", "jacobi"])
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--variant", default="auto", choices=["auto", "simple", "cluster", "patch", "quad"])
    ap.add_argument("--nondeterministic", action="store_true", help="Jacobi: LDS float atomics")
    ap.add_argument("--timing-steps", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=None, help="oracle steps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--learned-steps", type=int, default=50, help="steps of the learned-vs-plain comparison (0: off)")
    ap.add_argument("--learned-iters", type=int, default=3)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step eagerly (kernel timing mode) instead of replaying HIP graphs")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="MiB moved once by a plain device copy (scale for rocprof HBM byte counters)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "xpbd-tissue-sim_amd", "weights", "warmstart.npz"))
    return ap.parse_args()


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
    return {"value": nsteps / el_all, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{nsteps} steps of {len(idx)} of the {len(meshes)} instances (every {stride}th, "
                      f"{t_sample} of {t_all} tets), {cfg['iterations']} {cfg['solver']} iterations, box contacts, "
                      f"one host core, FP64 oracle; whole-batch time extrapolated by element count",
            "tet_proj_per_s": nsteps * cfg["iterations"] * t_sample / el,
            "host": platform.processor() or platform.machine(), "host_nproc": os.cpu_count()}


def cpu_baseline(meshes, cfg, sdf, nsteps):
    """FP64 oracle, one thread, on the first mesh of the workload (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    if isinstance(sdf, list):
        return cpu_baseline_batch(meshes, cfg, sdf, nsteps)
    mesh = meshes[0]
    o = orc.Sim(mesh, dt=cfg["dt"], iters=cfg["iterations"], mode=cfg["solver"])
    if sdf is not None:
        kind, pos, quat, prm, vel = sdf
        o.set_sdf(kind, pos, quat, prm)
    t0 = time.perf_counter()
    for n in range(nsteps):
        if sdf is not None:
            p = [pos[k] + vel[k] * n * cfg["dt"] for k in range(3)]
            o.set_sdf(kind, p, quat, prm)
        o.step(1, detect=sdf is not None)
    el = time.perf_counter() - t0
    return {"value": nsteps / el, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{nsteps} steps of the same workload ({mesh.nt} tets, {cfg['iterations']} "
                      f"{cfg['solver']} iterations) on one host core, FP64 oracle",
            "tet_proj_per_s": nsteps * cfg["iterations"] * mesh.nt / el,
            "host": platform.processor() or platform.machine(), "host_nproc": os.cpu_count()}


def cpu_baseline_parallel(meshes, cfg, sdf, nsteps):
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
    if sdf is not None:
        kind, pos, quat, prm, vel = sdf
        o.set_sdf(kind, pos, quat, prm)
    t0 = time.perf_counter()
    for n in range(nsteps):
        if sdf is not None:
            p = [pos[k] + vel[k] * n * cfg["dt"] for k in range(3)]
            o.set_sdf(kind, p, quat, prm)
        o.step(1, detect=sdf is not None)
    el = time.perf_counter() - t0
    how = (f"colour-ordered GS, {ncol} colours of element clusters" if cfg["solver"] == "gs"
           else "Jacobi, projections in parallel")
    return {"value": nsteps / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} steps of the same workload ({mesh.nt} tets, {cfg['iterations']} "
                      f"{cfg['solver']} iterations), FP64 oracle, {how}, {threads} OpenMP threads",
            "tet_proj_per_s": nsteps * cfg["iterations"] * mesh.nt / el,
            "host": platform.processor() or platform.machine(), "host_nproc": os.cpu_count()}


MLP_FLOP_PER_VERTEX = 2 * (24 * 128 + 128 * 128 + 128 * 3)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def learned_vs_plain(a, net_path):
    """Learned warm start (k iterations + network, fallback at 1.1x the plain
    residual) against the plain solver on the same workload and steps."""
    import numpy as np
    import train_warmstart as TW
    s_plain, _, cfg, _ = build_workload(a.workload, a.solver, a.iters, a.variant, 0 if a.nondeterministic else 1)
    s_plain.set_residual_policy(True)
    r_plain, t_plain = [], 0.0
    for _ in range(a.learned_steps):
        t0 = time.perf_counter()
        s_plain.step(1)
        r_plain.append(s_plain.last_step_residual())
        t_plain += time.perf_counter() - t0
    x_plain = s_plain.positions()
    trained_here = not os.path.exists(net_path)
    if trained_here:
        # no saved weights: train a small network on the device now
        x, y = TW.build_dataset(2, 60)
        import xpbd_amd as X
        net = X.Network()
        net.fit_scaling(x, y)
        net.train(x, y, batch=8192, epochs=10, lr=2e-3, seed=7)
    else:
        net = TW.load(net_path)
    out = {"iterations": a.learned_iters, "plain_iterations": cfg["iterations"], "fallback_ratio": 1.1,
           "steps": a.learned_steps, "weights": "trained in this run" if trained_here else os.path.relpath(net_path, ROOT),
           "plain_steps_per_s_with_residual_readback": a.learned_steps / t_plain}
    for mode, ref in (("with_fallback", r_plain), ("no_fallback", [0.0] * len(r_plain))):
        s, _, _, _ = build_workload(a.workload, a.solver, a.iters, a.variant, 0 if a.nondeterministic else 1)
        s.attach(net)
        s.set_kernel_timing(True)
        ratios, fb, t = [], 0, 0.0
        for n in range(a.learned_steps):
            t0 = time.perf_counter()
            f, r = s.step_learned(a.learned_iters, ref[n], 1.1)
            t += time.perf_counter() - t0
            fb += int(f)
            ratios.append(r / max(r_plain[n], 1e-300))
        st = s.learned_stats()
        x = s.positions()
        ratios = np.array(ratios)
        ms_inf = st["mlp_ms"] / max(st["mlp_calls"], 1)
        out[mode] = {"steps_per_s": a.learned_steps / t, "fallback_rate": fb / a.learned_steps,
                     "residual_ratio_median": float(np.median(ratios)), "residual_ratio_max": float(ratios.max()),
                     "rel_pos_err_vs_plain_end": float(np.max(np.abs(x - x_plain)) / np.max(np.abs(x_plain))),
                     "mlp_ms_per_inference": ms_inf}
        if mode == "no_fallback":
            flop = MLP_FLOP_PER_VERTEX * s.nv
            out["mlp_tflops"] = flop / (ms_inf * 1e-3) / 1e12
            out["mlp_mfma_util"] = out["mlp_tflops"] / MFMA_BF16_PEAK_TFLOPS
        s.close()
    out["learned_over_plain_at_equal_residual"] = (out["with_fallback"]["steps_per_s"] /
                                                   out["plain_steps_per_s_with_residual_readback"])
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md but not built)"}))
        sys.exit(2)
    import xpbd_amd as X  # noqa: F401  (fails loudly if the HIP library is missing)
    if a.copy_calibration > 0:
        # a copy of known size, so that FETCH_SIZE / WRITE_SIZE per dispatch can be scaled
        import torch
        n = int(a.copy_calibration * (1 << 20)) // 4
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": 4 * n}), file=sys.stderr)
    s, meshes, cfg, sdf = build_workload(a.workload, a.solver, a.iters, a.variant, 0 if a.nondeterministic else 1)
    if a.variant == "auto":
        # time the variants that share this handle's layout and keep the fastest
        s.autotune()
        cfg["autotuned"] = list(s.launch_config())
    cfg["variant"] = a.variant
    cfg["deterministic"] = not a.nondeterministic
    cfg["colors"] = s.num_colors()
    T, V = s.nt, s.nv
    iters = cfg["iterations"]

    if a.no_graph:
        s.set_kernel_timing(True)
    s.step(a.warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(a.steps)
    s.sync()
    el = time.perf_counter() - t0
    steps_per_s = a.steps / el
    tet_proj_per_s = steps_per_s * iters * T

    # projection-kernel timing with HIP events on the solver stream (no graph)
    s.set_kernel_timing(True)
    s.step(a.timing_steps)
    ms, nit, nlaunch = s.kernel_stats()
    s.set_kernel_timing(False)
    bytes_iter = 80.0 * T + 28.0 * V
    us_per_iter = 1e3 * ms / max(nit, 1)
    achieved = bytes_iter / (us_per_iter * 1e-6) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{a.workload}_{cfg['solver']}.json")
    traffic_kernel = None
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic, traffic_kernel = tj.get("bytes_per_iteration"), tj.get("kernel")

    res = s.residual()
    out = {
        "metric": "sim steps/sec & tet projections/sec at equal residual, learned vs plain, 1 GPU",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 / steps_per_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "tet_projections_per_s": tet_proj_per_s,
        "residual_after_run": res,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": KERNEL_NAMES.get(s.variant()[0], s.variant()[0]) + " (one solver iteration)",
                     "us_per_iteration": us_per_iter, "algorithmic_bytes_per_iteration": bytes_iter,
                     "iterations_timed": nit, "graph_replay": not a.no_graph,
                     "traffic_kernel": traffic_kernel},
    }
    if a.learned_steps > 0 and a.workload != "batch":
        s.close()
        out["learned"] = learned_vs_plain(a, a.weights)
    if not a.no_cpu:
        n = a.cpu_steps if a.cpu_steps is not None else {"block": 10, "batch": 40}.get(a.workload, 20)
        out["cpu_baseline"] = cpu_baseline(meshes, cfg, sdf, n)
        if a.workload != "batch":
            # second reported baseline: the same oracle on all host cores
            out["cpu_baseline_parallel"] = cpu_baseline_parallel(meshes, cfg, sdf, 4 * n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
