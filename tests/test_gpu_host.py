"""The C++ host layer on the GPU: the xpbd_sim driver (Simulation /
XPBDMeshObject / XPBDGPUSolver over the C-ABI) against the FP64 oracle.

Tolerance: 1e-5 relative vertex-position error after 100 steps (beam) and 50
steps (block with sphere contacts on the reference's collision cadence)."""
import os
import subprocess

import numpy as np
import pytest

import orc
import xpbd_amd as X
from helpers import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "xpbd-tissue-sim_amd", "host")
SIM = os.path.join(HOST, "xpbd_sim")


def drive(cfg, tmp_path, steps, *extra):
    pos, order, rank = tmp_path / "x.bin", tmp_path / "order.bin", tmp_path / "rank.bin"
    r = subprocess.run([SIM, os.path.join(HOST, "configs", cfg), "--steps", str(steps), "--dump-positions", str(pos),
                        "--dump-order", str(order), "--dump-face-rank", str(rank), *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return (np.fromfile(pos, np.float64).reshape(-1, 3), np.fromfile(order, np.int32), np.fromfile(rank, np.int32),
            r.stdout)


def test_beam_simulation_matches_oracle(tmp_path):
    x, order, _, log = drive("beam.yaml", tmp_path, 100, "--beam")
    xr, t = X.grid_mesh(4, 24, 4, 0.01)
    xr = xr - xr.mean(0)  # uniform grid: volume centroid = vertex mean; position (0, 0, 0)
    pinned = (np.abs(xr[:, 1] - xr[:, 1].min()) < 1e-12).astype(np.uint8)
    mesh = X.Mesh(x0=xr, tets=t, E=np.full(len(t), 1e4), nu=np.full(len(t), 0.49), rho=np.full(len(t), 1000.0),
                  pinned=pinned)
    o = orc.Sim(mesh, dt=1e-3, g=9.81, iters=10, mode="gs")
    o.set_order(order)
    o.step(100)
    err = rel_err(x, o.x())
    lines = log.strip().splitlines()
    print(f"beam via xpbd_sim: rel err {err:.2e}; log header {lines[0]!r}, last {lines[-1]!r}")
    assert err <= 1e-5
    assert len(lines) == 101 and lines[0].split()[1:] == ["Beam_Z-deflection[m]", "Beam_Constraint-Residual"]
    defl = [float(l.split()[1]) for l in lines[1:]]
    assert defl[-1] > 0 and np.all(np.isfinite(defl))


def test_block_with_sphere_on_reference_collision_cadence(tmp_path):
    steps = 50
    x, order, rank, _ = drive("block_sphere.yaml", tmp_path, steps)
    xr, t = X.grid_mesh(12, 12, 12, 0.01)
    pinned = (np.abs(xr[:, 2]) < 1e-12).astype(np.uint8)
    mesh = X.Mesh(x0=xr, tets=t, E=np.full(len(t), 3e3), nu=np.full(len(t), 0.45), rho=np.full(len(t), 1000.0),
                  pinned=pinned, faces=X.surface_faces(len(xr), t))
    o = orc.Sim(mesh, dt=1e-3, g=9.81, iters=20, mode="gs")
    o.set_order(order)
    o.set_contact_order(rank)
    # Simulation.cpp:197-276 cadence, same double arithmetic
    time, last, between, detections = 0.0, 0.0, 1.0 / 1000.0, 0
    for n in range(steps):
        o.set_sdf(orc.SDF_SPHERE, (0.06, 0.06, 0.1395 - 0.1 * n * 1e-3), (0, 0, 0, 1), (0.02, 0, 0))
        det = time - last > between
        o.step(1, detect=det)
        detections += det
        if det:
            last = time
        time += 1e-3
    err = rel_err(x, o.x())
    print(f"block via xpbd_sim: rel err {err:.2e}, {detections} detections in {steps} steps")
    assert 0 < detections < steps
    assert err <= 1e-5
