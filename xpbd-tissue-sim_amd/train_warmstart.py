"""Train the per-vertex warm-start network on the device.

Training data come from the plain Gauss-Seidel solver (20 iterations) on a
seeded family of meshes:
  * beams like the 200 x 20 x 20-cell stiff beam (0.01 m cells, E = 1e4 Pa,
    nu = 0.49, one end pinned) at reduced size, 24..60 x 4..8 x 4..8 cells,
    with seeded velocity noise and a seeded random load: a sphere collider
    driven into the beam at a random point and speed;
  * heterogeneous blocks (E log-uniform in [1e3, 5e4] Pa, nu in [0.45, 0.49],
    base pinned) of 8..14 cells per side pressed by a moving sphere, like the
    1M-element tissue block.
Every step yields one row per free vertex: the 24 features after the
inertial prediction and the solver's correction x_solved - x_predicted.  A
seeded fraction of each step's rows is kept (row subsampling), so that many
steps of many meshes fit one training set.  Collection and training (BF16
MFMA forward/backward, Adam) both run through the HIP library.

    python train_warmstart.py --out weights/warmstart.npz \
        [--beams 200 --blocks 40 --steps 250 --keep 0.0625 --epochs 20]

(200 beams x 250 steps = 50k steps of the beam family, plus 40 blocks x 250.)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import workloads as W  # noqa: E402
import xpbd_amd as X  # noqa: E402


def family_member(kind: str, seed: int):
    """(mesh, sphere pose/params/velocity) of one seeded training mesh."""
    rng = np.random.default_rng(seed)
    if kind == "beam":
        nx, ny = int(rng.integers(24, 61)), int(rng.integers(4, 9))
        mesh = W.grid_object(nx, ny, ny, 0.01, E=1e4, nu=0.49, pin_axis=0, noise=0.01, faces=True, rng=rng)
        L = np.array([nx, ny, ny]) * 0.01
        r = rng.uniform(0.01, 0.03)
        c = (rng.uniform(0.3, 1.0) * L[0], rng.uniform(0.2, 0.8) * L[1])
    else:
        n = int(rng.integers(8, 15))
        E, nu = W.heterogeneous(6 * n ** 3, rng)
        mesh = W.grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=0.01, faces=True, rng=rng)
        L = np.array([n, n, n]) * 0.01
        r = rng.uniform(0.015, 0.05)
        c = (rng.uniform(0.3, 0.7) * L[0], rng.uniform(0.3, 0.7) * L[1])
    pose = (c[0], c[1], L[2] + r - 0.0005)
    vel = (0.0, 0.0, -rng.uniform(0.05, 0.2))
    return mesh, (X.SDF_SPHERE, pose, (0.0, 0.0, 0.0, 1.0), (r, 0.0, 0.0), vel)


def collect(kind: str, seed: int, steps: int, iters: int = 20, keep: float = 1.0, chunk: int = 50):
    """Rows of `steps` plain steps of one family member; a seeded fraction
    `keep` of each chunk's rows is kept."""
    mesh, (k, pos, quat, prm, vel) = family_member(kind, seed)
    s = X.Solver(mesh, dt=1e-3, iters=iters, solver="gs")
    s.set_sdf(k, pos, quat, prm)
    s.set_sdf_velocity(vel)
    s.attach(X.Network())
    rng = np.random.default_rng(seed + 7919)
    xs, ys = [], []
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        x, y = s.collect_training(n)
        if keep < 1.0:
            sel = rng.random(len(x)) < keep
            x, y = x[sel], y[sel]
        xs.append(x)
        ys.append(y)
        done += n
    s.close()
    return np.concatenate(xs), np.concatenate(ys)


def build_dataset(beams: int, blocks: int, steps: int, keep: float, seed0: int = 1000):
    xs, ys = [], []
    members = [("beam", seed0 + 2 * i) for i in range(beams)] + [("block", seed0 + 2 * i + 1) for i in range(blocks)]
    t0 = time.time()
    for i, (kind, seed) in enumerate(members):
        x, y = collect(kind, seed, steps, keep=keep)
        xs.append(x)
        ys.append(y)
        if (i + 1) % 20 == 0 or i + 1 == len(members):
            print(f"collected {i + 1}/{len(members)} meshes, {sum(len(v) for v in xs)} rows, "
                  f"{time.time() - t0:.1f} s", flush=True)
    return np.concatenate(xs), np.concatenate(ys)


def save(path, net: X.Network, mean, std, scale, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=net.params(), mean=np.asarray(mean, np.float32), std=np.asarray(std, np.float32),
             scale=np.float32(scale), meta=np.array(json.dumps(meta)))
    # the same network as a file that C and C++ callers read with xpbd_mlp_load
    net.save(os.path.splitext(path)[0] + ".bin")


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart.npz")


def load(path) -> X.Network:
    z = np.load(path, allow_pickle=False)
    net = X.Network(z["params"])
    net.set_scaling(z["mean"], z["std"], float(z["scale"]))
    return net


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                                                 "warmstart.npz"))
    ap.add_argument("--beams", type=int, default=200, help="beam-family meshes")
    ap.add_argument("--blocks", type=int, default=40, help="block-family meshes")
    ap.add_argument("--steps", type=int, default=250, help="steps per mesh")
    ap.add_argument("--keep", type=float, default=0.0625, help="fraction of each step's rows kept")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    a = ap.parse_args()
    t0 = time.time()
    x, y = build_dataset(a.beams, a.blocks, a.steps, a.keep)
    t1 = time.time()
    xe, ye = collect("block", 99, 100)
    net = X.Network()
    mean, std, scale = net.fit_scaling(x, y)
    loss = net.train(x, y, batch=a.batch, epochs=a.epochs, lr=a.lr, seed=7)
    t2 = time.time()
    pred = net.forward(xe)
    err = float(np.linalg.norm(pred - ye) / np.linalg.norm(ye))
    err_prev = float(np.linalg.norm(xe[:, 15:18] - ye) / np.linalg.norm(ye))
    meta = dict(rows=int(len(x)), beams=a.beams, blocks=a.blocks, steps_per_mesh=a.steps,
                beam_steps=a.beams * a.steps, block_steps=a.blocks * a.steps, keep=a.keep,
                epochs=a.epochs, batch=a.batch, lr=a.lr,
                final_batch_loss=loss, heldout_rel_err=err, heldout_prev_correction_rel_err=err_prev,
                collect_s=t1 - t0, train_s=t2 - t1)
    save(a.out, net, mean, std, scale, meta)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
