"""FP64 CPU restatement of the warm-start network and its features.

TEST INFRASTRUCTURE ONLY (the checker of the device kernels): imported by
tests/ and never by the product path.

The network has no counterpart in the reference (the learned warm start is new
in this solver); these functions restate its definition in plain numpy:
  * forward / loss / gradient of the 24 -> 128 -> 128 -> 3 SiLU MLP, optionally
    with every GEMM operand rounded to bfloat16 (round-to-nearest-even) as the
    device's MFMA kernels do, accumulation in FP64;
  * the 24 per-vertex features the device computes after the inertial
    prediction.
Pinning: the gradient is checked against central finite differences of the
FP64 forward (tests/test_oracle_mlp.py); features are checked against
hand-computed values on small meshes.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 24, 128, 3
SIZES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_OUT, N_H)),
         ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round to bfloat16 (nearest, ties to even) through float32, returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _r(a, emulate):
    return bf16(a) if emulate else np.asarray(a, np.float64)


def forward(p, x, mean=None, std=None, scale=1.0, emulate_bf16=False, keep=False):
    """scale * MLP((x - mean) / std) for rows x (n x 24)."""
    P = unpack(p)
    x = np.asarray(x, np.float64)
    mean = np.zeros(N_IN) if mean is None else np.asarray(mean, np.float64)
    std = np.ones(N_IN) if std is None else np.asarray(std, np.float64)
    R = lambda a: _r(a, emulate_bf16)  # noqa: E731
    xs = (x - mean) / std
    z1 = R(xs) @ R(P["W1"]).T + P["b1"]
    h1 = silu(z1)
    z2 = R(h1) @ R(P["W2"]).T + P["b2"]
    h2 = silu(z2)
    y = R(h2) @ R(P["W3"]).T + P["b3"]
    out = scale * y
    if keep:
        return out, dict(xs=xs, z1=z1, h1=h1, z2=z2, h2=h2, y=y)
    return out


def loss_grad(p, x, t, mean=None, std=None, scale=1.0, emulate_bf16=False):
    """Loss in output units, mean over all n*3 outputs of (y - t / scale)^2 where
    scale * y is the network's output, and its gradient (packed layout)."""
    P = unpack(p)
    out, a = forward(p, x, mean, std, scale, emulate_bf16, keep=True)
    t = np.asarray(t, np.float64)
    n = len(t)
    d = a["y"] - t / scale
    loss = float(np.mean(d * d))
    R = lambda v: _r(v, emulate_bf16)  # noqa: E731
    dy = 2.0 * d / (n * N_OUT)
    g = {}
    g["W3"] = R(dy).T @ R(a["h2"])
    g["b3"] = dy.sum(0)
    dz2 = (R(dy) @ R(P["W3"])) * dsilu(a["z2"])
    g["W2"] = R(dz2).T @ R(a["h1"])
    g["b2"] = dz2.sum(0)
    dz1 = (R(dz2) @ R(P["W2"])) * dsilu(a["z1"])
    g["W1"] = R(dz1).T @ R(a["xs"])
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def one_ring(nv, tets):
    """Sorted unique 1-ring neighbour lists (CSR offsets, indices) from the element list."""
    t = np.asarray(tets, np.int64)
    a = np.repeat(t, 4, axis=1).reshape(-1)
    b = np.tile(t, (1, 4)).reshape(-1)
    keep = a != b
    pairs = np.unique(a[keep] * nv + b[keep])
    src, dst = pairs // nv, pairs % nv
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), dst


def features(x, v, cprev, rest, tets, pinned, E, nu, dt, g):
    """The 24 per-vertex features after the inertial prediction of one step.

    x, v, cprev: (V, 3) positions, velocities and previous solver correction at
    the start of the step; rest: (V, 3); E, nu per element; g gravity (acting
    along -z).  Columns: inertial displacement p = dt v + (0, 0, -g dt^2),
    velocity, 1-ring mean(u_j) - u_i with u = x - rest, 1-ring mean(p_j) - p_i,
    u_i, cprev_i, 1-ring mean(c_j) - c_i, mean log10 E and mean nu of the
    incident elements, fraction of 1-ring neighbours that are pinned."""
    x, v, c, X = (np.asarray(a, np.float64) for a in (x, v, cprev, rest))
    nv = len(x)
    off, nb = one_ring(nv, tets)
    cnt = np.diff(off)
    u = x - X
    p = dt * v
    p[:, 2] -= g * dt * dt
    def ring_mean(a):  # noqa: E306
        s = np.zeros_like(a)
        idx = np.repeat(np.arange(nv), cnt)
        np.add.at(s, idx, a[nb])
        out = np.zeros_like(a)
        has = cnt > 0
        out[has] = s[has] / cnt[has, None] - a[has]
        return out
    t = np.asarray(tets, np.int64)
    sE = np.zeros(nv)
    sN = np.zeros(nv)
    ct = np.zeros(nv)
    for j in range(4):
        np.add.at(sE, t[:, j], np.log10(np.asarray(E, np.float64)))
        np.add.at(sN, t[:, j], np.asarray(nu, np.float64))
        np.add.at(ct, t[:, j], 1.0)
    ct = np.maximum(ct, 1.0)
    pin = np.asarray(pinned, bool) if pinned is not None else np.zeros(nv, bool)
    pf = np.zeros(nv)
    idx = np.repeat(np.arange(nv), cnt)
    np.add.at(pf, idx, pin[nb].astype(np.float64))
    pf = np.where(cnt > 0, pf / np.maximum(cnt, 1), 0.0)
    return np.concatenate([p, v, ring_mean(u), ring_mean(p), u, c, ring_mean(c), (sE / ct)[:, None],
                           (sN / ct)[:, None], pf[:, None]], axis=1)
