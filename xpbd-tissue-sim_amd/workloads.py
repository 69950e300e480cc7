"""Synthetic inputs for the benchmark configurations.

All meshes are structured grids of hexahedral cells split 6 tets per cell
(conforming Kuhn split).  Random draws use numpy's PCG64 with the given seed,
so every caller (device solver, oracle, benchmark) sees identical arrays.

  cube      10x10x10 cells of 0.01 m (1,331 V / 6,000 T), E=3e3, nu=0.45,
            bottom face pinned, velocity noise N(0, 0.01 m/s)
  beam      200x20x20 cells of 0.01 m (88,641 V / 480,000 T), 2.0x0.2x0.2 m,
            E=1e4, nu=0.49, x=min end pinned, velocity noise
  block     56^3 cells of 0.01 m (185,193 V / 1,053,696 T), per-tet
            E ~ logU[1e3, 5e4], nu ~ U[0.45, 0.49], base pinned, sphere
            r = 0.05 m pressed down at 0.1 m/s
  batch     256 cubes of 4..24 cells per axis, per-instance E, nu, each with
            its own box collider
"""
from __future__ import annotations

import numpy as np

from xpbd_amd import Mesh, grid_mesh, surface_faces, SDF_SPHERE, SDF_BOX

RHO = 1000.0


def grid_object(nx, ny, nz, h=0.01, E=3e3, nu=0.45, pin_axis=2, pin_side="min", noise=0.01, seed=0,
                faces=False, origin=(0.0, 0.0, 0.0), rng=None):
    x, t = grid_mesh(nx, ny, nz, h, origin)
    rng = np.random.default_rng(seed) if rng is None else rng
    nt = len(t)
    Ea = np.full(nt, float(E)) if np.isscalar(E) else np.asarray(E, np.float64)
    nua = np.full(nt, float(nu)) if np.isscalar(nu) else np.asarray(nu, np.float64)
    pinned = None
    if pin_axis is not None:
        coord = x[:, pin_axis]
        ref = coord.min() if pin_side == "min" else coord.max()
        pinned = (np.abs(coord - ref) < 1e-9).astype(np.uint8)
    v0 = rng.normal(0.0, noise, size=x.shape) if noise else None
    f = surface_faces(len(x), t) if faces else None
    return Mesh(x0=x, tets=t, E=Ea, nu=nua, rho=np.full(nt, RHO), pinned=pinned, v0=v0, faces=f)


def cube(n=10, seed=1, **kw):
    kw.setdefault("E", 3e3)
    kw.setdefault("nu", 0.45)
    return grid_object(n, n, n, 0.01, pin_axis=2, seed=seed, **kw)


def beam(nx=200, ny=20, nz=20, seed=2, **kw):
    kw.setdefault("E", 1e4)
    kw.setdefault("nu", 0.49)
    return grid_object(nx, ny, nz, 0.01, pin_axis=0, seed=seed, **kw)


def heterogeneous(n_tets, rng, E_range=(1e3, 5e4), nu_range=(0.45, 0.49)):
    E = np.exp(rng.uniform(np.log(E_range[0]), np.log(E_range[1]), n_tets))
    nu = rng.uniform(nu_range[0], nu_range[1], n_tets)
    return E, nu


def block(n=56, seed=3, faces=True, noise=0.0):
    rng = np.random.default_rng(seed)
    nt = 6 * n ** 3
    E, nu = heterogeneous(nt, rng)
    m = grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=noise, faces=faces, rng=rng)
    return m


def block_sphere(n=56, r=0.05):
    """Sphere collider resting on the centre of the block's top face (pose, params, velocity)."""
    L = 0.01 * n
    return SDF_SPHERE, (L / 2, L / 2, L + r), (0.0, 0.0, 0.0, 1.0), (r, 0.0, 0.0), (0.0, 0.0, -0.1)


def random_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def batch(n_inst=256, lo=4, hi=24, seed=4, faces=True):
    """Independent cubes of uneven size, each with a box collider overlapping its top."""
    rng = np.random.default_rng(seed)
    meshes, sdfs = [], []
    for i in range(n_inst):
        nx, ny, nz = rng.integers(lo, hi + 1, size=3)
        E = float(np.exp(rng.uniform(np.log(1e3), np.log(5e4))))
        nu = float(rng.uniform(0.45, 0.49))
        m = grid_object(int(nx), int(ny), int(nz), 0.01, E=E, nu=nu, pin_axis=2, noise=0.0, faces=faces, rng=rng)
        meshes.append(m)
        ext = np.array([nx, ny, nz]) * 0.01
        size = np.full(3, 0.5 * ext.min())
        q = random_quat(rng)
        # centre above the top face so the box reaches about 10% of the height into the object
        pos = (ext[0] / 2, ext[1] / 2, ext[2] * 0.9 + 0.5 * size[0])
        sdfs.append((SDF_BOX, pos, tuple(q), tuple(size)))
    return meshes, sdfs
