"""Primary residual (the momentum balance of Equation (8) of the XPBD paper)
on the device against the FP64 oracle's, one step at a time from a shared
state.

Both record (switched on with xpbd_set_primary_residual on the device), at
the end of the global solve (before local collision passes
and the pin reset, where the reference's EVERY_SUBSTEP policy evaluates its
residuals, XPBDSolver.hpp:91-96),
    R_v = (x_v - xt_v) / w_v - sum grad_v C^T lambda
over both constraints of every element and every active contact around v,
zero on pinned and attached vertices.  Every step starts the oracle from the
device's positions and velocities and gives it the device's contact set.

Gate: per vertex, |R_gpu - R_ref| <= 1e-4 of the largest inertial term
max |(x - xt) / w| of the step (the FP32 state and multipliers against FP64),
and the norms within 1%.  R itself is 1e-3 to 1e-1 of the inertial term for
these solves, so a dropped or mis-signed contact or element term, or a
prediction taken at the wrong point, exceeds the gate by orders of magnitude.
"""
import numpy as np
import pytest

import orc
import workloads as W
import xpbd_amd as X

pytestmark = pytest.mark.gpu
DT, G = 1e-3, 9.81
TOL = 1e-4


def _block(seed=31, depth=0.002, n=6):
    rng = np.random.default_rng(seed)
    E, nu = W.heterogeneous(6 * n ** 3, rng, E_range=(1e3, 1e4))
    mesh = W.grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=0.005, faces=True, rng=rng)
    L = 0.01 * n
    sph = (X.SDF_SPHERE, (0.47 * L, 0.53 * L, L + 0.02 - depth), (0, 0, 0, 1), (0.02, 0, 0))
    return mesh, sph


def _collider(gpu, sph):
    gpu.set_sdf(*sph)
    gpu.set_sdf_velocity((0.0, 0.0, -0.05))


def _pair(mesh, solver, variant, iters, det=1, first_order=None):
    kw = {}
    if first_order is not None:
        kw = dict(first_order=True, damping_multiplier=first_order, adjust_damping=True)
    gpu = X.Solver(mesh, dt=DT, g=G, iters=iters, solver=solver, variant=variant, deterministic=det,
                   collision_every=1, **kw)
    o = orc.Sim(mesh, dt=DT, g=G, iters=iters, mode=solver)
    if first_order is not None:
        o.set_first_order(first_order, True)
    if solver == "gs":
        o.set_order(gpu.tet_order())
        o.set_contact_order(gpu.face_rank())
    gpu.set_primary_residual(True)
    o.set_primary_residual(True)
    return gpu, o


def _run(gpu, o, mesh, nsteps, sdf=None, grasp=None):
    """Per-step comparison; returns (max error / inertial scale, max |R| /
    inertial scale, norm ratios, contact counts)."""
    if sdf is not None:
        _collider(gpu, sdf)
    w = o.inverse_inertia()
    out = np.zeros(len(w), bool)
    if mesh.pinned is not None:
        out |= mesh.pinned.astype(bool)
    if grasp is not None:
        out[grasp] = True
    errs, rels, ratios, ncon = [], [], [], []
    for k in range(nsteps):
        if grasp is not None:
            tgt = mesh.x0[grasp] + np.array([0.004, -0.002, 0.01]) * min(k + 1, 20) / 20
            gpu.set_attachments(grasp, tgt)
            o.set_attachments(grasp, tgt)
        x0, v0 = gpu.positions(), gpu.velocities()
        gpu.step(1)
        o.set_state(x0, v0)
        if sdf is not None:
            face, cp, cn, cb = gpu.contacts()
            o.set_contacts(face, cp, cn, cb)
            ncon.append(len(face))
        o.step(1, detect=False)
        ng, Rg = gpu.last_step_primary_residual(vec=True)
        nr, Rr = o.last_primary_residual(vec=True)
        scale = np.abs((o.x() - o.last_prediction()) / w[:, None])[~out].max()
        assert np.all(Rg[out] == 0) and np.all(Rr[out] == 0)
        assert ng == pytest.approx(np.linalg.norm(Rg), rel=1e-10)
        errs.append(np.abs(Rg - Rr).max() / scale)
        rels.append(np.abs(Rr).max() / scale)
        ratios.append(ng / nr)
    return np.array(errs), np.array(rels), np.array(ratios), ncon


@pytest.mark.parametrize("solver,variant,det", [("gs", "quad", 1), ("gs", "cluster", 1), ("jacobi", "patch", 1),
                                                ("jacobi", "patch", 0)])
def test_primary_residual_with_contacts_matches_oracle(solver, variant, det):
    """A heterogeneous base-pinned block pressed by a descending sphere,
    contacts detected on the device every step."""
    mesh, sph = _block()
    gpu, o = _pair(mesh, solver, variant, 10, det)
    errs, rels, ratios, ncon = _run(gpu, o, mesh, 20, sdf=sph)
    print(f"primary residual {solver}/{variant}/det{det}: contacts {min(ncon)}..{max(ncon)}, "
          f"|R|/inertia {rels.min():.2e}..{rels.max():.2e}, max err/inertia {errs.max():.2e}, "
          f"norm ratio {ratios.min():.6f}..{ratios.max():.6f}")
    assert max(ncon) > 0
    assert rels.min() > 10 * TOL
    assert errs.max() <= TOL
    assert np.all(np.abs(ratios - 1) <= 1e-2)


@pytest.mark.parametrize("solver,variant", [("gs", "quad"), ("jacobi", "patch")])
def test_primary_residual_with_attachments_matches_oracle(solver, variant):
    """Three top vertices grasped and pulled: they read zero on both sides and
    the rest of the mesh matches."""
    mesh, _ = _block(seed=5)
    top = np.where(np.abs(mesh.x0[:, 2] - mesh.x0[:, 2].max()) < 1e-9)[0]
    grasp = top[[0, len(top) // 2, len(top) - 1]]
    gpu, o = _pair(mesh, solver, variant, 10)
    errs, rels, ratios, _ = _run(gpu, o, mesh, 20, grasp=grasp)
    print(f"primary residual with attachments {solver}/{variant}: |R|/inertia {rels.min():.2e}..{rels.max():.2e}, "
          f"max err/inertia {errs.max():.2e}, norm ratio {ratios.min():.6f}..{ratios.max():.6f}")
    assert errs.max() <= TOL
    assert np.all(np.abs(ratios - 1) <= 1e-2)


def test_primary_residual_first_order_matches_oracle():
    """First order: w is the inverse damping and the prediction the damped
    gravity move."""
    mesh, sph = _block(seed=9)
    gpu, o = _pair(mesh, "gs", "quad", 10, first_order=20.0)
    errs, rels, ratios, _ = _run(gpu, o, mesh, 20, sdf=sph)
    print(f"primary residual first order: |R|/inertia {rels.min():.2e}..{rels.max():.2e}, "
          f"max err/inertia {errs.max():.2e}, norm ratio {ratios.min():.6f}..{ratios.max():.6f}")
    assert errs.max() <= TOL
    assert np.all(np.abs(ratios - 1) <= 1e-2)


def test_primary_residual_switch_and_reproducible():
    """Off: XPBD_E_STATE.  On: zero before the first step; independent of the
    residual policy (the constraint residual stays off); two handles of the
    same configuration record the same bits (a fixed-order gather)."""
    mesh, sph = _block()
    gpu = X.Solver(mesh, dt=DT, iters=10, solver="gs", collision_every=1)
    with pytest.raises(X.XpbdError):
        gpu.last_step_primary_residual()
    gpu.set_primary_residual(True)
    assert gpu.last_step_primary_residual() == 0.0
    _collider(gpu, sph)
    twin = X.Solver(mesh, dt=DT, iters=10, solver="gs", collision_every=1)
    twin.set_primary_residual(True)
    twin.set_residual_policy(True)
    _collider(twin, sph)
    for _ in range(5):
        gpu.step(1)
        twin.step(1)
        a, ra = gpu.last_step_primary_residual(vec=True)
        b, rb = twin.last_step_primary_residual(vec=True)
        assert a > 0 and a == b and np.array_equal(ra, rb)
    with pytest.raises(X.XpbdError):
        gpu.last_step_residual()
    assert twin.last_step_residual() > 0
    gpu.set_primary_residual(False)
    with pytest.raises(X.XpbdError):
        gpu.last_step_primary_residual()


def test_primary_residual_of_the_learned_step_matches_oracle():
    """The learned step (network correction, multiplier start, 3 iterations,
    no fallback) against the oracle's orc_step_warm from the same state and
    correction: the prediction kept is the inertial one, before the
    correction.  The multiplier start breaks the momentum balance, so its
    residual is larger than the plain solve's from the same state."""
    import train_warmstart as TW
    mesh, sph = _block(seed=17, depth=0.004)
    gpu, o = _pair(mesh, "gs", "quad", 20)
    plain, _ = _pair(mesh, "gs", "quad", 20)
    net = TW.load(TW.DEFAULT_WEIGHTS)
    gpu.attach(net)
    _collider(gpu, sph)
    _collider(plain, sph)
    w = o.inverse_inertia()
    pinned = mesh.pinned.astype(bool)
    errs, ratios, over_plain = [], [], []
    for k in range(15):
        x0, v0 = gpu.positions(), gpu.velocities()
        corr = net.forward(gpu.features()).astype(np.float64)
        corr[pinned] = 0.0
        plain.set_state(x0, v0)
        plain.step(1)
        gpu.step_learned(3, ref_residual=0.0)
        face, cp, cn, cb = gpu.contacts()
        o.set_state(x0, v0)
        o.set_contacts(face, cp, cn, cb)
        o.step_warm(corr, 3, detect=False)
        ng, Rg = gpu.last_step_primary_residual(vec=True)
        nr, Rr = o.last_primary_residual(vec=True)
        scale = np.abs((o.x() - o.last_prediction()) / w[:, None])[~pinned].max()
        errs.append(np.abs(Rg - Rr).max() / scale)
        ratios.append(ng / nr)
        over_plain.append(ng / plain.last_step_primary_residual())
    errs, ratios = np.array(errs), np.array(ratios)
    print(f"primary residual, learned step: max err/inertia {errs.max():.2e}, norm ratio "
          f"{ratios.min():.6f}..{ratios.max():.6f}, learned / plain-20 median {np.median(over_plain):.1f}")
    assert errs.max() <= TOL
    assert np.all(np.abs(ratios - 1) <= 1e-2)
    assert np.median(over_plain) > 3


def test_primary_residual_of_a_batch_matches_each_instance():
    """Four uneven cubes in one batched handle, each with its own sphere;
    every instance's slice of the residual against its own oracle from the
    shared state and the device's contacts."""
    rng = np.random.default_rng(23)
    meshes, sdfs = [], []
    for _ in range(4):
        nx, ny, nz = (int(v) for v in rng.integers(3, 8, size=3))
        m = W.grid_object(nx, ny, nz, 0.01, E=float(rng.uniform(2e3, 2e4)), nu=0.47, pin_axis=2, noise=0.0,
                          faces=True, rng=rng)
        ext = np.array([nx, ny, nz]) * 0.01
        r = 0.6 * ext.min()
        meshes.append(m)
        sdfs.append((X.SDF_SPHERE, (ext[0] * 0.43, ext[1] * 0.57, ext[2] + r - 0.08 * ext[2]), (0, 0, 0, 1),
                     (r, 0, 0)))
    gpu = X.Solver(meshes, dt=DT, g=G, iters=10, solver="gs", variant="quad", collision_every=1)
    for i, (kind, pos, quat, prm) in enumerate(sdfs):
        gpu.set_sdf(kind, pos, quat, prm, instance=i)
    gpu.set_primary_residual(True)
    t_off = np.concatenate([[0], np.cumsum([m.nt for m in meshes])])
    v_off = np.concatenate([[0], np.cumsum([m.nv for m in meshes])])
    f_off = np.concatenate([[0], np.cumsum([len(m.faces) for m in meshes])])
    order, rank = gpu.tet_order(), gpu.face_rank()
    orcs = []
    for i, m in enumerate(meshes):
        o = orc.Sim(m, dt=DT, g=G, iters=10, mode="gs")
        o.set_order(order[(order >= t_off[i]) & (order < t_off[i + 1])] - t_off[i])
        o.set_contact_order(rank[f_off[i]:f_off[i + 1]])
        o.set_primary_residual(True)
        orcs.append(o)
    worst, ncon = 0.0, 0
    for _ in range(15):
        x0, v0 = gpu.positions(), gpu.velocities()
        gpu.step(1)
        face, cp, cn, cb = gpu.contacts()
        ncon = max(ncon, len(face))
        _, Rg = gpu.last_step_primary_residual(vec=True)
        for i, o in enumerate(orcs):
            vs = slice(v_off[i], v_off[i + 1])
            o.set_state(x0[vs], v0[vs])
            sel = (face >= f_off[i]) & (face < f_off[i + 1])
            o.set_contacts(face[sel] - f_off[i], cp[sel], cn[sel], cb[sel])
            o.step(1, detect=False)
            _, Rr = o.last_primary_residual(vec=True)
            free = ~meshes[i].pinned.astype(bool)
            scale = np.abs((o.x() - o.last_prediction()) / o.inverse_inertia()[:, None])[free].max()
            worst = max(worst, np.abs(Rg[vs] - Rr).max() / scale)
    print(f"primary residual, batch of {len(meshes)}: contacts up to {ncon}, max err/inertia {worst:.2e}")
    assert ncon > 0
    assert worst <= TOL
