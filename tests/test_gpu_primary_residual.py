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
