"""Device solver vs the FP64 oracle at the configurations' full sizes, one step
at a time from a shared state.

Every step starts the oracle from the device's state (positions and
velocities read back in FP64), steps both once and compares:
  - C2  beam 200x20x20 cells (480,000 tets), E = 1e4, nu = 0.49, 10 iterations,
        colour Gauss-Seidel and Jacobi (the reference's plain-sum Jacobi
        diverges on this beam; each step is still compared from the shared state)
  - C3  tissue block 56^3 cells (1,053,696 tets), heterogeneous E / nu, sphere
        r = 0.05 m moving down at 0.1 m/s, 20 colour-GS iterations
  - C4  256 cubes of 4..24 cells per axis with a box collider each, 10 GS
        iterations, every instance against its own oracle
Contacts: the oracle detects on the same state with the same collider pose and
must find the same contacts (face set; points, normals and barycentrics to
round-off), and then projects the device's contact set, so a detection
difference could not hide in the trajectory.  The oracle runs the device's
own visiting order (colour levels on OpenMP threads, bit for bit its
sequential pass).

Gate: 1e-5 relative position error per step (max |x - x_ref| / max |x_ref|).
The error in units of the largest displacement from rest (max |x - x_ref| /
max |x_ref - X|) is printed beside it.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import orc
import workloads as W
import xpbd_amd as X
from helpers import disp_err, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
DT = 1e-3
THREADS = max(1, min(16, os.cpu_count() or 1))


def _oracle(gpu, mesh, iters, mode):
    o = orc.Sim(mesh, dt=DT, iters=iters, mode=mode)
    if mode == "gs":
        o.set_order_levels(gpu.tet_order(), gpu.color_offsets())
        if mesh.faces is not None:
            o.set_contact_order(gpu.face_rank())
    o.set_threads(THREADS)
    return o


def _pose(sdf, n):
    """Collider pose at the n-th step after set_sdf, as the device forms it."""
    kind, pos, quat, prm, vel = sdf
    t = float(n) * DT
    return kind, [pos[k] + vel[k] * t for k in range(3)], quat, prm


def _compare_detection(o, face, cp, cn, cb, f0=0):
    """The oracle's own contacts on the current state against the device's
    (faces f0.. of the oracle's numbering).  Returns (set equal, max deviation)."""
    fo, po, no, bo = o.contacts()
    if len(fo) != len(face) or not np.array_equal(np.sort(fo), np.sort(face - f0)):
        return False, np.inf
    if len(fo) == 0:
        return True, 0.0
    a, b = np.argsort(fo), np.argsort(face)
    dev = max(np.max(np.abs(po[a] - cp[b])), np.max(np.abs(no[a] - cn[b])), np.max(np.abs(bo[a] - cb[b])))
    return True, float(dev)


def _shared_state_run(gpu, o, mesh, nsteps, sdf=None):
    errs, derrs, ncon, devs = [], [], [], []
    for n in range(nsteps):
        x0, v0 = gpu.positions(), gpu.velocities()
        gpu.step(1)
        o.set_state(x0, v0)
        if sdf is not None:
            face, cp, cn, cb = gpu.contacts()
            o.set_sdf(*_pose(sdf, n))
            o.detect()
            same, dev = _compare_detection(o, face, cp, cn, cb)
            assert same, f"step {n}: device and oracle detect different contact sets"
            devs.append(dev)
            o.set_contacts(face, cp, cn, cb)
            ncon.append(len(face))
        o.step(1, detect=False)
        x, xr = gpu.positions(), o.x()
        if not np.all(np.isfinite(xr)):
            break
        errs.append(rel_err(x, xr))
        derrs.append(disp_err(x, xr, mesh.x0))
    return np.array(errs), np.array(derrs), ncon, devs


@pytest.mark.parametrize("solver", ["gs", "jacobi"])
def test_c2_beam_per_step_from_shared_state(solver):
    mesh = W.beam(200, 20, 20)
    gpu = X.Solver(mesh, dt=DT, iters=10, solver=solver, collision_every=0)
    o = _oracle(gpu, mesh, 10, solver)
    errs, derrs, _, _ = _shared_state_run(gpu, o, mesh, 100)
    vmax = np.abs(gpu.velocities()).max()
    at = {k: f"{errs[k - 1]:.2e}" for k in (1, 10, 25, 50, 75, 100) if k <= len(errs)}
    print(f"C2 beam {solver} ({mesh.nt} tets), {len(errs)} steps: rel err max {errs.max():.3e} (by step {at}), "
          f"displacement-relative max {derrs.max():.3e} (last {derrs[-1]:.3e}); |v|max at the end {vmax:.3g} m/s")
    assert len(errs) >= 50
    assert errs.max() <= TOL


def test_c3_block_sphere_per_step_from_shared_state():
    mesh = W.block(56)
    gpu = X.Solver(mesh, dt=DT, iters=20, solver="gs", collision_every=1)
    sdf = W.block_sphere(56)
    kind, pos, quat, prm, vel = sdf
    gpu.set_sdf(kind, pos, quat, prm)
    gpu.set_sdf_velocity(vel)
    o = _oracle(gpu, mesh, 20, "gs")
    errs, derrs, ncon, devs = _shared_state_run(gpu, o, mesh, 100, sdf=sdf)
    print(f"C3 block ({mesh.nt} tets), {len(errs)} steps: rel err max {errs.max():.3e}, displacement-relative "
          f"max {derrs.max():.3e} (last {derrs[-1]:.3e}); contacts per step up to {max(ncon)}, detection equal on "
          f"every step, largest point/normal/barycentric deviation {max(devs):.3e}")
    # in the first 100 steps the block sags under gravity about as fast as the
    # sphere descends, so few faces touch; test_c3_sphere_over_grid_detection
    # presses the same sphere into the block for many contacts
    assert max(ncon) > 0
    assert errs.max() <= TOL
    assert max(devs) <= 1e-12


def test_large_block_ten_times_c3_per_step_from_shared_state():
    """Far past the configurations: a 120^3-cell block (10,368,000 tets,
    1,771,561 vertices, ten times C3) with C3's materials, the sphere pressed
    1 cm into its top face, 20 GS iterations; three steps, each from the
    shared state, against the oracle, with detection compared."""
    n = 120
    mesh = W.block(n)
    gpu = X.Solver(mesh, dt=DT, iters=20, solver="gs", collision_every=1)
    kind, pos, quat, prm, vel = W.block_sphere(n)
    sdf = (kind, (pos[0], pos[1], pos[2] - 0.01), quat, prm, vel)
    gpu.set_sdf(*sdf[:4])
    gpu.set_sdf_velocity(vel)
    o = _oracle(gpu, mesh, 20, "gs")
    errs, derrs, ncon, devs = _shared_state_run(gpu, o, mesh, 3, sdf=sdf)
    print(f"large block ({mesh.nt} tets, {mesh.nv} vertices, {gpu.launch_config()}), {len(errs)} steps: rel err max "
          f"{errs.max():.3e}, displacement-relative max {derrs.max():.3e}; contacts per step up to {max(ncon)}, "
          f"detection equal on every step, largest deviation {max(devs):.3e}")
    assert min(ncon) > 0
    assert errs.max() <= TOL
    assert max(devs) <= 1e-12


@pytest.mark.parametrize("solver,E_range,b", [("gs", (1e3, 5e4), 2e5), ("gs", (1e3, 5e3), 2e4),
                                              ("jacobi", (1e3, 3e3), 2e6)])
def test_first_order_with_contacts_per_step_from_shared_state(solver, E_range, b):
    """First-order XPBD with sphere contacts and two local collision passes,
    per step from the shared state, contacts handed over and detection
    compared, 30 steps of a 12^3 heterogeneous block pressed by a sphere.
    Whole trajectories cannot be compared here, and neither can single steps
    at low damping: there the FP64 oracle alone amplifies a state perturbation
    1e4x to 3e8x in one step (b = 200..2e4 on E up to 5e4, b = 2e3 on E up to
    5e3, Jacobi up to b = 2e5).  These damping multipliers are where one step
    amplifies it ~3.5x, which the gate needs."""
    n = 12
    rng = np.random.default_rng(41)
    E, nu = W.heterogeneous(6 * n ** 3, rng, E_range=E_range)
    mesh = W.grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=0.0, faces=True, rng=rng)
    L = 0.01 * n
    # the sphere keeps pressing: a first-order block yields at once, so a slow probe loses contact
    sdf = (X.SDF_SPHERE, (0.47 * L, 0.53 * L, L + 0.03 - 0.002), (0.0, 0.0, 0.0, 1.0), (0.03, 0.0, 0.0),
           (0.0, 0.0, -0.3))
    gpu = X.Solver(mesh, dt=DT, iters=10, solver=solver, collision_every=1, first_order=True, damping_multiplier=b,
                   local_collision_iters=2)
    gpu.set_sdf(*sdf[:4])
    gpu.set_sdf_velocity(sdf[4])
    o = _oracle(gpu, mesh, 10, solver)
    o.set_first_order(b, False)
    o.set_local_collision_iters(2)
    errs, derrs, ncon, devs = _shared_state_run(gpu, o, mesh, 30, sdf=sdf)
    touching = sum(k > 0 for k in ncon)
    print(f"first order {solver}, E {E_range}, b = {b:g}, contacts and 2 local passes: {len(errs)} steps, rel err "
          f"max {errs.max():.3e}, displacement-relative max {derrs.max():.3e}; contacts on {touching} steps, up to "
          f"{max(ncon)}, detection equal on every step")
    assert len(errs) == 30 and touching >= 10
    assert errs.max() <= TOL


@pytest.mark.parametrize("depth", [0.002, 0.01, 0.025])
def test_c3_sphere_over_grid_detection_equals_oracle(depth):
    """The C3 sphere centred over the block's grid of top faces (where
    Frank-Wolfe's vertex choice ties exactly), pressed `depth` into the block:
    device detection on the device's state equals the oracle's on the same
    state, contact for contact, at rest and after steps of deformation."""
    mesh = W.block(56)
    gpu = X.Solver(mesh, dt=DT, iters=20, solver="gs", collision_every=1)
    kind, pos, quat, prm, _ = W.block_sphere(56)
    p = (pos[0], pos[1], pos[2] - depth)
    gpu.set_sdf(kind, p, quat, prm)
    o = orc.Sim(mesh, dt=DT, iters=1, mode="gs")
    o.set_sdf(kind, p, quat, prm)
    counts, worst = [], 0.0
    for _ in range(4):
        o.set_state(gpu.positions(), None)
        gpu.detect()
        o.detect()
        face, cp, cn, cb = gpu.contacts()
        same, dev = _compare_detection(o, face, cp, cn, cb)
        assert same, f"depth {depth}: contact sets differ ({len(face)} on the device, {len(o.contacts()[0])} oracle)"
        counts.append(len(face))
        worst = max(worst, dev)
        gpu.step(5)
    print(f"C3 sphere {1e3 * depth:.0f} mm deep: contacts {counts}, equal sets, largest deviation {worst:.3e}")
    # the block springs back from the sphere within a few steps; the first
    # detection is the one on the undeformed grid
    assert counts[0] > 0
    assert worst <= 1e-12


def test_c4_batch_256_cubes_box_contacts_per_step():
    meshes, sdfs = W.batch(256)
    gpu = X.Solver(meshes, dt=DT, iters=10, solver="gs", collision_every=1)
    for i, (kind, pos, quat, prm) in enumerate(sdfs):
        gpu.set_sdf(kind, pos, quat, prm, instance=i)
    gpu.autotune()
    n = len(meshes)
    t_off = np.concatenate([[0], np.cumsum([m.nt for m in meshes])])
    v_off = np.concatenate([[0], np.cumsum([m.nv for m in meshes])])
    f_off = np.concatenate([[0], np.cumsum([len(m.faces) for m in meshes])])
    order, rank = gpu.tet_order(), gpu.face_rank()
    oracles = []
    for i, m in enumerate(meshes):
        o = orc.Sim(m, dt=DT, iters=10, mode="gs")
        o.set_order(order[(order >= t_off[i]) & (order < t_off[i + 1])] - t_off[i])
        o.set_contact_order(rank[f_off[i]:f_off[i + 1]])
        kind, pos, quat, prm = sdfs[i]
        o.set_sdf(kind, pos, quat, prm)
        oracles.append(o)
    worst = np.zeros(3)
    total_contacts, mismatched = 0, []
    with ThreadPoolExecutor(THREADS) as pool:
        for step in range(100):
            x0, v0 = gpu.positions(), gpu.velocities()
            gpu.step(1)
            face, cp, cn, cb = gpu.contacts()
            total_contacts = max(total_contacts, len(face))
            x = gpu.positions()

            def one(i):
                o = oracles[i]
                o.set_state(x0[v_off[i]:v_off[i + 1]], v0[v_off[i]:v_off[i + 1]])
                sel = (face >= f_off[i]) & (face < f_off[i + 1])
                o.detect()
                same, dev = _compare_detection(o, face[sel], cp[sel], cn[sel], cb[sel], f0=f_off[i])
                o.set_contacts(face[sel] - f_off[i], cp[sel], cn[sel], cb[sel])
                o.step(1, detect=False)
                xr = o.x()
                xi = x[v_off[i]:v_off[i + 1]]
                return same, dev, rel_err(xi, xr), disp_err(xi, xr, meshes[i].x0)

            for i, (same, dev, e, d) in enumerate(pool.map(one, range(n))):
                if not same:
                    mismatched.append((step, i))
                worst = np.maximum(worst, [dev if same else 0.0, e, d])
    print(f"C4 batch ({gpu.nt} tets in {n} instances, {gpu.launch_config()}), 100 steps: rel err max {worst[1]:.3e}, "
          f"displacement-relative max {worst[2]:.3e}; contacts per step up to {total_contacts}, detection "
          f"mismatches {len(mismatched)}, largest point/normal/barycentric deviation {worst[0]:.3e}")
    assert total_contacts > 1000
    assert not mismatched, mismatched[:5]
    assert worst[1] <= TOL
    assert worst[0] <= 1e-12


LEARNED_ITERS = 3   # "3 XPBD iterations plus MLP" against the plain solver's 20


def test_c5_learned_step_on_the_block_per_step_from_shared_state():
    """C5: the learned step (network correction, multiplier start, the 3
    iterations of C5, fallback at 1.1x the plain residual) on the 56^3 tissue block
    with its sphere.  Every step: a plain 20-iteration device step from the
    same state gives the reference residual; the oracle starts from the
    device's state with the device's contacts (checked against its own
    detection) and takes the step the device took -- the warm step
    (orc_step_warm) with the same correction (the network's output through the
    same inference kernel), or the plain step where the device fell back."""
    import train_warmstart as TW
    mesh = W.block(56)
    sdf = W.block_sphere(56)
    kind, pos, quat, prm, vel = sdf
    gpu = X.Solver(mesh, dt=DT, iters=20, solver="gs", collision_every=1)
    plain = X.Solver(mesh, dt=DT, iters=20, solver="gs", collision_every=1)
    for s in (gpu, plain):
        s.set_sdf(kind, pos, quat, prm)
        s.set_sdf_velocity(vel)
    plain.set_residual_policy(True)
    net = TW.load(TW.DEFAULT_WEIGHTS)
    gpu.attach(net)
    o = _oracle(gpu, mesh, 20, "gs")
    pinned = mesh.pinned.astype(bool)
    errs, derrs, ratios, fallbacks = [], [], [], 0
    for n in range(30):
        x0, v0 = gpu.positions(), gpu.velocities()
        corr = net.forward(gpu.features()).astype(np.float64)
        corr[pinned] = 0.0
        plain.set_state(x0, v0)
        plain.step(1)
        ref = plain.last_step_residual()
        fb, r = gpu.step_learned(LEARNED_ITERS, ref_residual=ref, ratio=1.1)
        assert r <= 1.1 * ref * (1 + 1e-12), n
        ratios.append(r / ref)
        fallbacks += int(fb)
        face, cp, cn, cb = gpu.contacts()
        o.set_state(x0, v0)
        o.set_sdf(*_pose(sdf, n))
        o.detect()
        same, _ = _compare_detection(o, face, cp, cn, cb)
        assert same, n
        o.set_contacts(face, cp, cn, cb)
        if fb:
            o.step(1, detect=False)
        else:
            o.step_warm(corr, LEARNED_ITERS, detect=False)
        x, xr = gpu.positions(), o.x()
        errs.append(rel_err(x, xr))
        derrs.append(disp_err(x, xr, mesh.x0))
    errs = np.array(errs)
    print(f"C5 learned step on the block ({mesh.nt} tets), 30 steps: rel err vs oracle max {errs.max():.3e}, "
          f"displacement-relative max {max(derrs):.3e}; residual / plain 20-iteration residual median "
          f"{np.median(ratios):.2e}, max {max(ratios):.2e}; {fallbacks} fallbacks")
    assert errs.max() <= TOL
    # near rest both residuals are the FP32 rounding of the multipliers, where
    # the 10% margin decides either way; once the block moves the start wins
    assert fallbacks <= 15
