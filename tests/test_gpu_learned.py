"""Learned warm start on the device: the MFMA network against the FP64 oracle,
gradients, training, features, the learned step and its fallback.

Tolerances: network outputs within 2e-3 (relative to the largest output) of
the oracle with the same bfloat16 operand rounding, 3e-2 of the unrounded FP64
network; gradients within 1e-2 relative L2 per parameter block of the rounded
oracle (5e-2 of FP64); features within 1e-5 relative per column group; the
learned step within the solver's 1e-5 position tolerance of the oracle's step
from the same corrected start; the fallback bit-identical to a plain step."""
import numpy as np
import pytest

import mlp_oracle as M
import orc
import workloads as W
import xpbd_amd as X
from helpers import rel_err

pytestmark = pytest.mark.gpu


def glorot(seed, bias=0.05):
    rng = np.random.default_rng(seed)
    P = {}
    for name, shp in M.SIZES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shp[0] + shp[1]))
            P[name] = rng.uniform(-a, a, shp)
        else:
            P[name] = rng.normal(0, bias, shp)
    return M.pack(P).astype(np.float32)


def scaled_inputs(seed, n):
    rng = np.random.default_rng(seed)
    mean = rng.normal(size=M.N_IN).astype(np.float32)
    std = rng.uniform(0.5, 3.0, M.N_IN).astype(np.float32)
    x = (mean + std * rng.normal(size=(n, M.N_IN))).astype(np.float32)
    return x, mean, std


def rel_max(a, ref):
    return float(np.max(np.abs(a - ref)) / max(np.max(np.abs(ref)), 1e-30))


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4097])
def test_network_forward_matches_oracle(n):
    p = glorot(1)
    x, mean, std = scaled_inputs(2, n)
    net = X.Network(p)
    net.set_scaling(mean, std, 2.5e-5)
    y = net.forward(x).astype(np.float64)
    ye = M.forward(p, x, mean, std, np.float32(2.5e-5), emulate_bf16=True)
    y64 = M.forward(p, x, mean, std, np.float32(2.5e-5))
    print(f"n={n}: vs bf16-rounded oracle {rel_max(y, ye):.2e}, vs FP64 {rel_max(y, y64):.2e}")
    assert np.all(np.isfinite(y))
    assert rel_max(y, ye) <= 2e-3
    assert rel_max(y, y64) <= 3e-2


def test_network_params_round_trip():
    p = glorot(3)
    net = X.Network(p)
    assert np.array_equal(net.params(), p)
    with pytest.raises(X.XpbdError):
        bad = p.copy()
        bad[7] = np.nan
        net.set_params(bad)


def _block_rel_l2(g, ref):
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    out = {}
    for i, (name, _) in enumerate(M.SIZES):
        a, b = g[offs[i]:offs[i + 1]], ref[offs[i]:offs[i + 1]]
        out[name] = (float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)),
                     float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)))
    return out


@pytest.mark.parametrize("n", [100, 3000])
def test_network_gradient_matches_oracle(n):
    """Forward + backward through the MFMA GEMM kernels (split-K for n = 3000)."""
    p = glorot(4)
    x, mean, std = scaled_inputs(5, n)
    t = np.random.default_rng(6).normal(size=(n, 3)).astype(np.float32)
    net = X.Network(p)
    net.set_scaling(mean, std, 1.5)
    loss, g = net.loss_grad(x, t)
    le, ge = M.loss_grad(p, x, t, mean, std, np.float32(1.5), emulate_bf16=True)
    l64, g64 = M.loss_grad(p, x, t, mean, std, np.float32(1.5))
    be, b64 = _block_rel_l2(g.astype(np.float64), ge), _block_rel_l2(g.astype(np.float64), g64)
    print(f"n={n} loss {loss:.6e} oracle {le:.6e}; per block (rel L2, cos) vs rounded {be}; vs FP64 {b64}")
    assert abs(loss - le) <= 1e-3 * le
    for name in be:
        assert be[name][0] <= 1e-2 and be[name][1] >= 0.9999, (name, be[name])
        assert b64[name][0] <= 5e-2, (name, b64[name])


def test_training_reduces_loss():
    rng = np.random.default_rng(7)
    n = 20000
    x = rng.normal(size=(n, M.N_IN)).astype(np.float32)
    B = rng.normal(size=(M.N_IN, 3)) / np.sqrt(M.N_IN)
    t = np.tanh(x @ B).astype(np.float32)
    net = X.Network()
    net.fit_scaling(x, t)
    loss0, _ = net.loss_grad(x, t)
    loss1 = net.train(x, t, batch=1024, epochs=15, lr=2e-3, seed=1)
    loss_all, _ = net.loss_grad(x, t)
    print(f"training loss {loss0:.4e} -> last batch {loss1:.4e}, full set {loss_all:.4e}")
    assert loss_all < 0.05 * loss0


def _mesh_state(seed=7, n=6, steps=4):
    mesh = W.cube(n, seed=seed)
    gpu = X.Solver(mesh, dt=1e-3, iters=5, solver="gs")
    gpu.step(steps)
    return mesh, gpu


def test_features_match_oracle():
    mesh, gpu = _mesh_state()
    net = X.Network()
    gpu.attach(net)
    f = gpu.features().astype(np.float64)
    x, v = gpu.positions(), gpu.velocities()
    fe = M.features(x, v, np.zeros_like(x), mesh.x0, mesh.tets, mesh.pinned, mesh.E, mesh.nu, 1e-3, 9.81)
    for a, b in [(0, 3), (3, 6), (6, 9), (9, 12), (12, 15), (15, 18), (18, 21), (21, 24)]:
        err = np.max(np.abs(f[:, a:b] - fe[:, a:b])) / max(np.max(np.abs(fe[:, a:b])), 1e-12)
        print(f"features {a}-{b - 1}: rel err {err:.2e}")
        assert err <= 1e-5 or np.max(np.abs(fe[:, a:b])) == 0.0 and np.max(np.abs(f[:, a:b])) == 0.0
    # the prediction inside features() is undone: stepping continues unchanged
    ref = X.Solver(mesh, dt=1e-3, iters=5, solver="gs")
    ref.step(5)
    gpu.step(1)
    assert np.array_equal(gpu.positions(), ref.positions())


@pytest.mark.parametrize("solver,variant,start", [("gs", "cluster", "positions"), ("jacobi", "patch", "positions"),
                                                  ("gs", "cluster", "multipliers"), ("gs", "quad", "multipliers"),
                                                  ("jacobi", "patch", "multipliers"), ("gs", "quad", "consistent"),
                                                  ("gs", "cluster", "consistent"), ("jacobi", "patch", "consistent")])
def test_learned_step_matches_oracle_from_the_same_corrected_start(solver, variant, start):
    """The device's learned step against the oracle's from the same network
    correction: 'positions' starts every multiplier at zero (orc_step_corrected),
    'multipliers' at -C / alpha~ of the corrected positions (orc_step_warm),
    'consistent' at those multipliers with the positions moved by
    M^-1 grad C^T lambda0 from the prediction (orc_step_learned), whose primary
    residual (momentum balance) is compared too."""
    mesh = W.cube(6, seed=8, E=3e3 if solver == "gs" else 1e3)
    iters = 10
    gpu = X.Solver(mesh, dt=1e-3, iters=iters, solver=solver, variant=variant)
    o = orc.Sim(mesh, dt=1e-3, iters=iters, mode="gs" if solver == "gs" else "jacobi")
    if solver == "gs":
        o.set_order(gpu.tet_order())
    net = X.Network(glorot(9))
    net.set_scaling(np.zeros(24), np.full(24, 1e-3), 2e-6)
    gpu.attach(net)
    gpu.set_learned_start(start)
    if start == "consistent":
        gpu.set_primary_residual(True)
        o.set_primary_residual(True)
    for _ in range(3):
        gpu.step(1)
        o.step(1)
    for k in range(3):
        corr = net.forward(gpu.features()).astype(np.float64)
        corr[mesh.pinned.astype(bool)] = 0.0
        assert 1e-8 < np.max(np.abs(corr)) < 1e-3
        fb, r = gpu.step_learned(3, ref_residual=0.0)
        assert not fb
        if start == "positions":
            o.step_corrected(corr, 3)
        elif start == "consistent":
            o.step_learned(corr, 3)
        else:
            o.step_warm(corr, 3)
        err = rel_err(gpu.positions(), o.x())
        ro = o.last_step_residual()
        print(f"{solver} {start} learned step {k}: rel err {err:.2e}, residual {r:.3e} (oracle {ro:.3e})")
        assert err <= 1e-5
        if start == "consistent":
            pg, po = gpu.last_step_primary_residual(), o.last_primary_residual()
            print(f"  primary residual {pg:.4e} (oracle {po:.4e})")
            assert abs(pg - po) <= 1e-2 * po
            assert abs(r - ro) <= 1e-2 * ro + 1e-9
        elif start == "positions":
            assert abs(r - ro) <= 1e-3 * ro + 1e-9
        else:
            # both start at zero elastic residual; what remains is rounding of the
            # FP32 state and multipliers (device) or FP64 (oracle)
            assert r <= 1e-4 and ro <= 1e-4


@pytest.mark.parametrize("solver,variant", [("gs", "cluster"), ("jacobi", "patch")])
def test_fallback_reproduces_the_plain_step_exactly(solver, variant):
    """A corrupted network (huge weights) must trigger the fallback, which redoes
    the step as the plain solver: positions bit-identical to a plain run."""
    mesh = W.cube(6, seed=10, E=3e3 if solver == "gs" else 1e3)
    sph = W.block_sphere(6, r=0.02)
    mk = lambda: X.Solver(mesh, dt=1e-3, iters=10, solver=solver, variant=variant)  # noqa: E731
    plain, learned = mk(), mk()
    for s in (plain, learned):
        s.set_sdf(sph[0], (0.03, 0.03, 0.06 + 0.02 - 0.0005), sph[2], sph[3])
        s.set_sdf_velocity(sph[4])
        s.set_residual_policy(True)
    p = glorot(11) * 1e3
    net = X.Network(p)
    net.set_scaling(np.zeros(24), np.full(24, 1e-6), 1.0)
    learned.attach(net)
    # with multipliers from zero the residual tells a bad start apart
    learned.set_learned_start("positions")
    for k in range(5):
        plain.step(1)
        ref = plain.last_step_residual()
        fb, r = learned.step_learned(2, ref_residual=ref, ratio=1.1)
        assert fb, k
        assert np.array_equal(learned.positions(), plain.positions()), k
        assert np.array_equal(learned.velocities(), plain.velocities()), k
        assert r == pytest.approx(ref, rel=1e-12)
    st = learned.learned_stats()
    assert st["steps"] == 5 and st["fallbacks"] == 5


@pytest.mark.parametrize("solver,variant", [("gs", "quad"), ("jacobi", "patch")])
def test_captured_learned_step_equals_the_eager_one(solver, variant):
    """The learned step replays a captured graph; in kernel timing mode it
    launches eagerly.  Both give the same bits, step after step, through
    fallbacks forced on every third step, detection steps, and a change of the
    network's output scale after the graphs were captured."""
    import train_warmstart as TW
    if solver == "gs":
        mesh, sph = _held_out_block(seed=79, n=6)
    else:  # the reference's plain-sum Jacobi is stable on soft material only
        mesh = W.cube(6, seed=10, E=1e3)
        b = W.block_sphere(6, r=0.02)
        sph = (b[0], (0.03, 0.03, 0.06 + 0.02 - 0.0005), b[2], b[3], b[4])
    graph, eager = (X.Solver(mesh, dt=1e-3, iters=6, solver=solver, variant=variant) for _ in range(2))
    net = TW.load(TW.DEFAULT_WEIGHTS)
    for s in (graph, eager):
        s.set_sdf(*sph[:4])
        s.set_sdf_velocity(sph[4])
        s.set_residual_policy(True)
        s.attach(net)
    eager.set_kernel_timing(True)
    z = np.load(TW.DEFAULT_WEIGHTS, allow_pickle=False)
    fbs = 0
    for k in range(12):
        if k == 6:
            net.set_scaling(z["mean"], z["std"], 0.5 * float(z["scale"]))
        ref = 1e-30 if k % 3 == 2 else 1e30
        fg, rg = graph.step_learned(3, ref_residual=ref, ratio=1.1)
        fe, re = eager.step_learned(3, ref_residual=ref, ratio=1.1)
        assert fg == fe == (k % 3 == 2), k
        assert rg == re, k
        assert np.array_equal(graph.positions(), eager.positions()), k
        assert np.array_equal(graph.velocities(), eager.velocities()), k
        fbs += fg
    assert graph.learned_stats()["fallbacks"] == fbs == 4
    assert eager.learned_stats()["mlp_calls"] == 12


@pytest.mark.parametrize("start", ["multipliers", "consistent"])
def test_first_order_learned_step_matches_oracle(start):
    """First-order XPBD (damping inertia, alpha~ = alpha / dt, the damped
    gravity move) under the learned step's warm starts, against the oracle's
    from the same correction."""
    rng = np.random.default_rng(41)
    n = 6
    E, nu = W.heterogeneous(6 * n ** 3, rng, E_range=(1e3, 5e4))
    mesh = W.grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=0.0, rng=rng)
    gpu = X.Solver(mesh, dt=1e-3, iters=10, solver="gs", variant="quad", first_order=True, damping_multiplier=20.0,
                   adjust_damping=True)
    o = orc.Sim(mesh, dt=1e-3, iters=10, mode="gs")
    o.set_first_order(20.0, True)
    o.set_order(gpu.tet_order())
    net = X.Network(glorot(13))
    net.set_scaling(np.zeros(24), np.full(24, 1e-3), 2e-6)
    gpu.attach(net)
    gpu.set_learned_start(start)
    for _ in range(3):
        gpu.step(1)
        o.step(1)
    for k in range(3):
        corr = net.forward(gpu.features()).astype(np.float64)
        corr[mesh.pinned.astype(bool)] = 0.0
        fb, r = gpu.step_learned(3, ref_residual=0.0)
        assert not fb
        if start == "consistent":
            o.step_learned(corr, 3)
        else:
            o.step_warm(corr, 3)
        err = rel_err(gpu.positions(), o.x())
        print(f"first order {start} learned step {k}: rel err {err:.2e}, residual {r:.3e} "
              f"(oracle {o.last_step_residual():.3e})")
        assert err <= 1e-5


def test_switching_learned_start_modes_replays_the_right_graph():
    """One captured graph per start mode: a handle that switches modes from
    step to step (positions, multipliers, consistent, in a cycle that returns
    to each) takes the same bits as the same sequence launched eagerly."""
    import train_warmstart as TW
    mesh, sph = _held_out_block(seed=81, n=6)
    graph, eager = (X.Solver(mesh, dt=1e-3, iters=6, solver="gs", variant="quad") for _ in range(2))
    net = TW.load(TW.DEFAULT_WEIGHTS)
    for s in (graph, eager):
        s.set_sdf(*sph[:4])
        s.set_sdf_velocity(sph[4])
        s.attach(net)
    eager.set_kernel_timing(True)
    modes = ["positions", "multipliers", "consistent", "multipliers", "positions", "consistent", "consistent",
             "positions", "multipliers"]
    for k, mode in enumerate(modes):
        for s in (graph, eager):
            s.set_learned_start(mode)
        fg, rg = graph.step_learned(3, ref_residual=0.0)
        fe, re = eager.step_learned(3, ref_residual=0.0)
        assert rg == re, (k, mode)
        assert np.array_equal(graph.positions(), eager.positions()), (k, mode)
        assert np.array_equal(graph.velocities(), eager.velocities()), (k, mode)


def test_learned_step_requires_a_network():
    mesh = W.cube(4, seed=1)
    gpu = X.Solver(mesh, dt=1e-3, iters=5, solver="gs")
    with pytest.raises(X.XpbdError):
        gpu.step_learned(3)
    net = X.Network()
    gpu.attach(net)
    gpu.attach(None)
    with pytest.raises(X.XpbdError):
        gpu.step_learned(3)


def _trajectory(seed, n=8, steps=40):
    """A small heterogeneous block pressed by a moving sphere (training family)."""
    rng = np.random.default_rng(seed)
    E, nu = W.heterogeneous(6 * n ** 3, rng)
    mesh = W.grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=0.01, faces=True, rng=rng)
    gpu = X.Solver(mesh, dt=1e-3, iters=20, solver="gs")
    L = 0.01 * n
    r = rng.uniform(0.015, 0.03)
    c = rng.uniform(0.3 * L, 0.7 * L, 2)
    gpu.set_sdf(X.SDF_SPHERE, (c[0], c[1], L + r - 0.0005), (0, 0, 0, 1), (r, 0, 0))
    gpu.set_sdf_velocity((0.0, 0.0, -rng.uniform(0.05, 0.2)))
    net = X.Network()
    gpu.attach(net)
    return gpu.collect_training(steps)


def test_training_learns_the_solver_correction():
    """Training on the device learns the solver's correction: on a held-out
    trajectory the network's c_hat is closer to the plain solver's correction
    than the inertial prediction (c_hat = 0) or the previous step's correction."""
    xs, ys = zip(*[_trajectory(s) for s in (21, 22, 23, 24)])
    x, y = np.concatenate(xs), np.concatenate(ys)
    xe, ye = _trajectory(99)
    net = X.Network()
    net.fit_scaling(x, y)
    net.train(x, y, batch=2048, epochs=40, lr=2e-3, seed=3)
    pred = net.forward(xe)
    err_net = np.linalg.norm(pred - ye) / np.linalg.norm(ye)
    err_prev = np.linalg.norm(xe[:, 15:18] - ye) / np.linalg.norm(ye)
    print(f"warm-start error / |correction|: network {err_net:.3f}, previous correction {err_prev:.3f}, "
          f"none 1.000 ({len(x)} training rows, {len(xe)} held-out rows)")
    assert err_net < 0.5
    assert err_net < err_prev


def test_network_file_round_trip(tmp_path):
    """xpbd_mlp_save / xpbd_mlp_load keep parameters and scaling exactly; a
    truncated file or a network without scaling is refused."""
    rng = np.random.default_rng(21)
    a = X.Network(glorot(5))
    mean, std = rng.normal(size=X.MLP_IN).astype(np.float32), rng.uniform(0.5, 2.0, X.MLP_IN).astype(np.float32)
    a.set_scaling(mean, std, 3e-4)
    path = str(tmp_path / "net.bin")
    a.save(path)
    b = X.Network()
    b.load(path)
    assert np.array_equal(a.params(), b.params())
    x = rng.normal(size=(1000, X.MLP_IN)).astype(np.float32) * std + mean
    assert np.array_equal(a.forward(x), b.forward(x))
    raw = open(path, "rb").read()
    assert len(raw) == 8 + 8 + 4 * (X.MLP_PARAMS + 2 * X.MLP_IN + 1)
    bad = tmp_path / "short.bin"
    bad.write_bytes(raw[:-3])
    with pytest.raises(X.XpbdError):
        X.Network().load(str(bad))
    with pytest.raises(X.XpbdError):
        X.Network(glorot(6)).save(str(tmp_path / "unscaled.bin"))


def test_learned_step_on_a_batch_falls_back_to_the_plain_batch_step():
    """The learned step drives a batched handle as one object (features over the
    concatenated vertices, one network pass, one residual over all elements);
    its fallback reproduces the plain batched step bit for bit."""
    meshes = [W.cube(4, seed=40), W.cube(5, seed=41), W.cube(3, seed=42)]
    mk = lambda: X.Solver(meshes, dt=1e-3, iters=8, solver="gs", variant="quad")  # noqa: E731
    plain, learned = mk(), mk()
    for s in (plain, learned):
        s.set_residual_policy(True)
    net = X.Network(glorot(12) * 1e3)
    net.set_scaling(np.zeros(24), np.full(24, 1e-6), 1.0)
    learned.attach(net)
    learned.set_learned_start("positions")
    for k in range(4):
        plain.step(1)
        fb, r = learned.step_learned(2, ref_residual=plain.last_step_residual(), ratio=1.1)
        assert fb, k
        for i in range(len(meshes)):
            assert np.array_equal(learned.instance_positions(i), plain.instance_positions(i)), (k, i)
    # a sane network on the same batch: finite positions, residual reported
    good = X.Network(glorot(13))
    good.set_scaling(np.zeros(24), np.ones(24), 1e-6)
    learned.attach(good)
    fb, r = learned.step_learned(8, ref_residual=1e30, ratio=1.1)
    assert not fb and np.isfinite(r)
    assert all(np.all(np.isfinite(learned.instance_positions(i))) for i in range(len(meshes)))


def _held_out_block(seed=77, n=10):
    rng = np.random.default_rng(seed)
    E, nu = W.heterogeneous(6 * n ** 3, rng)
    mesh = W.grid_object(n, n, n, 0.01, E=E, nu=nu, pin_axis=2, noise=0.01, faces=True, rng=rng)
    L = 0.01 * n
    sph = (X.SDF_SPHERE, (0.45 * L, 0.55 * L, L + 0.02 - 0.001), (0, 0, 0, 1), (0.02, 0, 0), (0, 0, -0.15))
    return mesh, sph


def _solver(mesh, sph, iters=20):
    s = X.Solver(mesh, dt=1e-3, iters=iters, solver="gs")
    s.set_sdf(*sph[:4])
    s.set_sdf_velocity(sph[4])
    s.set_residual_policy(True)
    return s


def test_multiplier_start_reaches_the_plain_residual():
    """From the same state, the learned step with the multiplier start and one
    iteration ends at or below the residual of the plain 20-iteration step,
    step after step (a block pressed by a sphere, contacts every step)."""
    import train_warmstart as TW
    mesh, sph = _held_out_block()
    plain, learned = _solver(mesh, sph), _solver(mesh, sph)
    learned.attach(TW.load(TW.DEFAULT_WEIGHTS))
    ratios = []
    for k in range(30):
        x0, v0 = plain.positions(), plain.velocities()
        plain.step(1)
        learned.set_state(x0, v0)
        fb, r = learned.step_learned(1, ref_residual=plain.last_step_residual(), ratio=1.1)
        assert not fb, k
        ratios.append(r / plain.last_step_residual())
    print(f"multiplier start, 1 iteration: residual / plain 20-iteration residual: median {np.median(ratios):.2e}, "
          f"max {max(ratios):.2e}")
    assert max(ratios) <= 1.0


def test_trained_network_gains_trajectory_accuracy():
    """What the shipped network gains.  With the multiplier start the residual
    test is met whatever the correction (see the test above), so the network's
    contribution is the trajectory: on a held-out block, 3 learned iterations
    per step with the network's correction end closer to the plain
    20-iteration trajectory than the same learned step without a correction
    (c_hat = 0: inertial prediction plus multiplier start)."""
    import train_warmstart as TW
    mesh, sph = _held_out_block(seed=78)
    plain = _solver(mesh, sph)
    runs = {"network": TW.load(TW.DEFAULT_WEIGHTS), "no correction": X.Network(np.zeros(X.MLP_PARAMS, np.float32))}
    runs["no correction"].set_scaling(np.zeros(24), np.ones(24), 1.0)
    solvers = {}
    for name, net in runs.items():
        s = _solver(mesh, sph)
        s.attach(net)
        solvers[name] = s
    errs = {name: [] for name in runs}
    for _ in range(40):
        plain.step(1)
        xp = plain.positions()
        for name, s in solvers.items():
            s.step_learned(3, ref_residual=0.0)
            errs[name].append(float(np.max(np.abs(s.positions() - xp)) / np.max(np.abs(xp - mesh.x0))))
    for name in runs:
        e = errs[name]
        print(f"{name}: trajectory error vs plain (max |dx| / max displacement) after 10 / 20 / 40 steps: "
              f"{e[9]:.3e} / {e[19]:.3e} / {e[39]:.3e}")
    assert errs["network"][-1] < errs["no correction"][-1]
