"""Pinning of the FP64 network/feature oracle (CPU only).

The network oracle is checked against an independent statement of the same
network in torch (float64 autograd) and against central finite differences;
its bfloat16 rounding against torch's float32 -> bfloat16 conversion; the
feature oracle against values computed by hand on small meshes."""
import numpy as np
import pytest
import torch

import mlp_oracle as M
import xpbd_amd as X


def _torch_net(p):
    P = M.unpack(p)
    return [torch.tensor(P[k], dtype=torch.float64, requires_grad=True) for k, _ in M.SIZES]


def _torch_loss(ts, x, t, mean, std, scale):
    W1, b1, W2, b2, W3, b3 = ts
    xs = (torch.tensor(x) - torch.tensor(mean)) / torch.tensor(std)
    h1 = torch.nn.functional.silu(xs @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    y = scale * (h2 @ W3.T + b3)
    return torch.mean(((y - torch.tensor(t)) / scale) ** 2), y


def _problem(seed=0, n=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.15, M.N_PARAMS)
    x = rng.normal(size=(n, M.N_IN)) * 3 + 1
    t = rng.normal(size=(n, M.N_OUT))
    mean = rng.normal(size=M.N_IN)
    std = rng.uniform(0.5, 4, M.N_IN)
    return p, x, t, mean, std, 1.7


def test_param_count():
    assert M.N_PARAMS == 20099


def test_forward_and_gradient_match_torch_autograd():
    p, x, t, mean, std, sc = _problem()
    ts = _torch_net(p)
    loss_t, y_t = _torch_loss(ts, x, t, mean, std, sc)
    loss_t.backward()
    g_t = np.concatenate([a.grad.numpy().reshape(-1) for a in ts])
    loss, g = M.loss_grad(p, x, t, mean, std, sc)
    y = M.forward(p, x, mean, std, sc)
    assert np.allclose(y, y_t.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(loss - loss_t.item()) <= 1e-12 * loss
    assert np.max(np.abs(g - g_t)) <= 1e-11 * np.max(np.abs(g_t))


def test_gradient_matches_central_differences():
    p, x, t, mean, std, sc = _problem(1, n=9)
    _, g = M.loss_grad(p, x, t, mean, std, sc)
    rng = np.random.default_rng(2)
    # a sample from every parameter block
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    idx = np.concatenate([rng.choice(np.arange(offs[i], offs[i + 1]), min(8, offs[i + 1] - offs[i]), replace=False)
                          for i in range(len(M.SIZES))])
    eps = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        lp = M.loss_grad(pp, x, t, mean, std, sc)[0]
        pp[i] -= 2 * eps
        lm = M.loss_grad(pp, x, t, mean, std, sc)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * np.max(np.abs(g)) + 1e-7 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-8, 8, 4000),
                        [0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 2 ** -8 + 2 ** -9, 1 + 3 * 2 ** -8, -(1 + 2 ** -8)]])
    ref = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(a), ref)


def test_bf16_emulation_changes_results_by_bf16_amounts():
    p, x, t, mean, std, sc = _problem(4)
    y = M.forward(p, x, mean, std, sc)
    yb = M.forward(p, x, mean, std, sc, emulate_bf16=True)
    d = np.max(np.abs(y - yb)) / np.max(np.abs(y))
    assert 1e-4 < d < 5e-2


def test_one_ring_of_single_tet_and_shared_face():
    off, nb = M.one_ring(5, np.array([[0, 1, 2, 3], [1, 2, 3, 4]]))
    ring = [list(nb[off[i]:off[i + 1]]) for i in range(5)]
    assert ring == [[1, 2, 3], [0, 2, 3, 4], [0, 1, 3, 4], [0, 1, 2, 4], [1, 2, 3]]


def test_features_by_hand_on_two_tets():
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], float)
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    rng = np.random.default_rng(5)
    u = rng.normal(size=(5, 3))
    v = rng.normal(size=(5, 3))
    c = rng.normal(size=(5, 3))
    E = np.array([1e3, 1e5])
    nu = np.array([0.45, 0.49])
    pinned = np.array([1, 0, 0, 0, 0], np.uint8)
    dt, g = 1e-3, 9.81
    f = M.features(X + u, v, c, X, tets, pinned, E, nu, dt, g)
    assert f.shape == (5, 24)
    p = dt * v + np.array([0, 0, -g * dt * dt])
    assert np.allclose(f[:, 0:3], p)
    assert np.allclose(f[:, 3:6], v)
    # vertex 0 neighbours 1,2,3; vertex 4 neighbours 1,2,3; vertex 1 neighbours 0,2,3,4
    assert np.allclose(f[0, 6:9], u[[1, 2, 3]].mean(0) - u[0])
    assert np.allclose(f[1, 9:12], p[[0, 2, 3, 4]].mean(0) - p[1])
    assert np.allclose(f[4, 18:21], c[[1, 2, 3]].mean(0) - c[4])
    assert np.allclose(f[:, 12:15], u)
    assert np.allclose(f[:, 15:18], c)
    assert np.allclose(f[:, 21], [3, 4, 4, 4, 5])
    assert np.allclose(f[:, 22], [0.45, 0.47, 0.47, 0.47, 0.49])
    assert np.allclose(f[:, 23], [0, 0.25, 1 / 4, 1 / 4, 0])


def test_ring_laplacian_of_affine_field_on_grid_interior():
    # for u = A X the ring feature is A (mean of neighbour rest positions - X_i): it vanishes
    # at interior vertices of the 6-tet grid, whose 1-ring is point-symmetric
    x, tets = X.grid_mesh(3, 3, 3, 0.1)
    A = np.array([[0.1, 0.2, -0.3], [0.0, 0.5, 0.1], [0.2, -0.1, 0.3]])
    u = x @ A.T
    f = M.features(x + u, np.zeros_like(x), np.zeros_like(x), x, tets, None, np.ones(len(tets)),
                   np.full(len(tets), 0.3), 1e-3, 0.0)
    interior = np.all((x > 0.05) & (x < 0.25), axis=1)
    assert interior.sum() == 8
    assert np.max(np.abs(f[interior, 6:9])) < 1e-12
    assert np.max(np.abs(f[~interior, 6:9])) > 1e-3
