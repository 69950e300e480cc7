"""Independent float64 restatement of the element constraint maths with autograd.

Written from the constraint definitions alone (Stable Neo-Hookean split of
Macklin & Mueller: C_D = |F|_F, C_H = log J - gamma with the cubic Taylor
series of log J below J = 1), with gradients by automatic differentiation and
the rest-shape inverse by torch.linalg.inv, so it shares no code path with the
hand-derived gradients and adjugate inverse of the C oracle.
"""
import numpy as np
import torch

torch.set_default_dtype(torch.float64)


def rest(X):
    X = torch.as_tensor(np.asarray(X, np.float64)).reshape(4, 3)
    Dm = torch.stack([X[0] - X[3], X[1] - X[3], X[2] - X[3]], dim=1)
    return torch.linalg.inv(Dm), torch.abs(torch.linalg.det(Dm)) / 6.0


def constraints(x, Q, gamma):
    x = x.reshape(4, 3)
    D = torch.stack([x[0] - x[3], x[1] - x[3], x[2] - x[3]], dim=1)
    F = D @ Q
    CD = torch.sqrt((F * F).sum())
    J = torch.linalg.det(F)
    if J >= 1:
        CH = torch.log(J) - gamma
    else:
        a = J - 1
        CH = a - a ** 2 / 2 + a ** 3 / 3 - gamma
    return CD, CH


def grads(x12, Q, gamma):
    x = torch.tensor(np.asarray(x12, np.float64).reshape(12), requires_grad=True)
    Qt = torch.as_tensor(np.asarray(Q, np.float64))
    CD, CH = constraints(x, Qt, gamma)
    gD, = torch.autograd.grad(CD, x, retain_graph=True)
    gH, = torch.autograd.grad(CH, x)
    return float(CD.detach()), float(CH.detach()), gD.numpy(), gH.numpy()


def project(x12, w4, Q, aD, aH, gamma, dt, lam):
    """One XPBD 2x2 block step for the pair (C_D, C_H) of one element."""
    CD, CH, gD, gH = grads(x12, Q, gamma)
    W = np.repeat(np.asarray(w4, np.float64), 3)
    G = np.stack([gD, gH])                         # 2 x 12
    at = np.array([aD, aH]) / dt ** 2
    A = (G * W) @ G.T + np.diag(at)
    b = -np.array([CD, CH]) - at * np.asarray(lam, np.float64)
    dlam = np.linalg.solve(A, b)
    dx = W * (G.T @ dlam)
    return dx, np.asarray(lam, np.float64) + dlam
