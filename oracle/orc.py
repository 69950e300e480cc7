"""ctypes wrapper of liboracle.so, the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline legs of bench.py.  It is the checker, never the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `make -C {_HERE}`")
lib = C.CDLL(LIB_PATH)

GS, JACOBI = 0, 1
SDF_NONE, SDF_SPHERE, SDF_BOX = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


def _sig(name, args, res=None):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("orc_lame", [C.c_double, C.c_double, _dp, _dp])
_sig("orc_rest_tet", [_dp, _dp, _dp, _dp, _dp, _dp], C.c_int)
_sig("orc_eval_tet", [_dp, _dp, C.c_double, _dp, _dp])
_sig("orc_deformation_gradient", [_dp, _dp, _dp])
_sig("orc_project_tet", [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp])
_sig("orc_sdf_eval", [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp])
_sig("orc_frank_wolfe", [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp])
_sig("orc_barycentric", [_dp, _dp, _dp, _dp, _dp])
_sig("orc_create", [C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp, _u8p, C.c_double, C.c_double, C.c_int,
                    C.c_int], C.c_void_p)
_sig("orc_destroy", [C.c_void_p])
_sig("orc_set_order", [C.c_void_p, _ip])
_sig("orc_set_mode", [C.c_void_p, C.c_int, C.c_int])
_sig("orc_set_faces", [C.c_void_p, C.c_int, _ip])
_sig("orc_set_sdf", [C.c_void_p, C.c_int, _dp, _dp, _dp])
_sig("orc_set_state", [C.c_void_p, _dp, _dp])
_sig("orc_detect", [C.c_void_p], C.c_int)
_sig("orc_num_contacts", [C.c_void_p], C.c_int)
_sig("orc_get_contacts", [C.c_void_p, _ip, _dp, _dp, _dp])
_sig("orc_step", [C.c_void_p, C.c_int])
_sig("orc_predict", [C.c_void_p])
_sig("orc_get_x", [C.c_void_p, _dp])
_sig("orc_get_xprev", [C.c_void_p, _dp])
_sig("orc_get_v", [C.c_void_p, _dp])
_sig("orc_get_lambda", [C.c_void_p, _dp])
_sig("orc_get_masses", [C.c_void_p, _dp])
_sig("orc_get_rest", [C.c_void_p, _dp, _dp])
_sig("orc_get_compliance", [C.c_void_p, _dp, _dp, _dp])
_sig("orc_residual", [C.c_void_p, _dp], C.c_double)
_sig("orc_energy", [C.c_void_p], C.c_double)
_sig("orc_set_contact_order", [C.c_void_p, _ip])
_sig("orc_last_step_residual", [C.c_void_p], C.c_double)
_sig("orc_project_contact", [_dp, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int)
_sig("orc_step_corrected", [C.c_void_p, C.c_int, _dp, C.c_int])
_sig("orc_step_dual_warm", [C.c_void_p, C.c_int, _dp, C.c_int])
_sig("orc_step_learned", [C.c_void_p, C.c_int, _dp, C.c_int])
_sig("orc_set_first_order", [C.c_void_p, C.c_double, C.c_int])
_sig("orc_get_inverse_inertia", [C.c_void_p, _dp])
_sig("orc_set_attachments", [C.c_void_p, C.c_int, _ip, _dp])
_sig("orc_set_local_collision_iters", [C.c_void_p, C.c_int], None)
_sig("orc_last_local_list_size", [C.c_void_p], C.c_int)
_sig("orc_local_list_size", [C.c_void_p], C.c_int)
_sig("orc_get_contact_lambdas", [C.c_void_p, _dp], None)
_sig("orc_color_elements", [C.c_void_p, C.c_int], C.c_int)
_sig("orc_set_threads", [C.c_void_p, C.c_int], None)
_sig("orc_get_colors", [C.c_void_p, _ip, _ip, _ip], C.c_int)
_sig("orc_set_order_levels", [C.c_void_p, _ip, C.c_int, _ip], C.c_int)
_sig("orc_set_contacts", [C.c_void_p, C.c_int, _ip, _dp, _dp, _dp])
_sig("orc_momentum_residual", [C.c_void_p, _dp, _dp], C.c_double)
_sig("orc_set_primary_residual", [C.c_void_p, C.c_int], None)
_sig("orc_last_primary_residual", [C.c_void_p, _dp], C.c_double)
_sig("orc_last_prediction", [C.c_void_p, _dp])
_sig("orc_step_warm", [C.c_void_p, C.c_int, _dp, C.c_int])
_sig("orc_solve_elements", [C.c_void_p, C.c_int, _ip, C.c_int, C.c_int, C.c_int])
_sig("orc_set_lambdas", [C.c_void_p, _dp])


def project_contact(x9, w3, cp, cn, bary, lam=0.0):
    lam_c = C.c_double(lam)
    dx = np.zeros(9)
    active = lib.orc_project_contact(p(_d(x9).reshape(9)), p(_d(w3)), p(_d(cp)), p(_d(cn)), p(_d(bary)),
                                     C.byref(lam_c), p(dx))
    return bool(active), dx, lam_c.value


def p(a, ct=C.c_double):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    return a.ctypes.data_as(C.POINTER(ct))


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def lame(E, nu):
    mu, lam = C.c_double(), C.c_double()
    lib.orc_lame(E, nu, C.byref(mu), C.byref(lam))
    return mu.value, lam.value


def rest_tet(X):
    X = _d(X).reshape(4, 3)
    Q = np.empty(9)
    vol = C.c_double()
    rc = lib.orc_rest_tet(p(X[0].copy()), p(X[1].copy()), p(X[2].copy()), p(X[3].copy()), p(Q), C.byref(vol))
    return Q.reshape(3, 3).T.copy(), vol.value, rc  # Q as a 3x3 (row, col) matrix


def _qcm(Q):
    """3x3 matrix -> column-major 9 vector."""
    return _d(np.asarray(Q).T.reshape(9))


def eval_tet(x12, Q, gamma):
    C2 = np.empty(2)
    g = np.empty(24)
    lib.orc_eval_tet(p(_d(x12).reshape(12)), p(_qcm(Q)), gamma, p(C2), p(g))
    return C2, g


def deformation_gradient(x12, Q):
    F = np.empty(9)
    lib.orc_deformation_gradient(p(_d(x12).reshape(12)), p(_qcm(Q)), p(F))
    return F.reshape(3, 3).T.copy()


def project_tet(x12, w4, Q, aD, aH, gamma, dt, lam2):
    lam = _d(lam2).copy()
    dx = np.empty(12)
    lib.orc_project_tet(p(_d(x12).reshape(12)), p(_d(w4)), p(_qcm(Q)), aD, aH, gamma, dt, p(lam), p(dx))
    return dx, lam


def sdf_eval(kind, pos, quat, prm, x):
    d = C.c_double()
    g = np.empty(3)
    lib.orc_sdf_eval(kind, p(_d(pos)), p(_d(quat)), p(_d(prm)), p(_d(x)), C.byref(d), p(g))
    return d.value, g


def frank_wolfe(kind, pos, quat, prm, p1, p2, p3):
    x = np.empty(3)
    lib.orc_frank_wolfe(kind, p(_d(pos)), p(_d(quat)), p(_d(prm)), p(_d(p1)), p(_d(p2)), p(_d(p3)), p(x))
    return x


def barycentric(pt, a, b, c):
    out = np.empty(3)
    lib.orc_barycentric(p(_d(pt)), p(_d(a)), p(_d(b)), p(_d(c)), p(out))
    return out


class Sim:
    """FP64 reference-semantics simulation of one object."""

    def __init__(self, mesh, dt=1e-3, g=9.81, iters=10, mode="gs"):
        self.nv, self.nt = len(mesh.x0), len(mesh.tets)
        self._x0 = _d(mesh.x0)
        self._v0 = None if mesh.v0 is None else _d(mesh.v0)
        self._t = np.ascontiguousarray(mesh.tets, np.int32)
        self._E, self._nu, self._rho = _d(mesh.E), _d(mesh.nu), _d(mesh.rho)
        self._pin = None if mesh.pinned is None else np.ascontiguousarray(mesh.pinned, np.uint8)
        self.s = lib.orc_create(self.nv, self.nt, p(self._x0), p(self._v0), p(self._t, C.c_int32), p(self._E),
                                p(self._nu), p(self._rho), p(self._pin, C.c_uint8), dt, g, iters,
                                GS if mode == "gs" else JACOBI)
        self.nf = 0
        if mesh.faces is not None and len(mesh.faces):
            self._f = np.ascontiguousarray(mesh.faces, np.int32)
            self.nf = len(self._f)
            lib.orc_set_faces(self.s, self.nf, p(self._f, C.c_int32))

    def __del__(self):
        if getattr(self, "s", None):
            lib.orc_destroy(self.s)
            self.s = None

    def set_order(self, order):
        o = np.ascontiguousarray(order, np.int32)
        lib.orc_set_order(self.s, p(o, C.c_int32))

    def set_order_levels(self, order, level_off):
        """Gauss-Seidel order cut into levels (the colours of a colour-ordered
        solver): with set_threads > 1 the connected units of a level run on
        separate threads, bit for bit the sequential pass over `order`.
        Returns the number of units."""
        o = np.ascontiguousarray(order, np.int32)
        lo = np.ascontiguousarray(level_off, np.int32)
        n = lib.orc_set_order_levels(self.s, p(o, C.c_int32), len(lo) - 1, p(lo, C.c_int32))
        if n < 0:
            raise ValueError("bad level offsets")
        self._ncolor = len(lo) - 1
        return n

    def set_contacts(self, face, cp, cn, bary):
        """Replace the contact set with the given contacts (multipliers 0)."""
        f = np.ascontiguousarray(face, np.int32)
        self._cset = (f, _d(cp).reshape(-1), _d(cn).reshape(-1), _d(bary).reshape(-1))
        lib.orc_set_contacts(self.s, len(f), p(f, C.c_int32), p(self._cset[1]), p(self._cset[2]), p(self._cset[3]))

    def color_elements(self, cluster=True):
        """Greedy colouring of units (runs of consecutive elements over at most 8
        vertices when cluster is set, single elements otherwise); the GS order
        becomes colour-major.  Returns the number of colours (-1 if more than 64
        would be needed)."""
        self._ncolor = lib.orc_color_elements(self.s, int(bool(cluster)))
        return self._ncolor

    def set_threads(self, n):
        """OpenMP threads of the colour-parallel GS / parallel Jacobi projection."""
        lib.orc_set_threads(self.s, int(n))

    def colors(self):
        """(colour-major order, unit offsets into it, colour offsets into the
        units) after color_elements."""
        nu = lib.orc_get_colors(self.s, None, None, None)
        order = np.empty(self.nt, np.int32)
        uoff = np.empty(nu + 1, np.int32)
        coff = np.empty(self._ncolor + 1, np.int32)
        lib.orc_get_colors(self.s, p(order, C.c_int32), p(uoff, C.c_int32), p(coff, C.c_int32))
        return order, uoff, coff

    def set_contact_order(self, face_rank):
        r = np.ascontiguousarray(face_rank, np.int32)
        lib.orc_set_contact_order(self.s, p(r, C.c_int32))

    def set_mode(self, mode, iters):
        lib.orc_set_mode(self.s, GS if mode == "gs" else JACOBI, iters)

    def set_first_order(self, damping_multiplier=1.0, adjust_to_material=False):
        """Switch to first-order XPBD (damping B as constraint inertia, alpha / dt)."""
        lib.orc_set_first_order(self.s, float(damping_multiplier), int(adjust_to_material))

    def set_attachments(self, verts, targets):
        """Attach vertices verts[i] to the points targets[i] (world coordinates)."""
        v = np.ascontiguousarray(verts, np.int32)
        t = _d(targets).reshape(-1)
        self._att = (v, t)
        lib.orc_set_attachments(self.s, len(v), p(v, C.c_int32), p(t))

    def inverse_inertia(self):
        return self._get(lib.orc_get_inverse_inertia, (self.nv,))

    def set_local_collision_iters(self, n):
        """Extra passes over the elements around active contacts, then all
        contacts, after the global solve (num-local-collision-iters)."""
        lib.orc_set_local_collision_iters(self.s, int(n))

    def last_local_list_size(self):
        """Element projections per local pass in the last step (repeats included)."""
        return lib.orc_last_local_list_size(self.s)

    def local_list_size(self):
        """The same count for the current contact multipliers, by brute force."""
        return lib.orc_local_list_size(self.s)

    def contact_lambdas(self):
        n = lib.orc_num_contacts(self.s)
        out = np.zeros(max(n, 1))
        lib.orc_get_contact_lambdas(self.s, p(out))
        return out[:n]

    def set_sdf(self, kind, pos, quat=(0, 0, 0, 1), params=(0, 0, 0)):
        lib.orc_set_sdf(self.s, kind, p(_d(pos)), p(_d(quat)), p(_d(params)))

    def set_state(self, x=None, v=None):
        lib.orc_set_state(self.s, p(None if x is None else _d(x)), p(None if v is None else _d(v)))

    def detect(self):
        return lib.orc_detect(self.s)

    def contacts(self):
        n = lib.orc_num_contacts(self.s)
        face = np.empty(n, np.int32)
        pp, nn, bb = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        lib.orc_get_contacts(self.s, p(face, C.c_int32), p(pp), p(nn), p(bb))
        return face, pp, nn, bb

    def step(self, n=1, detect=False):
        for _ in range(n):
            lib.orc_step(self.s, 1 if detect else 0)

    def predict(self):
        lib.orc_predict(self.s)

    def step_learned(self, corr, iters, detect=False):
        """Learned-warm-start step: multipliers and positions started from the
        predicted correction `corr` (V x 3, or None for zero)."""
        c = None if corr is None else _d(corr).reshape(-1)
        lib.orc_step_learned(self.s, 1 if detect else 0, p(c), iters)

    def step_dual_warm(self, lam0, iters, detect=False):
        """One step whose multipliers start at lam0 (T x 2) with the consistent position start."""
        l0 = None if lam0 is None else _d(lam0).reshape(-1)
        lib.orc_step_dual_warm(self.s, 1 if detect else 0, p(l0), iters)

    def step_warm(self, corr, iters, detect=False):
        """One step started at the predicted positions x~ + corr (V x 3, or None)
        with every element's multipliers at -C / alpha~ there, then `iters`
        iterations without a reset (the device's learned step)."""
        c = None if corr is None else _d(corr).reshape(-1)
        lib.orc_step_warm(self.s, 1 if detect else 0, p(c), iters)

    def momentum_residual(self, xt, vec=False):
        """||(x - xt) / w - sum grad C^T lambda|| (elements and contacts) over
        the vertices that are neither pinned nor attached, for the inertial
        prediction xt (V x 3)."""
        out = np.empty((self.nv, 3)) if vec else None
        n = lib.orc_momentum_residual(self.s, p(_d(xt)), p(out))
        return (n, out) if vec else n

    def set_primary_residual(self, on=True):
        """Record the primary residual at the end of every solve (as the
        residual policy EVERY_SUBSTEP does)."""
        lib.orc_set_primary_residual(self.s, int(bool(on)))

    def last_primary_residual(self, vec=False):
        """Norm (and, with vec, the V x 3 vector) of the primary residual
        recorded at the end of the last solve."""
        out = np.empty((self.nv, 3)) if vec else None
        n = lib.orc_last_primary_residual(self.s, p(out))
        return (n, out) if vec else n

    def last_prediction(self):
        """The last step's inertial prediction xt (V x 3), kept while the
        primary residual is recorded."""
        out = np.zeros((self.nv, 3))
        lib.orc_last_prediction(self.s, p(out))
        return out

    def set_lambdas(self, lam):
        """Overwrite the element multipliers (T x 2: lambda_D, lambda_H)."""
        lib.orc_set_lambdas(self.s, p(_d(lam).reshape(-1)))

    def solve_elements(self, elements, num_iter=1, initialize=False, with_contacts=False):
        """XPBDSolver::solve(projector_references, num_iter, initialize) on a list
        of elements (repeats allowed), optionally followed by every contact."""
        e = np.ascontiguousarray(elements, np.int32)
        lib.orc_solve_elements(self.s, len(e), p(e, C.c_int32), int(num_iter), int(bool(initialize)),
                               int(bool(with_contacts)))

    def step_corrected(self, corr, iters, detect=False):
        """One step with `corr` (V x 3) added to the predicted positions before `iters` iterations."""
        c = None if corr is None else _d(corr).reshape(-1)
        lib.orc_step_corrected(self.s, 1 if detect else 0, p(c), iters)

    def _get(self, fn, shape):
        out = np.empty(shape)
        fn(self.s, p(out))
        return out

    def x(self):
        return self._get(lib.orc_get_x, (self.nv, 3))

    def xprev(self):
        return self._get(lib.orc_get_xprev, (self.nv, 3))

    def v(self):
        return self._get(lib.orc_get_v, (self.nv, 3))

    def lambdas(self):
        return self._get(lib.orc_get_lambda, (self.nt, 2))

    def masses(self):
        return self._get(lib.orc_get_masses, (self.nv,))

    def rest(self):
        Q = np.empty((self.nt, 9))
        vol = np.empty(self.nt)
        lib.orc_get_rest(self.s, p(Q), p(vol))
        return Q, vol

    def compliance(self):
        aD, aH, g = np.empty(self.nt), np.empty(self.nt), np.empty(self.nt)
        lib.orc_get_compliance(self.s, p(aD), p(aH), p(g))
        return aD, aH, g

    def residual(self, vec=False):
        out = np.empty((self.nt, 2)) if vec else None
        n = lib.orc_residual(self.s, p(out))
        return (n, out) if vec else n

    def energy(self):
        return lib.orc_energy(self.s)

    def last_step_residual(self):
        return lib.orc_last_step_residual(self.s)
