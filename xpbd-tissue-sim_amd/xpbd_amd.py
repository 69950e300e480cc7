"""ctypes binding of libxpbd_hip.so (the C-ABI in include/xpbd_hip.h).

Thin Python plumbing for tests, benchmarks and scripting: every compute call
goes through the HIP library.  There is no CPU fallback — if the library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxpbd_hip.so")

XPBD_SOLVER_JACOBI = 0
XPBD_SOLVER_GS_COLORED = 1
SDF_NONE, SDF_SPHERE, SDF_BOX = 0, 1, 2
MLP_PARAMS = 24 * 128 + 128 + 128 * 128 + 128 + 3 * 128 + 3

ERRORS = {
    -1: "XPBD_E_INVALID_ARG", -2: "XPBD_E_HIP", -3: "XPBD_E_NONFINITE", -4: "XPBD_E_DEGENERATE",
    -5: "XPBD_E_UNSUPPORTED", -6: "XPBD_E_NOMEM", -7: "XPBD_E_STATE", -8: "XPBD_E_IO",
}


class XpbdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `make -C {_HERE}` (or __graft_entry__.build())")
lib = C.CDLL(LIB_PATH)

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("num_tets", C.c_int32), ("x0", _dp), ("v0", _dp),
                ("tets", _ip), ("E", _dp), ("nu", _dp), ("rho", _dp), ("pinned", _u8p),
                ("num_faces", C.c_int32), ("faces", _ip)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("g", C.c_double), ("iters", C.c_int32), ("solver", C.c_int32),
                ("deterministic", C.c_int32), ("collision_every", C.c_int32), ("variant", C.c_int32),
                ("first_order", C.c_int32), ("damping_multiplier", C.c_double), ("adjust_damping", C.c_int32),
                ("local_collision_iters", C.c_int32)]


def _sig(name, args, res=C.c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("xpbd_create", [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(_vp)])
_sig("xpbd_batch_create", [C.c_int32, C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(_vp)])
_sig("xpbd_destroy", [_vp])
_sig("xpbd_last_error", [_vp], C.c_char_p)
_sig("xpbd_step", [_vp, C.c_int32])
_sig("xpbd_set_iterations", [_vp, C.c_int32])
_sig("xpbd_synchronize", [_vp])
_sig("xpbd_get_positions", [_vp, _dp])
_sig("xpbd_get_positions_f32", [_vp, _fp])
_sig("xpbd_get_velocities", [_vp, _dp])
_sig("xpbd_set_state", [_vp, _dp, _dp])
_sig("xpbd_get_masses", [_vp, _dp])
_sig("xpbd_get_lambdas", [_vp, _dp])
_sig("xpbd_constraint_residual", [_vp, _dp, _dp])
_sig("xpbd_strain_energy", [_vp, _dp])
_sig("xpbd_set_residual_policy", [_vp, C.c_int32])
_sig("xpbd_last_step_residual", [_vp, _dp])
_sig("xpbd_set_primary_residual", [_vp, C.c_int32])
_sig("xpbd_last_step_primary_residual", [_vp, _dp, _dp])
_sig("xpbd_num_colors", [_vp, _ip])
_sig("xpbd_get_tet_order", [_vp, _ip])
_sig("xpbd_get_face_rank", [_vp, _ip])
_sig("xpbd_get_color_offsets", [_vp, _ip])
_sig("xpbd_set_learned_start", [_vp, C.c_int32])
_sig("xpbd_mlp_forward_noalloc", [_vp, C.c_int32, _fp, _fp])
_sig("xpbd_set_contacts", [_vp, C.c_int32, _ip, _dp, _dp, _dp])
_sig("xpbd_solve_elements", [_vp, C.c_int32, _ip, C.c_int32, C.c_int32, C.c_int32])
_sig("xpbd_set_sdf", [_vp, C.c_int32, _dp, _dp])
_sig("xpbd_set_sdf_velocity", [_vp, _dp])
_sig("xpbd_detect_contacts", [_vp])
_sig("xpbd_get_contacts", [_vp, C.c_int32, _ip, _ip, _dp, _dp, _dp])
_sig("xpbd_set_kernel_timing", [_vp, C.c_int32])
_sig("xpbd_kernel_stats", [_vp, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
_sig("xpbd_batch_num_instances", [_vp, _ip])
_sig("xpbd_batch_set_sdf", [_vp, C.c_int32, C.c_int32, _dp, _dp])
_sig("xpbd_batch_get_positions", [_vp, C.c_int32, _dp])
_sig("xpbd_autotune", [_vp, C.c_char_p])
_sig("xpbd_get_variant", [_vp, _ip, _ip, _ip])
_sig("xpbd_set_attachments", [_vp, C.c_int32, _ip, _dp])
_sig("xpbd_mesh_grid", [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, _dp, _dp, _ip])
_sig("xpbd_mesh_surface", [C.c_int32, C.c_int32, _ip, _ip, _ip])
_sig("xpbd_color_items", [C.c_int32, C.c_int32, C.c_int32, _ip, _ip, _ip])
_sig("xpbd_read_msh", [C.c_char_p, _ip, _ip, _ip, _dp, _ip, _ip])
_sig("xpbd_read_element_classes", [C.c_char_p, C.c_int32, _ip, _ip])
_sig("xpbd_mlp_create", [C.POINTER(_vp)])
_sig("xpbd_mlp_destroy", [_vp])
_sig("xpbd_mlp_num_params", [_ip])
_sig("xpbd_mlp_set_params", [_vp, _fp])
_sig("xpbd_mlp_get_params", [_vp, _fp])
_sig("xpbd_mlp_set_scaling", [_vp, _fp, _fp, C.c_float])
_sig("xpbd_mlp_save", [_vp, C.c_char_p])
_sig("xpbd_mlp_load", [_vp, C.c_char_p])
_sig("xpbd_batch_step", [_vp, C.c_int32])
_sig("xpbd_mlp_forward", [_vp, C.c_int32, _fp, _fp])
_sig("xpbd_mlp_loss_grad", [_vp, C.c_int32, _fp, _fp, _dp, _fp])
_sig("xpbd_mlp_train", [_vp, C.c_int32, _fp, _fp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _dp])
_sig("xpbd_collect_training", [_vp, C.c_int32, C.c_int32, _fp, _fp, _ip])
_sig("xpbd_features", [_vp, _fp])
_sig("xpbd_attach_mlp", [_vp, _vp])
_sig("xpbd_step_learned", [_vp, C.c_int32, C.c_double, C.c_double, _ip, _dp])
_sig("xpbd_learned_stats", [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _dp, C.POINTER(C.c_int64)])
_sig("xpbd_mlp_last_error", [_vp], C.c_char_p)


def ptr(a: Optional[np.ndarray], ct):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    return a.ctypes.data_as(C.POINTER(ct))


def _check(rc: int, handle=None):
    if rc != 0:
        msg = lib.xpbd_last_error(handle)
        raise XpbdError(rc, msg.decode() if msg else "")


# ---------------------------------------------------------------------------
# host-side mesh utilities
# ---------------------------------------------------------------------------
def grid_mesh(nx: int, ny: int, nz: int, h, origin=(0.0, 0.0, 0.0)):
    """(x[V,3] float64, tets[T,4] int32) of a 6-tet-per-cell structured grid."""
    hx, hy, hz = (h, h, h) if np.isscalar(h) else h
    nv = (nx + 1) * (ny + 1) * (nz + 1)
    x = np.empty((nv, 3), np.float64)
    t = np.empty((6 * nx * ny * nz, 4), np.int32)
    o = np.asarray(origin, np.float64)
    _check(lib.xpbd_mesh_grid(nx, ny, nz, hx, hy, hz, ptr(o, C.c_double), ptr(x, C.c_double), ptr(t, C.c_int32)))
    return x, t


def surface_faces(nv: int, tets: np.ndarray) -> np.ndarray:
    tets = np.ascontiguousarray(tets, np.int32)
    f = np.empty((4 * len(tets), 3), np.int32)
    nf = C.c_int32()
    _check(lib.xpbd_mesh_surface(nv, len(tets), ptr(tets, C.c_int32), C.byref(nf), ptr(f, C.c_int32)))
    return f[: nf.value].copy()


def color_items(nv: int, items: np.ndarray):
    items = np.ascontiguousarray(items, np.int32)
    col = np.empty(len(items), np.int32)
    nc = C.c_int32()
    _check(lib.xpbd_color_items(nv, len(items), items.shape[1], ptr(items, C.c_int32), ptr(col, C.c_int32),
                                C.byref(nc)))
    return col, nc.value


def read_msh(path: str):
    nv, nt, nf = C.c_int32(), C.c_int32(), C.c_int32()
    null_d, null_i = ptr(None, C.c_double), ptr(None, C.c_int32)
    _check(lib.xpbd_read_msh(path.encode(), C.byref(nv), C.byref(nt), C.byref(nf), null_d, null_i, null_i))
    x = np.empty((nv.value, 3), np.float64)
    t = np.empty((nt.value, 4), np.int32)
    f = np.empty((nf.value, 3), np.int32)
    _check(lib.xpbd_read_msh(path.encode(), C.byref(nv), C.byref(nt), C.byref(nf), ptr(x, C.c_double),
                             ptr(t, C.c_int32), ptr(f, C.c_int32)))
    return x, t, f


def read_element_classes(path: str) -> np.ndarray:
    """Element classes file: one integer per line (element-classes-filename)."""
    n = C.c_int32()
    _check(lib.xpbd_read_element_classes(path.encode(), 0, ptr(None, C.c_int32), C.byref(n)))
    out = np.empty(n.value, np.int32)
    _check(lib.xpbd_read_element_classes(path.encode(), n.value, ptr(out, C.c_int32), C.byref(n)))
    return out


def materials_from_classes(classes, materials):
    """Per-element (E, nu, rho) from element classes and a list of materials
    (dicts with E, nu, density), as XPBDMeshObject.cpp:287-300 maps them: a
    class at or past the end of the list takes the last material."""
    cls = np.asarray(classes, np.int64)
    idx = np.where((cls < 0) | (cls >= len(materials)), len(materials) - 1, cls)
    E = np.array([m["E"] for m in materials], np.float64)[idx]
    nu = np.array([m["nu"] for m in materials], np.float64)[idx]
    rho = np.array([m["density"] for m in materials], np.float64)[idx]
    return E, nu, rho


# ---------------------------------------------------------------------------
# objects
# ---------------------------------------------------------------------------
@dataclass
class Mesh:
    """One tetrahedral object (all arrays host-side, caller's ordering)."""
    x0: np.ndarray                      # (V,3) float64
    tets: np.ndarray                    # (T,4) int32
    E: np.ndarray                       # (T,) float64
    nu: np.ndarray                      # (T,)
    rho: np.ndarray                     # (T,)
    pinned: Optional[np.ndarray] = None  # (V,) uint8
    v0: Optional[np.ndarray] = None     # (V,3)
    faces: Optional[np.ndarray] = None  # (F,3) int32
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nv(self):
        return len(self.x0)

    @property
    def nt(self):
        return len(self.tets)

    def desc(self) -> MeshDesc:
        arrs = dict(
            x0=np.ascontiguousarray(self.x0, np.float64),
            tets=np.ascontiguousarray(self.tets, np.int32),
            E=np.ascontiguousarray(self.E, np.float64),
            nu=np.ascontiguousarray(self.nu, np.float64),
            rho=np.ascontiguousarray(self.rho, np.float64),
            pinned=None if self.pinned is None else np.ascontiguousarray(self.pinned, np.uint8),
            v0=None if self.v0 is None else np.ascontiguousarray(self.v0, np.float64),
            faces=None if self.faces is None else np.ascontiguousarray(self.faces, np.int32),
        )
        self._keep = list(arrs.values())
        d = MeshDesc()
        d.num_vertices = len(arrs["x0"])
        d.num_tets = len(arrs["tets"])
        d.x0 = ptr(arrs["x0"], C.c_double)
        d.v0 = ptr(arrs["v0"], C.c_double)
        d.tets = ptr(arrs["tets"], C.c_int32)
        d.E = ptr(arrs["E"], C.c_double)
        d.nu = ptr(arrs["nu"], C.c_double)
        d.rho = ptr(arrs["rho"], C.c_double)
        d.pinned = ptr(arrs["pinned"], C.c_uint8)
        d.num_faces = 0 if arrs["faces"] is None else len(arrs["faces"])
        d.faces = ptr(arrs["faces"], C.c_int32)
        return d


SOLVERS = {"jacobi": XPBD_SOLVER_JACOBI, "gs": XPBD_SOLVER_GS_COLORED}


VARIANTS = {"auto": 0, "simple": 1, "cluster": 2, "patch": 3, "quad": 4}
QUAD_SPLIT = 1000  # XPBD_QUAD_SPLIT: added to a quad block size, the iterations keep the deltas in their own array


def make_params(dt=1e-3, g=9.81, iters=10, solver="gs", deterministic=1, collision_every=1,
                variant="auto", first_order=False, damping_multiplier=1.0, adjust_damping=False,
                local_collision_iters=0) -> Params:
    p = Params()
    p.dt, p.g, p.iters = dt, g, iters
    p.solver = SOLVERS[solver] if isinstance(solver, str) else solver
    p.deterministic, p.collision_every = deterministic, collision_every
    p.variant = VARIANTS[variant] if isinstance(variant, str) else variant
    p.first_order, p.damping_multiplier, p.adjust_damping = int(first_order), damping_multiplier, int(adjust_damping)
    p.local_collision_iters = int(local_collision_iters)
    return p


class Solver:
    """A device solver for one object (or a batch of objects)."""

    def __init__(self, mesh, dt=1e-3, g=9.81, iters=10, solver="gs", deterministic=1, collision_every=1,
                 variant="auto", first_order=False, damping_multiplier=1.0, adjust_damping=False,
                 local_collision_iters=0):
        meshes = mesh if isinstance(mesh, (list, tuple)) else [mesh]
        self.meshes = list(meshes)
        self.params = make_params(dt, g, iters, solver, deterministic, collision_every, variant, first_order,
                                  damping_multiplier, adjust_damping, local_collision_iters)
        descs = (MeshDesc * len(meshes))(*[m.desc() for m in meshes])
        h = _vp()
        if len(meshes) == 1 and not isinstance(mesh, (list, tuple)):
            rc = lib.xpbd_create(descs, C.byref(self.params), C.byref(h))
        else:
            rc = lib.xpbd_batch_create(len(meshes), descs, C.byref(self.params), C.byref(h))
        _check(rc, None)
        self.h = h
        self.nv = sum(m.nv for m in meshes)
        self.nt = sum(m.nt for m in meshes)
        self.nf = sum(0 if m.faces is None else len(m.faces) for m in meshes)

    def close(self):
        if getattr(self, "h", None):
            lib.xpbd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.h)

    def step(self, n=1):
        self._c(lib.xpbd_step(self.h, n))

    def set_iterations(self, n):
        self._c(lib.xpbd_set_iterations(self.h, n))

    def sync(self):
        self._c(lib.xpbd_synchronize(self.h))

    def positions(self) -> np.ndarray:
        out = np.empty((self.nv, 3), np.float64)
        self._c(lib.xpbd_get_positions(self.h, ptr(out, C.c_double)))
        return out

    def positions_f32(self) -> np.ndarray:
        out = np.empty((self.nv, 3), np.float32)
        self._c(lib.xpbd_get_positions_f32(self.h, ptr(out, C.c_float)))
        return out

    def velocities(self) -> np.ndarray:
        out = np.empty((self.nv, 3), np.float64)
        self._c(lib.xpbd_get_velocities(self.h, ptr(out, C.c_double)))
        return out

    def set_state(self, x=None, v=None):
        xa = None if x is None else np.ascontiguousarray(x, np.float64)
        va = None if v is None else np.ascontiguousarray(v, np.float64)
        self._c(lib.xpbd_set_state(self.h, ptr(xa, C.c_double), ptr(va, C.c_double)))

    def masses(self):
        out = np.empty(self.nv, np.float64)
        self._c(lib.xpbd_get_masses(self.h, ptr(out, C.c_double)))
        return out

    def lambdas(self):
        out = np.empty((self.nt, 2), np.float64)
        self._c(lib.xpbd_get_lambdas(self.h, ptr(out, C.c_double)))
        return out

    def residual(self, vec=False):
        n = C.c_double()
        out = np.empty((self.nt, 2), np.float64) if vec else None
        self._c(lib.xpbd_constraint_residual(self.h, C.byref(n), ptr(out, C.c_double)))
        return (n.value, out) if vec else n.value

    def set_residual_policy(self, every_step: bool):
        self._c(lib.xpbd_set_residual_policy(self.h, int(every_step)))

    def last_step_residual(self):
        r = C.c_double()
        self._c(lib.xpbd_last_step_residual(self.h, C.byref(r)))
        return r.value

    def set_primary_residual(self, on: bool):
        self._c(lib.xpbd_set_primary_residual(self.h, int(on)))

    def last_step_primary_residual(self, vec=False):
        """Norm (and, with vec, the V x 3 vector) of the primary residual
        (x - xt) / w - sum grad C^T lambda recorded at the end of the last
        step's solve (set_primary_residual on)."""
        n = C.c_double()
        out = np.empty((self.nv, 3), np.float64) if vec else None
        self._c(lib.xpbd_last_step_primary_residual(self.h, C.byref(n), ptr(out, C.c_double)))
        return (n.value, out) if vec else n.value

    def energy(self):
        e = C.c_double()
        self._c(lib.xpbd_strain_energy(self.h, C.byref(e)))
        return e.value

    def num_colors(self):
        n = C.c_int32()
        self._c(lib.xpbd_num_colors(self.h, C.byref(n)))
        return n.value

    def tet_order(self):
        out = np.empty(self.nt, np.int32)
        self._c(lib.xpbd_get_tet_order(self.h, ptr(out, C.c_int32)))
        return out

    def face_rank(self):
        out = np.empty(self.nf, np.int32)
        self._c(lib.xpbd_get_face_rank(self.h, ptr(out, C.c_int32)))
        return out

    def color_offsets(self):
        """Colour boundaries (ncolors + 1 positions) in the tet_order() array."""
        out = np.empty(self.num_colors() + 1, np.int32)
        self._c(lib.xpbd_get_color_offsets(self.h, ptr(out, C.c_int32)))
        return out

    def set_sdf(self, kind, pos, quat=(0, 0, 0, 1), params=(0, 0, 0), instance=None):
        pose = np.asarray(list(pos) + list(quat), np.float64)
        prm = np.asarray(params, np.float64)
        if instance is None:
            self._c(lib.xpbd_set_sdf(self.h, kind, ptr(pose, C.c_double), ptr(prm, C.c_double)))
        else:
            self._c(lib.xpbd_batch_set_sdf(self.h, instance, kind, ptr(pose, C.c_double), ptr(prm, C.c_double)))

    def set_attachments(self, vertices, targets):
        """Pull vertices[i] to the world points targets[i] (replaces the list; empty clears)."""
        v = np.ascontiguousarray(vertices, np.int32).reshape(-1)
        t = np.ascontiguousarray(targets, np.float64).reshape(-1)
        self._c(lib.xpbd_set_attachments(self.h, len(v), ptr(v, C.c_int32), ptr(t, C.c_double)))

    def set_sdf_velocity(self, vel):
        v = np.asarray(vel, np.float64)
        self._c(lib.xpbd_set_sdf_velocity(self.h, ptr(v, C.c_double)))

    def detect(self):
        self._c(lib.xpbd_detect_contacts(self.h))

    def contacts(self):
        cap = max(self.nf, 1)
        cnt = C.c_int32()
        face = np.empty(cap, np.int32)
        p = np.empty((cap, 3))
        n = np.empty((cap, 3))
        b = np.empty((cap, 3))
        self._c(lib.xpbd_get_contacts(self.h, cap, C.byref(cnt), ptr(face, C.c_int32), ptr(p, C.c_double),
                                      ptr(n, C.c_double), ptr(b, C.c_double)))
        k = cnt.value
        return face[:k], p[:k], n[:k], b[:k]

    def set_contacts(self, faces, p, n, bary):
        """Replace the contact set (clearCollisionConstraints +
        addStaticCollisionConstraint per contact); empty arrays clear it."""
        f = np.ascontiguousarray(faces, np.int32).reshape(-1)
        pp = np.ascontiguousarray(p, np.float64).reshape(-1)
        nn = np.ascontiguousarray(n, np.float64).reshape(-1)
        bb = np.ascontiguousarray(bary, np.float64).reshape(-1)
        self._c(lib.xpbd_set_contacts(self.h, len(f), ptr(f, C.c_int32), ptr(pp, C.c_double), ptr(nn, C.c_double),
                                      ptr(bb, C.c_double)))

    def solve_elements(self, elements, num_iter=1, initialize=False, with_contacts=False):
        """XPBDSolver::solve(projector_references, num_iter, initialize) on a list
        of elements (repeats allowed), optionally followed by every contact."""
        e = np.ascontiguousarray(elements, np.int32).reshape(-1)
        self._c(lib.xpbd_solve_elements(self.h, len(e), ptr(e, C.c_int32), int(num_iter), int(bool(initialize)),
                                        int(bool(with_contacts))))

    def set_kernel_timing(self, on: bool):
        self._c(lib.xpbd_set_kernel_timing(self.h, int(on)))

    def kernel_stats(self):
        ms, it, ln = C.c_double(), C.c_int64(), C.c_int64()
        self._c(lib.xpbd_kernel_stats(self.h, C.byref(ms), C.byref(it), C.byref(ln)))
        return ms.value, it.value, ln.value

    # batch helpers
    def instance_positions(self, i):
        out = np.empty((self.meshes[i].nv, 3), np.float64)
        self._c(lib.xpbd_batch_get_positions(self.h, i, ptr(out, C.c_double)))
        return out

    def autotune(self, cache_path: Optional[str] = None):
        self._c(lib.xpbd_autotune(self.h, None if cache_path is None else cache_path.encode()))
        return self.variant()

    def variant(self):
        """(variant name, workgroup size or 0) of the kernels in use."""
        return self.launch_config()[:2]

    def launch_config(self):
        """(variant name, workgroup size or 0, XCD-contiguous mapping on) of the kernels in use."""
        v, b, x = C.c_int32(), C.c_int32(), C.c_int32()
        self._c(lib.xpbd_get_variant(self.h, C.byref(v), C.byref(b), C.byref(x)))
        names = {val: k for k, val in VARIANTS.items()}
        return names.get(v.value, str(v.value)), b.value, bool(x.value)

    # learned warm start
    def attach(self, net: Optional["Network"]):
        self._net = net
        self._c(lib.xpbd_attach_mlp(self.h, None if net is None else net.m))

    def features(self) -> np.ndarray:
        out = np.empty((self.nv, MLP_IN), np.float32)
        self._c(lib.xpbd_features(self.h, ptr(out, C.c_float)))
        return out

    def collect_training(self, nsteps: int, cap: Optional[int] = None):
        """(features[rows, 24], corrections[rows, 3]) of nsteps plain steps (free vertices)."""
        nfree = int(sum((m.pinned == 0).sum() if m.pinned is not None else m.nv for m in self.meshes))
        cap = nsteps * nfree if cap is None else cap
        x = np.empty((max(cap, 1), MLP_IN), np.float32)
        y = np.empty((max(cap, 1), 3), np.float32)
        rows = C.c_int32()
        self._c(lib.xpbd_collect_training(self.h, nsteps, cap, ptr(x, C.c_float), ptr(y, C.c_float),
                                          C.byref(rows)))
        return x[:rows.value], y[:rows.value]

    def set_learned_start(self, mode: str):
        """'multipliers' (default: element multipliers at -C / alpha~ at the
        corrected positions), 'positions' (multipliers from zero) or
        'consistent' (those multipliers with the positions that keep the
        momentum balance, xt + M^-1 grad C^T lambda0)."""
        self._c(lib.xpbd_set_learned_start(self.h, {"positions": 0, "multipliers": 1, "consistent": 2}[mode]))

    def step_learned(self, iters: int, ref_residual: float = 0.0, ratio: float = 1.1):
        """One learned step; returns (fell_back, residual)."""
        fb, r = C.c_int32(), C.c_double()
        self._c(lib.xpbd_step_learned(self.h, iters, ref_residual, ratio, C.byref(fb), C.byref(r)))
        return bool(fb.value), r.value

    def learned_stats(self):
        st, fb, n = C.c_int64(), C.c_int64(), C.c_int64()
        ms = C.c_double()
        self._c(lib.xpbd_learned_stats(self.h, C.byref(st), C.byref(fb), C.byref(ms), C.byref(n)))
        return {"steps": st.value, "fallbacks": fb.value, "mlp_ms": ms.value, "mlp_calls": n.value}


MLP_IN, MLP_HIDDEN, MLP_OUT = 24, 128, 3


class Network:
    """The per-vertex warm-start MLP (24 -> 128 -> 128 -> 3, SiLU) on the device."""

    def __init__(self, params: Optional[np.ndarray] = None):
        m = _vp()
        _check(lib.xpbd_mlp_create(C.byref(m)), None)
        self.m = m
        if params is not None:
            self.set_params(params)

    def close(self):
        if getattr(self, "m", None):
            lib.xpbd_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        if rc != 0:
            msg = lib.xpbd_mlp_last_error(self.m)
            raise XpbdError(rc, msg.decode() if msg else "")

    def params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        self._c(lib.xpbd_mlp_get_params(self.m, ptr(out, C.c_float)))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32).reshape(MLP_PARAMS)
        self._c(lib.xpbd_mlp_set_params(self.m, ptr(a, C.c_float)))

    def set_scaling(self, mean, std, scale_out):
        mu = np.ascontiguousarray(mean, np.float32).reshape(MLP_IN)
        sd = np.ascontiguousarray(std, np.float32).reshape(MLP_IN)
        self._c(lib.xpbd_mlp_set_scaling(self.m, ptr(mu, C.c_float), ptr(sd, C.c_float), float(scale_out)))

    def save(self, path: str):
        """Write parameters and scaling as a network file (xpbd_mlp_save)."""
        self._c(lib.xpbd_mlp_save(self.m, os.fsencode(path)))

    def load(self, path: str):
        """Read parameters and scaling from a network file (xpbd_mlp_load)."""
        self._c(lib.xpbd_mlp_load(self.m, os.fsencode(path)))

    def forward(self, x) -> np.ndarray:
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        out = np.empty((len(xa), MLP_OUT), np.float32)
        self._c(lib.xpbd_mlp_forward(self.m, len(xa), ptr(xa, C.c_float), ptr(out, C.c_float)))
        return out

    def loss_grad(self, x, y):
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        ya = np.ascontiguousarray(y, np.float32).reshape(-1, MLP_OUT)
        loss = C.c_double()
        g = np.empty(MLP_PARAMS, np.float32)
        self._c(lib.xpbd_mlp_loss_grad(self.m, len(xa), ptr(xa, C.c_float), ptr(ya, C.c_float), C.byref(loss),
                                       ptr(g, C.c_float)))
        return loss.value, g

    def train(self, x, y, batch=4096, epochs=10, lr=1e-3, seed=0):
        xa = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        ya = np.ascontiguousarray(y, np.float32).reshape(-1, MLP_OUT)
        loss = C.c_double()
        self._c(lib.xpbd_mlp_train(self.m, len(xa), ptr(xa, C.c_float), ptr(ya, C.c_float), batch, epochs,
                                   float(lr), seed, C.byref(loss)))
        return loss.value

    def fit_scaling(self, x, y):
        """Standardise inputs by their mean/std and outputs by their RMS (std floor 1e-12)."""
        x = np.asarray(x, np.float64)
        mu = x.mean(0)
        sd = x.std(0)
        # a (near-)constant input keeps a unit scale instead of amplifying noise
        sd = np.where(sd > 1e-6 * np.abs(mu) + 1e-12, sd, np.maximum(np.abs(mu), 1.0))
        sc = float(max(np.sqrt(np.mean(np.asarray(y, np.float64) ** 2)), 1e-12))
        self.set_scaling(mu, sd, sc)
        return mu, sd, sc
