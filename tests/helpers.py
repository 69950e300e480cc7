"""Shared helpers for the parity tests."""
import numpy as np


def rel_err(x, ref):
    """max_i |x_i - ref_i| / max(|ref|_inf, 1e-12): the per-step position error measure."""
    return float(np.max(np.abs(np.asarray(x) - np.asarray(ref))) / max(np.max(np.abs(ref)), 1e-12))


def run_pair(gpu, orc, steps, detect=False, sdf_motion=None, check_every=1, contacts=None):
    """Step the device solver and the oracle together; return the per-step errors.
    If `contacts` is a list, the device contact count after each step is appended."""
    errs = []
    for n in range(steps):
        if sdf_motion is not None:
            kind, pos0, quat, prm, vel = sdf_motion
            pos = np.asarray(pos0) + np.asarray(vel) * n * gpu.params.dt
            orc.set_sdf(kind, pos, quat, prm)
        gpu.step(1)
        orc.step(1, detect=detect)
        if contacts is not None:
            contacts.append(len(gpu.contacts()[0]))
        if (n + 1) % check_every == 0:
            errs.append(rel_err(gpu.positions(), orc.x()))
    return np.array(errs)


def disp_err(x, ref, rest):
    """max_i |x_i - ref_i| / max(|ref - rest|_inf, 1e-300): the position error in
    units of the largest displacement from the rest shape, which does not hide
    an error of one step's motion behind the size of the coordinates."""
    ref = np.asarray(ref)
    return float(np.max(np.abs(np.asarray(x) - ref)) / max(np.max(np.abs(ref - np.asarray(rest))), 1e-300))
