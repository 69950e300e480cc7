"""Host-side logic and the C-ABI boundary (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import xpbd_amd as X

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "xpbd_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xpbd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) > 40
    lib = C.CDLL(X.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


# ---------------------------------------------------------------- mesh tools
@pytest.mark.parametrize("n", [(1, 1, 1), (3, 2, 4), (10, 10, 10)])
def test_grid_counts_volume_and_conformity(n):
    nx, ny, nz = n
    x, t = X.grid_mesh(nx, ny, nz, 0.01)
    assert len(x) == (nx + 1) * (ny + 1) * (nz + 1) and len(t) == 6 * nx * ny * nz
    D = np.stack([x[t[:, 0]] - x[t[:, 3]], x[t[:, 1]] - x[t[:, 3]], x[t[:, 2]] - x[t[:, 3]]], axis=2)
    vol = np.abs(np.linalg.det(D)) / 6
    np.testing.assert_allclose(vol, 1e-6 / 6, rtol=1e-9)             # 6 equal tets per cell
    assert vol.sum() == pytest.approx(nx * ny * nz * 1e-6, rel=1e-12)
    f = X.surface_faces(len(x), t)
    # conforming split: every boundary quad is exactly two triangles, all on the box surface
    assert len(f) == 4 * (nx * ny + ny * nz + nx * nz)
    assert np.all(_on_box_surface(x, f))


def _on_box_surface(x, f):
    lo, hi = x.min(0), x.max(0)
    P = x[f]                                                  # (F, 3 vertices, 3 coords)
    return np.any([(np.isclose(P[:, :, a], lo[a]).all(1) | np.isclose(P[:, :, a], hi[a]).all(1))
                   for a in range(3)], axis=0)


def test_surface_rejects_non_manifold():
    t = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]], np.int32)
    with pytest.raises(X.XpbdError):
        X.surface_faces(6, t)


@pytest.mark.parametrize("n", [4, 9])
def test_coloring_is_valid_and_balanced(n):
    x, t = X.grid_mesh(n, n, n, 0.01)
    col, nc = X.color_items(len(x), t)
    assert nc >= 24                          # interior vertices touch 24 tets
    for c in range(nc):
        sel = t[col == c].ravel()
        assert len(np.unique(sel)) == len(sel), f"colour {c} shares a vertex"
    counts = np.bincount(col)
    assert counts.min() > 0.5 * counts.mean()


def test_face_coloring_is_valid():
    x, t = X.grid_mesh(5, 4, 3, 0.01)
    f = X.surface_faces(len(x), t)
    col, nc = X.color_items(len(x), f)
    for c in range(nc):
        sel = f[col == c].ravel()
        assert len(np.unique(sel)) == len(sel)


@pytest.mark.parametrize("name,nv,nt,nf", [("single.msh", 4, 1, 4), ("cube2.msh", 27, 40, 48),
                                           ("cube4.msh", 125, 320, 192)])
def test_msh_reader_fixture_counts(name, nv, nt, nf):
    x, t, f = X.read_msh(os.path.join(GOLDEN, "msh", name))
    assert (len(x), len(t), len(f)) == (nv, nt, nf)
    assert t.min() >= 0 and t.max() < nv
    D = np.stack([x[t[:, 0]] - x[t[:, 3]], x[t[:, 1]] - x[t[:, 3]], x[t[:, 2]] - x[t[:, 3]]], axis=2)
    vol = np.abs(np.linalg.det(D)).sum() / 6
    P = x[f]
    area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1).sum()
    if name == "single.msh":
        # one tet, its four faces listed as the surface
        assert vol == pytest.approx(np.sqrt(3) * 1.73205080757 / 3, rel=1e-10)
        assert {tuple(sorted(r)) for r in f} == {tuple(sorted(r)) for r in X.surface_faces(4, t)}
    else:
        # unit cube: tets fill it, triangles cover its surface
        ext = x.max(0) - x.min(0)
        assert vol == pytest.approx(np.prod(ext), rel=1e-12)
        assert area == pytest.approx(2 * (ext[0] * ext[1] + ext[1] * ext[2] + ext[0] * ext[2]), rel=1e-12)
        assert np.all(_on_box_surface(x, f))


def test_msh_reader_missing_file():
    with pytest.raises(X.XpbdError) as e:
        X.read_msh("/nonexistent/file.msh")
    assert e.value.code == -8


# ---------------------------------------------------------------- argument validation
def _desc(mesh):
    return mesh.desc()


def _create(mesh, **kw):
    p = X.make_params(**kw)
    h = C.c_void_p()
    rc = X.lib.xpbd_create(C.byref(_desc(mesh)), C.byref(p), C.byref(h))
    return rc, h


def _small():
    import workloads as W
    return W.grid_object(2, 2, 2, 0.01, noise=0.0)


def test_create_rejects_bad_arguments_before_touching_the_device():
    m = _small()
    assert _create(m, dt=0.0)[0] == -1
    assert _create(m, solver=7)[0] == -1
    assert _create(m, iters=-1)[0] == -1
    bad = _small()
    bad.tets = bad.tets.copy()
    bad.tets[0, 0] = 10_000
    assert _create(bad)[0] == -1
    assert b"out of range" in X.lib.xpbd_last_error(None)
    nan = _small()
    nan.x0 = nan.x0.copy()
    nan.x0[3, 1] = np.nan
    assert _create(nan)[0] == -3
    mat = _small()
    mat.nu = np.full(mat.nt, 0.5)
    assert _create(mat)[0] == -1
    flat = _small()
    flat.x0 = flat.x0.copy()
    flat.x0[:, 2] = 0.0
    rc, _ = _create(flat)
    assert rc in (-4, -2)       # zero volume (or no device here)


def test_empty_inputs_are_rejected_before_touching_the_device():
    """A mesh without elements or vertices, and a batch of no meshes, are
    argument errors, not empty solvers."""
    m = _small()
    empty = X.Mesh(x0=np.zeros((0, 3)), tets=np.zeros((0, 4), np.int32), E=np.zeros(0), nu=np.zeros(0),
                   rho=np.zeros(0))
    assert _create(empty)[0] == -1
    no_tets = _small()
    no_tets.tets = np.zeros((0, 4), np.int32)
    no_tets.E, no_tets.nu, no_tets.rho = np.zeros(0), np.zeros(0), np.zeros(0)
    assert _create(no_tets)[0] == -1
    p = X.make_params()
    h = C.c_void_p()
    assert X.lib.xpbd_batch_create(0, C.byref(_desc(m)), C.byref(p), C.byref(h)) == -1
    assert b"n > 0" in X.lib.xpbd_last_error(None)


def test_null_handles_are_rejected():
    assert X.lib.xpbd_step(None, 1) == -1
    assert X.lib.xpbd_set_contacts(None, 0, None, None, None, None) == -1
    assert X.lib.xpbd_solve_elements(None, 0, None, 1, 1, 0) == -1
    assert X.lib.xpbd_step_learned(None, 1, 0.0, 1.1, None, None) == -1
    assert X.lib.xpbd_destroy(None) == 0
    assert X.lib.xpbd_last_error(None) is not None


@pytest.mark.parametrize("name,nv,nt,nf", [("Prostate.msh", 2472, 9127, 3594),
                                           ("tissue_block1x1x0.1_4subdiv.msh", 5417, 20952, 7680),
                                           ("neuroma_segmentation_decimated_2000.msh", 1276, 4448, 2000),
                                           ("2x1x1_block16.msh", 9537, 40960, 5120)])
def test_reference_tissue_meshes_read(name, nv, nt, nf, tmp_path):
    """Tissue meshes shipped with the reference (resource/tissue,
    resource/demos/palpation), read by xpbd_read_msh: counts, element volumes
    positive after orientation, the boundary of the tets is closed (every edge
    of the extracted surface is shared by two faces) and covers the file's
    triangles."""
    import gzip
    p = tmp_path / name
    p.write_bytes(gzip.open(os.path.join(GOLDEN, "msh", name + ".gz")).read())
    x, t, f = X.read_msh(str(p))
    assert (len(x), len(t), len(f)) == (nv, nt, nf)
    a, b, c, d = (x[t[:, k]] for k in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) / 6.0
    assert np.all(np.abs(vol) > 0)
    s = X.surface_faces(len(x), t)
    edges = np.sort(np.concatenate([s[:, [0, 1]], s[:, [1, 2]], s[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    on_surface = {tuple(sorted(tri)) for tri in s.tolist()}
    assert sum(tuple(sorted(tri)) in on_surface for tri in f.tolist()) >= 0.99 * len(f)


def test_truncated_and_corrupt_msh_files_fail_cleanly(tmp_path):
    """A file cut anywhere either reads as a valid prefix (every element index
    in range) or raises; counts that claim more than the header declares, or
    more than the file holds, raise instead of allocating or looping on them."""
    src = open(os.path.join(GOLDEN, "msh", "cube4.msh"), "rb").read()
    rng = np.random.default_rng(5)
    p = tmp_path / "cut.msh"
    for cut in sorted(set(rng.integers(0, len(src), 40).tolist()) | {len(src) // 2, len(src) - 1}):
        p.write_bytes(src[:cut])
        try:
            x, t, _ = X.read_msh(str(p))
        except X.XpbdError:
            continue
        assert t.size == 0 or (t.min() >= 0 and t.max() < len(x))
    text = src.decode()
    lines = text.split("\n")
    k = lines.index("$Nodes")
    header = lines[k + 1].split()
    for bad in (f"{header[0]} 99999999999999 {header[2]} {header[3]}",    # more nodes than the file holds
                f"{header[0]} -5 {header[2]} {header[3]}"):
        q = lines[:k + 1] + [bad] + lines[k + 2:]
        p.write_text("\n".join(q))
        with pytest.raises(X.XpbdError):
            X.read_msh(str(p))
    blk = lines[k + 2].split()                    # first node block: dim, tag, parametric, count
    q = lines[:k + 2] + [" ".join(blk[:3] + ["1000000000000"])] + lines[k + 3:]
    p.write_text("\n".join(q))
    with pytest.raises(X.XpbdError):
        X.read_msh(str(p))


def test_msh_without_nodes_section_is_refused(tmp_path):
    """A file with a $MeshFormat header but no $Nodes / $Elements sections (as
    resource/tissue/liver.msh, whose body is a bare node and element list) is
    not read as an empty mesh."""
    p = tmp_path / "bare.msh"
    p.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n1 3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n")
    with pytest.raises(X.XpbdError, match="Nodes"):
        X.read_msh(str(p))


def test_element_classes_reader(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("0\n2\n1\n  7\nnot a class\n3\n")
    assert X.read_element_classes(str(p)).tolist() == [0, 2, 1, 7]   # stops at the first non-integer line
    E, nu, rho = X.materials_from_classes([0, 2, 1, 7, -1], [{"E": 1.0, "nu": 0.4, "density": 10.0},
                                                             {"E": 2.0, "nu": 0.45, "density": 20.0}])
    assert E.tolist() == [1.0, 2.0, 2.0, 2.0, 2.0] and rho.tolist() == [10.0, 20.0, 20.0, 20.0, 20.0]
    with pytest.raises(X.XpbdError):
        X.read_element_classes(str(tmp_path / "missing.txt"))
