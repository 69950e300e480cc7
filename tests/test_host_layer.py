"""The C++ host layer (xpbd-tissue-sim_amd/host): config parsing and error
behaviour of the xpbd_sim driver, without touching the GPU (--print-config)."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "xpbd-tissue-sim_amd", "host")
SIM = os.path.join(HOST, "xpbd_sim")


def run(cfg_text, tmp_path, *args):
    p = tmp_path / "cfg.yaml"
    p.write_text(cfg_text)
    return subprocess.run([SIM, str(p), *args], capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(SIM):
        pytest.fail("xpbd_sim is not built (make -C xpbd-tissue-sim_amd/host)")


@pytest.mark.parametrize("name", ["beam.yaml", "block_sphere.yaml"])
def test_shipped_configs_parse(name):
    r = subprocess.run([SIM, os.path.join(HOST, "configs", name), "--print-config"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout)
    assert cfg["time_step"] == 1e-3 and len(cfg["objects"]) == 1


REFERENCE_STYLE = """
## comment lines, options commented out, a dash on its own line, quoted strings
name: "Style_Test"   # trailing comment
description: 'single # quoted'
g-accel: 9.5
time-step: 5e-4
collision-rate: 250
end-time: 2
objects:
  -
    name: "Cube1"
    # type: "FirstOrderXPBDMeshObject"
    type: "XPBDMeshObject"
    filename: "mesh.msh"
    material:
      name: "Soft Tissue"
      density: 1100
      E: 6e4
      nu: 0.45
    max-size: 0.5
    position: [0, 0, 2]
    velocity: [0, 0, 0.0]
    num-solver-iters: 7
    num-local-collision-iters: 3
    solver-type: "Parallel-Jacobi"
    # solver-type: "Jacobi"
    residual-policy: "Every-Substep"
  - name: "sphere"
    type: "RigidSphere"
    radius: 0.5
    position: [0, 1, -0.25]
    fixed: true
    collisions: true
"""


def test_reference_style_config(tmp_path):
    r = run(REFERENCE_STYLE, tmp_path, "--print-config")
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout)
    assert c["name"] == "Style_Test" and c["g_accel"] == 9.5 and c["time_step"] == 5e-4
    assert c["collision_rate"] == 250 and c["rigid_objects"] == 1
    o = c["objects"][0]
    assert o == {"name": "Cube1", "solver_type": "Parallel-Jacobi", "iters": 7, "local_collision_iters": 3,
                 "E": 6e4, "nu": 0.45,
                 "density": 1100, "collisions": False, "source": "mesh.msh", "position": [0, 0, 2],
                 "first_order": False, "damping_multiplier": 1, "adjust_damping": False,
                 "materials": [{"name": "Soft Tissue", "E": 6e4, "nu": 0.45, "density": 1100}], "element_classes": ""}


@pytest.mark.parametrize("bad,msg", [
    ('objects:\n  - name: a\n    type: "RigidMeshObject"\n    grid: [1, 1, 1]\n', "not on the device path"),
    ('objects:\n  - name: a\n    type: "FirstOrderXPBDMeshObject"\n    grid: [1, 1, 1]\n    damping-multiplier: 0\n',
     "damping-multiplier must be > 0"),
    ('objects:\n  - name: a\n    grid: [1, 1, 1]\n    solver-type: "Newton"\n', "unknown solver-type"),
    ('objects:\n  - name: a\n    grid: [1, 1, 1]\n    residual-policy: "Sometimes"\n', "unknown residual-policy"),
    ('objects:\n  - name: a\n    grid: [1, 1]\n', "grid needs 3"),
    ('objects:\n  - name: a\n', "needs a filename or a grid"),
    ('time-step: 0\n', "time-step"),
    ('name: a\nname: b\n', "duplicate key"),
    ('objects:\n\t- name: a\n', "tab in indentation"),
    ('key: &anchor 3\n', "unsupported YAML"),
])
def test_config_errors(tmp_path, bad, msg):
    r = run(bad, tmp_path, "--print-config")
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr


def test_missing_config_file(tmp_path):
    r = subprocess.run([SIM, str(tmp_path / "nope.yaml"), "--print-config"], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_grid_object_is_loaded_and_placed(tmp_path):
    r = subprocess.run([SIM, os.path.join(HOST, "configs", "block_sphere.yaml"), "--print-config", "--load-meshes"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)["objects"][0]
    assert (o["vertices"], o["elements"], o["surface_faces"]) == (13 ** 3, 6 * 12 ** 3, 2 * 6 * 12 ** 2)
    # centre of mass moved to the position (0.06, 0.06, 0.06): the 0.12 m cube spans [0, 0.12]
    assert o["bbox"][0] == pytest.approx([0, 0, 0], abs=1e-12)
    assert o["bbox"][1] == pytest.approx([0.12, 0.12, 0.12], abs=1e-12)


def test_msh_object_is_scaled_and_moved(tmp_path):
    msh = os.path.join(ROOT, "tests", "golden", "msh", "cube4.msh")
    cfg = f"""
objects:
  - name: "c"
    type: "XPBDMeshObject"
    filename: "{msh}"
    max-size: 0.5
    position: [1, 2, 3]
    collisions: true
"""
    r = run(cfg, tmp_path, "--print-config", "--load-meshes")
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)["objects"][0]
    lo, hi = o["bbox"]
    ext = [h - l for l, h in zip(lo, hi)]
    assert max(ext) == pytest.approx(0.5, rel=1e-12)          # MeshObject.hpp:41-44 max-size
    centre = [(l + h) / 2 for l, h in zip(lo, hi)]
    assert centre == pytest.approx([1, 2, 3], abs=1e-9)        # symmetric cube: centroid = box centre
    assert o["surface_faces"] > 0 and o["elements"] > 0


def test_first_order_object_config(tmp_path):
    cfg = """
objects:
  - name: "fo"
    type: "FirstOrderXPBDMeshObject"
    grid: [2, 2, 2]
    damping-multiplier: 20000
    adjust-damping-to-material: true
    num-solver-iters: 3
"""
    r = run(cfg, tmp_path, "--print-config")
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)["objects"][0]
    assert o["first_order"] is True and o["damping_multiplier"] == 20000 and o["adjust_damping"] is True


def _gunzip(name, tmp_path):
    import gzip
    src = os.path.join(ROOT, "tests", "golden", "msh", name + ".gz")
    dst = tmp_path / name
    dst.write_bytes(gzip.open(src).read())
    return str(dst)


PALPATION_MATERIALS = """
materials:
  -
    name: "Soft Tissue"
    density: 1000
    E: 6e4
    nu: 0.45
  - {name: "Stiffer Tissue", density: 1030, E: 3e5, nu: 0.49}
"""


def test_element_classes_select_materials(tmp_path):
    """The reference's palpation demo: a 40,960-element block whose classes file
    marks 6,161 elements as class 1 (resource/demos/palpation, config
    config/demos/palpation/palpation.yaml).  Class c takes the object's c-th
    material (XPBDMeshObject.cpp:287-300); a vertex's class is the largest of its
    elements' (XPBDMeshObject.cpp:135-148)."""
    msh = _gunzip("2x1x1_block16.msh", tmp_path)
    cls = _gunzip("2x1x1_block16_element_classes.txt", tmp_path)
    cfg = PALPATION_MATERIALS + f"""
objects:
  - name: "Cube1"
    type: "XPBDMeshObject"
    filename: "{msh}"
    element-classes-filename: "{cls}"
    materials: ["Soft Tissue", "Stiffer Tissue"]
    max-size: 0.025
"""
    r = run(cfg, tmp_path, "--print-config", "--load-meshes")
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)["objects"][0]
    assert [m["name"] for m in o["materials"]] == ["Soft Tissue", "Stiffer Tissue"]
    assert o["materials"][1] == {"name": "Stiffer Tissue", "E": 3e5, "nu": 0.49, "density": 1030}
    assert o["elements"] == 40960 and o["elements_per_material"] == [34799, 6161]
    # independent count from the files: vertices of a class-1 element are class 1
    import xpbd_amd as X
    x, t, _ = X.read_msh(msh)
    c = X.read_element_classes(cls)
    v1 = np.zeros(len(x), bool)
    v1[t[c == 1].ravel()] = True
    assert o["vertices_per_class"] == [int((~v1).sum()), int(v1.sum())]


def test_classes_past_the_material_list_take_the_last_material(tmp_path):
    """One material for two classes: the reference warns and uses the last
    material for every class past the end of the list."""
    msh = _gunzip("2x1x1_block16.msh", tmp_path)
    cls = _gunzip("2x1x1_block16_element_classes.txt", tmp_path)
    cfg = PALPATION_MATERIALS + f"""
objects:
  - name: "Cube1"
    type: "XPBDMeshObject"
    filename: "{msh}"
    element-classes-filename: "{cls}"
    materials: ["Stiffer Tissue"]
"""
    r = run(cfg, tmp_path, "--print-config", "--load-meshes")
    assert r.returncode == 0, r.stderr
    assert "WARNING" in r.stderr
    assert json.loads(r.stdout)["objects"][0]["elements_per_material"] == [40960]


@pytest.mark.parametrize("bad,msg", [("materials: [\"Nope\"]", "no material named 'Nope'"),
                                     ("element-classes-filename: \"SHORT\"", "entries for 40960 elements")])
def test_material_and_class_errors(tmp_path, bad, msg):
    msh = _gunzip("2x1x1_block16.msh", tmp_path)
    short = tmp_path / "short.txt"
    short.write_text("0\n1\n1\nnot a number\n0\n")
    cfg = PALPATION_MATERIALS + f"""
objects:
  - name: "Cube1"
    type: "XPBDMeshObject"
    filename: "{msh}"
    {bad.replace("SHORT", str(short))}
"""
    r = run(cfg, tmp_path, "--print-config", "--load-meshes")
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr
