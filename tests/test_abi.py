"""The C-ABI's structures as the C compiler lays them out against the ctypes
descriptions that call it: the binding the tests and the bench use
(xpbd_amd.py) and the ctypes stub INTEGRATION.md gives a maintainer.

A C probe built with gcc from include/xpbd_hip.h prints the offset of every
field the ctypes side names and the size of the structure; a field missing
from the header fails the build, a field added, moved or resized changes an
offset or the size.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

import xpbd_amd as X

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINDINGS = {"xpbd_mesh_desc": X.MeshDesc, "xpbd_params": X.Params}


def _c_layout(tmp_path, struct, names):
    src = tmp_path / f"probe_{struct}.c"
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "xpbd_hip.h"', "int main(void) {"]
    for n in names:
        lines.append(f'    printf("%s %zu\\n", "{n}", offsetof({struct}, {n}));')
    lines.append(f'    printf("sizeof %zu\\n", sizeof({struct}));')
    lines += ["    return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / f"probe_{struct}"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    return {k: int(v) for k, v in (l.split() for l in out if l)}


def _ctypes_layout(cls):
    lay = {name: getattr(cls, name).offset for name, _ in cls._fields_}
    lay["sizeof"] = C.sizeof(cls)
    return lay


@pytest.mark.parametrize("struct", sorted(BINDINGS))
def test_binding_matches_the_header(tmp_path, struct):
    cls = BINDINGS[struct]
    names = [n for n, _ in cls._fields_]
    assert _c_layout(tmp_path, struct, names) == _ctypes_layout(cls)


def _integration_stub():
    """The ctypes Structure classes of INTEGRATION.md's Python section."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = re.search(r"## 3\. ctypes.*?```python\n(.*?)```", text, re.S).group(1)
    classes = re.findall(r"^class \w+\(C\.Structure\):.*?\]\n", block, re.S | re.M)
    ns = {"C": C}
    for c in classes:
        exec(c, ns)
    return {"xpbd_mesh_desc": ns["MeshDesc"], "xpbd_params": ns["Params"]}


@pytest.mark.parametrize("struct", sorted(BINDINGS))
def test_integration_stub_matches_the_header(tmp_path, struct):
    cls = _integration_stub()[struct]
    names = [n for n, _ in cls._fields_]
    assert names == [n for n, _ in BINDINGS[struct]._fields_]
    assert _c_layout(tmp_path, struct, names) == _ctypes_layout(cls)
