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


# C parameter types of the header -> the ctypes type that passes them
_CTYPES = {
    "int32_t": C.c_int32, "int64_t": C.c_int64, "uint64_t": C.c_uint64, "double": C.c_double, "float": C.c_float,
    "int": C.c_int,
    "int32_t*": C.POINTER(C.c_int32), "int64_t*": C.POINTER(C.c_int64), "double*": C.POINTER(C.c_double),
    "float*": C.POINTER(C.c_float), "uint8_t*": C.POINTER(C.c_uint8), "char*": C.c_char_p,
    "xpbd_handle*": C.c_void_p, "xpbd_batch*": C.c_void_p, "xpbd_mlp*": C.c_void_p,
    "xpbd_handle**": C.POINTER(C.c_void_p), "xpbd_batch**": C.POINTER(C.c_void_p), "xpbd_mlp**": C.POINTER(C.c_void_p),
    "xpbd_mesh_desc*": C.POINTER(X.MeshDesc), "xpbd_params*": C.POINTER(X.Params),
}


def _header_prototypes():
    text = open(os.path.join(ROOT, "include", "xpbd_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    protos = {}
    for ret, name, args in re.findall(r"([A-Za-z_][\w\s\*]*?)\b(xpbd_\w+)\s*\(([^;{]*?)\)\s*;", text):
        def canon(decl):
            decl = re.sub(r"\bconst\b", " ", decl)
            stars = decl.count("*")
            words = decl.replace("*", " ").split()
            return words[0] + "*" * stars if words else ""
        # each parameter: its type, then its name
        params = [] if args.strip() in ("", "void") else [canon(re.match(r"(.*?)\w+\s*$", a.strip()).group(1))
                                                           for a in args.split(",")]
        # the return type: the stars and the last word before the name
        toks = ret.replace("*", " * ").split()
        stars, word = 0, ""
        for t in reversed(toks):
            if t == "*":
                stars += 1
            elif t != "const":
                word = t
                break
        protos[name] = (word + "*" * stars, params)
    return protos


def test_binding_signatures_match_the_header():
    """Every function the binding registers has the header's parameter list
    (count and C types) and return type."""
    protos = _header_prototypes()
    checked = 0
    for name, (ret, params) in sorted(protos.items()):
        fn = getattr(X.lib, name, None)
        if fn is None or fn.argtypes is None:
            continue
        assert [_CTYPES[p] for p in params] == list(fn.argtypes), name
        assert fn.restype == (C.c_char_p if ret == "char*" else _CTYPES[ret]), name
        checked += 1
    assert checked >= 60, checked
