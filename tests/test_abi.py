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


def _split_args(s):
    """Top-level comma split of a call's argument text (template argument
    lists are not tracked: no call in the checked code has one)."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        if ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def _calls(code, prefix="xpbd_"):
    """(name, argument count) of every call of a function named prefix* in code."""
    found = []
    for m in re.finditer(r"\b(%s\w+)\s*\(" % prefix, code):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(code[i], 0)
            i += 1
        found.append((m.group(1), len(_split_args(code[m.end():i - 1]))))
    return found


def test_integration_shim_calls_match_the_header():
    """The C++ shim for the reference's tree and the
    ctypes stub call only functions the header declares, each with the
    header's number of arguments."""
    protos = _header_prototypes()
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```(?:cpp|c\+\+|python)\n(.*?)```", text, re.S)
    # comments dropped: they name functions without calling them
    calls = [c for b in blocks for c in _calls(re.sub(r"(//|#)[^\n]*", " ", b))]
    assert len(calls) >= 15, calls
    for name, nargs in calls:
        assert name in protos, f"{name} is not declared in xpbd_hip.h"
        assert nargs == len(protos[name][1]), (name, nargs, protos[name][1])


def test_integration_shim_aggregates_fit_the_structures():
    """The shim's brace initialisers of xpbd_mesh_desc and xpbd_params give no
    more values than the structures have fields, in the header's order of
    kinds (pointer fields get pointers or nullptr, counts get counts)."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    shim = re.search(r"```cpp\n(.*?)```", text, re.S).group(1)
    for struct, cls in BINDINGS.items():
        m = re.search(r"\b%s\s+\w+\s*\{(.*?)\};" % struct, shim, re.S)
        assert m, struct
        vals = [v.strip() for v in _split_args(re.sub(r"/\*.*?\*/", "", m.group(1)))]
        assert 0 < len(vals) <= len(cls._fields_), (struct, vals)
        for v, (fname, ftype) in zip(vals, cls._fields_):
            is_ptr = issubclass(ftype, (C._Pointer, C.c_void_p, C.c_char_p))
            looks_ptr = v == "nullptr" or v.endswith(".data()")
            assert is_ptr == looks_ptr, (struct, fname, v)
