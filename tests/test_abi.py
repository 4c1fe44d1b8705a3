"""The C-ABI library loads and exports every function include/cns.h declares;
host-only entry points behave (no device work here)."""
import ctypes as C
import os
import re

import pytest

from cudanavierstokes_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(cns_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS)


def test_config_struct_layout_matches_header():
    txt = open(os.path.join(ROOT, "include", "cns.h")).read()
    body = txt[txt.index("typedef struct cns_config {"):txt.index("} cns_config;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    decl = []
    for line in body.splitlines()[1:]:
        line = line.strip().rstrip(";")
        if not line:
            continue
        typ, rest = line.split(None, 1)
        for nm in rest.split(","):
            decl.append((nm.strip(), typ))
    fields = [(n, "int" if t is C.c_int else "double") for n, t in _lib.CnsConfig._fields_]
    assert decl == fields


def test_config_defaults():
    L = _lib.load()
    c = _lib.CnsConfig()
    assert L.cns_config_init(C.byref(c)) == 0
    assert (c.nx, c.ny, c.nz, c.stencil_size, c.scheme, c.precision) == (64, 64, 64, 2, 0, 0)
    assert c.cfl == 0.5 and c.reynolds == 1600.0


def test_create_rejects_bad_config_before_touching_device():
    L = _lib.load()
    c = _lib.CnsConfig()
    L.cns_config_init(C.byref(c))
    c.stencil_size = 5
    h = C.c_void_p()
    rc = L.cns_create(C.byref(c), C.byref(h))
    assert rc == -1
    assert b"stencil_size" in L.cns_last_error(None)
    c.stencil_size, c.stencil_visc = 2, 3
    assert L.cns_create(C.byref(c), C.byref(h)) == -1
    c.stencil_visc, c.nx = 2, 3
    assert L.cns_create(C.byref(c), C.byref(h)) == -1


def test_null_arguments_are_errors():
    L = _lib.load()
    assert L.cns_config_init(None) == -1
    assert L.cns_step(None, 1, None, None, None) == -1
    assert L.cns_destroy(None) == 0


def test_missing_library_raises(tmp_path):
    with pytest.raises(_lib.CnsError):
        _lib.load(str(tmp_path / "nope.so"))
