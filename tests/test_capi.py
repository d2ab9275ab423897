"""The C-ABI library loads on a machine without a GPU and exports every function that
include/paddi.h declares (no compute calls here)."""
import ctypes
import os
import re

from paddi_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(paddi_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_the_runner_api():
    names = declared_functions()
    for n in ("paddi_create", "paddi_destroy", "paddi_range_discords", "paddi_discords",
              "paddi_discords_batch", "paddi_compute_statistics", "paddi_status_str"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_layer_binds_all_exports():
    L = engine.load_library()
    for n in engine.EXPORTS:
        assert hasattr(L, n)
    assert set(engine.EXPORTS) <= set(declared_functions())


def test_status_strings_without_device():
    L = engine.load_library()
    assert L.paddi_status_str(0) == b"ok"
    assert L.paddi_status_str(1) == b"invalid argument"
    assert b"gfx950" in L.paddi_version()


def test_create_fails_cleanly_without_device():
    import torch
    if torch.cuda.device_count() > 0:  # counts without initialising a device
        return
    L = engine.load_library()
    p = ctypes.c_void_p()
    st = L.paddi_create(0, ctypes.byref(p))
    assert st != 0 and not p


def _header_struct(name):
    """Fields (type, name, count) of `typedef struct { ... } name;` in include/paddi.h, in
    order (pointers as type "ptr")."""
    txt = open(os.path.join(ROOT, "include", "paddi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    m = re.search(r"typedef struct\s*\{([^}]*)\}\s*" + name + r"\s*;", txt)
    assert m, name
    out = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        f = re.fullmatch(r"(?:const )?(u?int(?:32|64)_t|float|double|char)\s*(\*?)\s*"
                         r"([a-z_0-9]+)(?:\[(\d+)\])?", decl)
        assert f, (name, decl)
        t, ptr, n, k = f.groups()
        out.append(("ptr" if ptr else t, n, int(k) if k else 1))
    return out


def test_python_structs_match_the_header():
    """The ctypes mirrors of the ABI's structs have the header's fields, order and widths (a
    field added to the header and not to the binding would shift every field after it)."""
    width = {"int32_t": 4, "uint32_t": 4, "int64_t": 8, "uint64_t": 8, "float": 4, "double": 8,
             "ptr": 8}
    for cname, py in (("paddi_discord", engine.Discord), ("paddi_opts", engine.Opts),
                      ("paddi_stats", engine.Stats), ("paddi_tunables", engine.Tunables)):
        fields = _header_struct(cname)
        assert [n for _, n, _ in fields] == [f[0] for f in py._fields_], cname
        assert [width[t] * k for t, _, k in fields] == [ctypes.sizeof(f[1]) for f in py._fields_], cname
        size, align = 0, 1
        for t, _, k in fields:  # C layout: each field at a multiple of its width
            w = width[t]
            size = (size + w - 1) // w * w + w * k
            align = max(align, w)
        assert ctypes.sizeof(py) == (size + align - 1) // align * align, cname
