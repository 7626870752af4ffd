"""The C ABI library loads and exports every entry point of
include/vlbm3d.h, include/vlbm2d.h, include/vlbm_geom.h and include/vlbm_nn.h;
descriptor layouts agree with the library; the network entry points reject
bad arguments before any launch.  No GPU work."""

import ctypes as C
import re
from pathlib import Path

import pytest

from curved_lbm import _native

INCLUDE = Path(__file__).resolve().parents[1] / "include"
HEADER = INCLUDE / "vlbm3d.h"
HEADER2 = INCLUDE / "vlbm2d.h"
HEADER_GEOM = INCLUDE / "vlbm_geom.h"
HEADER_NN = INCLUDE / "vlbm_nn.h"


def declared_functions():
    names = set()
    for h in (HEADER, HEADER2, HEADER_GEOM, HEADER_NN):
        names |= set(re.findall(r"^\s*(?:int|const char \*)\s*(vlbm\w*)\s*\(", h.read_text(), flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTED)


def test_descriptor_layout_matches_library():
    """vlbm3d_create checks struct_size before anything else; with all
    pointers null it must fail on the pointer check, not on the size."""
    lib = _native.load_library()
    d = _native.Vlbm3dDesc()
    d.struct_size = C.sizeof(_native.Vlbm3dDesc)
    d.abi_version = _native.ABI_VERSION
    d.n_active = 10
    d.disp_sweeps = 1
    h = C.c_void_p()
    rc = lib.vlbm3d_create(C.byref(d), C.byref(h))
    assert rc == -1
    assert b"null" in lib.vlbm_last_error()
    d.struct_size -= 8
    assert lib.vlbm3d_create(C.byref(d), C.byref(h)) == -1
    assert b"size mismatch" in lib.vlbm_last_error()


def test_descriptor2_layout_matches_library():
    lib = _native.load_library()
    d = _native.Vlbm2dDesc()
    d.struct_size = C.sizeof(_native.Vlbm2dDesc)
    d.abi_version = _native.ABI2_VERSION
    d.nx = d.ny = 8
    d.n_active = 10
    d.disp_sweeps = 1
    d.recon = 12
    h = C.c_void_p()
    assert lib.vlbm2d_create(C.byref(d), C.byref(h)) == -1
    assert b"null" in lib.vlbm_last_error()
    d.recon = 14
    assert lib.vlbm2d_create(C.byref(d), C.byref(h)) == -1
    assert b"unknown reconstruction" in lib.vlbm_last_error()
    d.recon = 12
    d.collision_model = 1
    d.ghost_omega = 1.0
    assert lib.vlbm2d_create(C.byref(d), C.byref(h)) == -1
    assert b"ghost" in lib.vlbm_last_error()
    d.collision_model = 0
    d.struct_size -= 8
    assert lib.vlbm2d_create(C.byref(d), C.byref(h)) == -1
    assert b"size mismatch" in lib.vlbm_last_error()


@pytest.mark.parametrize("header,tag,binding", [(HEADER, "vlbm3d_desc", "Vlbm3dDesc"),
                                                (HEADER2, "vlbm2d_desc", "Vlbm2dDesc")])
def test_header_struct_fields_match_binding(header, tag, binding):
    text = header.read_text()
    body = text[text.index("typedef struct {"): text.index("} " + tag + ";")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl or decl.startswith("typedef"):
            decl = decl.replace("typedef struct {", "").strip()
            if not decl:
                continue
        parts = decl.replace("*", " ").split()
        tail = " ".join(parts[1:]) if parts[0] not in ("const",) else " ".join(parts[2:])
        if parts[0] == "const":
            tail = " ".join(parts[2:])
        for item in tail.split(","):
            item = item.strip()
            if item:
                names.append(re.sub(r"\[.*\]", "", item).strip())
    assert names == [f[0] for f in getattr(_native, binding)._fields_]


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _native.load_library(tmp_path / "libvlbm.so")


def test_library_binds_to_the_hip_runtime_torch_loads():
    """load_library() imports torch before dlopen, so libvlbm resolves its
    HIP runtime to the one torch has loaded (a second runtime loaded first
    finds no device once torch owns it).  Checked in a fresh interpreter."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "from curved_lbm import _native\n"
            "assert 'torch' not in sys.modules\n"
            "_native.load_library()\n"
            "assert 'torch' in sys.modules\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "assert len(libs) == 1, libs\n" % str(_native.Path(_native.__file__).resolve().parents[1]))
    subprocess.run([sys.executable, "-c", code], check=True, timeout=300)


def test_nn_entry_points_reject_bad_arguments():
    lib = _native.load_library()
    g = _native.NNGrid(2, 4, 4, 4)
    bad = _native.NNGrid(0, 4, 4, 4)
    assert lib.vlbm_nn_conv_forward(C.byref(bad), 1, 1, 1, 0, 1, 1, None, None) == -1
    assert b"positive" in lib.vlbm_last_error()
    assert lib.vlbm_nn_conv_forward(C.byref(g), None, 1, 1, 0, 1, 1, None, None) == -1
    assert lib.vlbm_nn_conv_forward(C.byref(g), 1, 1, 1, 7, 1, 1, None, None) == -1
    assert b"mode" in lib.vlbm_last_error()
    assert lib.vlbm_nn_conv_forward(C.byref(g), 1, 1, 1, 1, None, None, None, None) == -1
    assert lib.vlbm_nn_conv_backward_weight(C.byref(g), 1, 1, 1, 0, 1, 1, None) == -1
    assert b"n_partials" in lib.vlbm_last_error()
    assert lib.vlbm_nn_loss(C.byref(g), 1, 1, 1, 16, 1.0, 1, 1, 1, 1, None) == -1
    assert b"n_out" in lib.vlbm_last_error()
    assert lib.vlbm_nn_adam(-1, None, None, None, None, 0, 0, 0, 0, 0, 0, None) == -1
    assert lib.vlbm_nn_adam(0, None, None, None, None, 0, 0, 0, 0, 0, 0, None) == 0
