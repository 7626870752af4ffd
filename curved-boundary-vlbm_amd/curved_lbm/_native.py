"""ctypes binding of libvlbm.so, the HIP (gfx950) implementation of the time step.

The library is loaded from ``curved_lbm/_lib/libvlbm.so`` (built in-tree by
``build_native.py``).  There is no fallback: if the library or a GPU is
missing the solver refuses to construct.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libvlbm.so"

ABI_VERSION = 9
ABI2_VERSION = 8
PREC_F64 = 0
PREC_F32 = 1

RED_SUM_V2, RED_MAX_V, RED_MIN_J, RED_MAX_J, RED_MAX_U, RED_NONFINITE, RED_INVALID, RED_SUM_DU2 = range(8)
RED_COUNT = 8
RED_SCRATCH = 1024 * 8 + 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class NNGrid(C.Structure):
    """Mirror of ``vlbm_nn_grid`` in include/vlbm_nn.h."""

    _fields_ = [("batch", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32)]


class Vlbm3dDesc(C.Structure):
    """Mirror of ``vlbm3d_desc`` in include/vlbm3d.h (field order matters)."""

    _fields_ = [
        ("struct_size", C.c_uint32),
        ("abi_version", C.c_int32),
        ("n_active", C.c_int32),
        ("n_links", C.c_int32),
        ("precision", C.c_int32),
        ("model", C.c_int32),
        ("lam", C.c_double),
        ("mu", C.c_double),
        ("p0", C.c_double),
        ("p1", C.c_double),
        ("p2", C.c_double),
        ("j_floor", C.c_double),
        ("dt", C.c_double),
        ("c", C.c_double),
        ("omega", C.c_double),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("dz", C.c_double),
        ("source_mode", C.c_int32),
        ("source_params", C.c_double * 12),
        ("recon", C.c_int32),
        ("normal_solver", C.c_int32),
        ("disp_interval", C.c_int32),
        ("disp_sweeps", C.c_int32),
        ("disp_relax", C.c_double),
        ("disp_weight", C.c_double),
        ("disp_boundary_id", C.c_int32),
        ("compat_interval", C.c_int32),
        ("compat_blend", C.c_double),
        ("has_char", C.c_int32),
        ("n_anchor_links", C.c_int32),
        ("n_char_links", C.c_int32),
        ("n_char_nodes", C.c_int32),
        ("U", C.c_void_p),
        ("u", C.c_void_p),
        ("u_star", C.c_void_p),
        ("u_next", C.c_void_p),
        ("record", C.c_void_p),
        ("invalid", C.c_void_p),
        ("reduce_scratch", C.c_void_p),
        ("neighbor", C.c_void_p),
        ("active_i", C.c_void_p),
        ("active_j", C.c_void_p),
        ("active_k", C.c_void_p),
        ("node_link_off", C.c_void_p),
        ("node_link_ids", C.c_void_p),
        ("link_p", C.c_void_p),
        ("link_up", C.c_void_p),
        ("link_q", C.c_void_p),
        ("link_bid", C.c_void_p),
        ("link_eta", C.c_void_p),
        ("link_xb", C.c_void_p),
        ("link_yb", C.c_void_p),
        ("link_zb", C.c_void_p),
        ("link_nx", C.c_void_p),
        ("link_ny", C.c_void_p),
        ("link_nz", C.c_void_p),
        ("char_link", C.c_void_p),
        ("char_nodes", C.c_void_p),
        ("char_pack", C.c_void_p),
        ("char_pack_slot", C.c_void_p),
        ("char_gidx", C.c_void_p),
        ("char_gw", C.c_void_p),
        ("char_pidx", C.c_void_p),
        ("char_pw", C.c_void_p),
        ("char_state", C.c_void_p),
        ("tan_p", C.c_void_p),
        ("tan_w1", C.c_void_p),
        ("tan_w2", C.c_void_p),
        ("tan_valid", C.c_void_p),
        ("kind_by_id", C.c_void_p),
        ("comp_kind_by_id", C.c_void_p),
        ("mode_by_id", C.c_void_p),
        ("params_by_id", C.c_void_p),
        ("dyn_enabled", C.c_void_p),
        ("dyn_traction", C.c_void_p),
        ("lkind", C.c_void_p),
        ("lcomp_kind", C.c_void_p),
        ("lbv", C.c_void_p),
        ("lFb", C.c_void_p),
        ("lPb", C.c_void_p),
        ("box", C.c_int32),
        ("box_n", C.c_int32 * 3),
        ("box_side_kind", C.c_int32 * 6),
        ("box_value_mode", C.c_int32 * 6),
        ("box_params", C.c_double * 96),
        ("box_length", C.c_double * 3),
        ("box_speed", C.c_double * 3),
        ("stream", C.c_void_p),
    ]


class Vlbm2dDesc(C.Structure):
    """Mirror of ``vlbm2d_desc`` in include/vlbm2d.h (field order matters)."""

    _fields_ = [
        ("struct_size", C.c_uint32),
        ("abi_version", C.c_int32),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("n_active", C.c_int32),
        ("n_links", C.c_int32),
        ("precision", C.c_int32),
        ("model", C.c_int32),
        ("lam", C.c_double),
        ("mu", C.c_double),
        ("p0", C.c_double),
        ("p1", C.c_double),
        ("p2", C.c_double),
        ("j_floor", C.c_double),
        ("dt", C.c_double),
        ("c", C.c_double),
        ("omega", C.c_double),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("collision_model", C.c_int32),
        ("ghost_omega", C.c_double),
        ("kinetic_filter", C.c_double),
        ("source_mode", C.c_int32),
        ("source_params", C.c_double * 6),
        ("recon", C.c_int32),
        ("normal_solver", C.c_int32),
        ("disp_interval", C.c_int32),
        ("disp_sweeps", C.c_int32),
        ("disp_relax", C.c_double),
        ("disp_weight", C.c_double),
        ("compat_interval", C.c_int32),
        ("compat_blend", C.c_double),
        ("compat_interior_only", C.c_int32),
        ("n_anchor_links", C.c_int32),
        ("U", C.c_void_p),
        ("u", C.c_void_p),
        ("u_star", C.c_void_p),
        ("u_next", C.c_void_p),
        ("record", C.c_void_p),
        ("invalid", C.c_void_p),
        ("ghost", C.c_void_p),
        ("reduce_scratch", C.c_void_p),
        ("active", C.c_void_p),
        ("active_list", C.c_void_p),
        ("anchor_off", C.c_void_p),
        ("anchor_ids", C.c_void_p),
        ("link_p", C.c_void_p),
        ("link_q", C.c_void_p),
        ("link_bid", C.c_void_p),
        ("link_eta", C.c_void_p),
        ("link_xb", C.c_void_p),
        ("link_yb", C.c_void_p),
        ("link_nx", C.c_void_p),
        ("link_ny", C.c_void_p),
        ("lsq_p", C.c_void_p),
        ("lsq_wb", C.c_void_p),
        ("lsq_wx", C.c_void_p),
        ("lsq_wy", C.c_void_p),
        ("lsq_wt", C.c_void_p),
        ("lsq_valid", C.c_void_p),
        ("link_state", C.c_void_p),
        ("kind_by_id", C.c_void_p),
        ("mode_by_id", C.c_void_p),
        ("params_by_id", C.c_void_p),
        ("dyn_tx", C.c_void_p),
        ("dyn_ty", C.c_void_p),
        ("use_dyn", C.c_int32),
        ("box", C.c_int32),
        ("box_side_kind", C.c_int32 * 4),
        ("box_value_mode", C.c_int32 * 4),
        ("box_params", C.c_double * 32),
        ("length_x", C.c_double),
        ("length_y", C.c_double),
        ("stream", C.c_void_p),
    ]


EXPORTED = (
    "vlbm3d_create",
    "vlbm3d_destroy",
    "vlbm3d_step",
    "vlbm3d_set_graph_steps",
    "vlbm3d_refresh",
    "vlbm3d_apply_boundary",
    "vlbm3d_init_populations",
    "vlbm3d_set_ustar",
    "vlbm3d_reduce",
    "vlbm3d_reduce_noalloc",
    "vlbm3d_populations_to_f64",
    "vlbm3d_populations_from_f64",
    "vlbm3d_timing_enable",
    "vlbm3d_timing_read",
    "vlbm2d_create",
    "vlbm2d_destroy",
    "vlbm2d_step",
    "vlbm2d_refresh",
    "vlbm2d_apply_boundary",
    "vlbm2d_init_populations",
    "vlbm2d_set_ustar",
    "vlbm2d_stress",
    "vlbm2d_post_collision",
    "vlbm2d_kinetic_filter",
    "vlbm2d_set_box_sides",
    "vlbm2d_set_dynamic_traction",
    "vlbm2d_set_graph_steps",
    "vlbm2d_reduce",
    "vlbm2d_batch_create",
    "vlbm2d_batch_select",
    "vlbm2d_batch_set_graph_steps",
    "vlbm2d_batch_parity",
    "vlbm2d_batch_step",
    "vlbm2d_batch_destroy",
    "vlbm2d_batch_autotune",
    "vlbm2d_batch_set_blocks",
    "vlbm2d_populations_to_f64",
    "vlbm2d_populations_from_f64",
    "vlbm2d_timing_enable",
    "vlbm2d_timing_read",
    "vlbm_geom3d_char_weights",
    "vlbm_geom2d_lsq_weights",
    "vlbm_geom3d_tangent_weights",
    "vlbm_nn_conv_forward",
    "vlbm_nn_conv_backward_data",
    "vlbm_nn_conv_backward_weight",
    "vlbm_nn_loss",
    "vlbm_nn_pack_weights",
    "vlbm_nn_adam",
    "vlbm_last_error",
    "vlbm_abi_version",
)

SLOTS = ("collide", "sweep", "commit", "record", "stream", "gather", "prepare", "links", "step")

_lib = None


def load_library(path: Path | None = None) -> C.CDLL:
    """Load libvlbm.so once and declare its signatures.  Raises if missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"HIP library {p} is missing; build it with `python curved-boundary-vlbm_amd/build_native.py`"
        )
    # torch first: its HIP runtime is then the one libvlbm binds to (same
    # soname), so the process holds one HIP runtime and the library works on
    # torch's device memory and streams.  Loading libvlbm before torch would
    # pull in a second runtime that finds no device once torch owns it.
    import torch  # noqa: F401

    lib = C.CDLL(str(p))
    vp = C.c_void_p
    lib.vlbm_last_error.restype = C.c_char_p
    lib.vlbm_last_error.argtypes = []
    lib.vlbm_abi_version.restype = C.c_int
    lib.vlbm_abi_version.argtypes = []
    sigs = {
        "vlbm3d_create": [C.POINTER(Vlbm3dDesc), C.POINTER(vp)],
        "vlbm3d_destroy": [vp],
        "vlbm3d_step": [vp, vp, vp, C.c_double, C.c_int64, C.c_int32],
        "vlbm3d_set_graph_steps": [vp, C.c_int32],
        "vlbm3d_refresh": [vp, vp],
        "vlbm3d_apply_boundary": [vp, vp, vp, C.c_double, C.c_int32],
        "vlbm3d_init_populations": [vp, vp],
        "vlbm3d_set_ustar": [vp, vp],
        "vlbm3d_reduce": [vp, vp, _dp],
        "vlbm3d_reduce_noalloc": [vp, vp, vp],
        "vlbm3d_populations_to_f64": [vp, vp, vp],
        "vlbm3d_populations_from_f64": [vp, vp, vp],
        "vlbm3d_timing_enable": [vp, C.c_int32],
        "vlbm3d_timing_read": [vp, _dp, _ip, C.c_int32],
        "vlbm2d_create": [C.POINTER(Vlbm2dDesc), C.POINTER(vp)],
        "vlbm2d_destroy": [vp],
        "vlbm2d_step": [vp, vp, vp, C.c_double, C.c_int64, C.c_int32],
        "vlbm2d_refresh": [vp, vp],
        "vlbm2d_apply_boundary": [vp, vp, vp, C.c_double],
        "vlbm2d_init_populations": [vp, vp],
        "vlbm2d_set_ustar": [vp, vp],
        "vlbm2d_stress": [vp, vp, vp],
        "vlbm2d_post_collision": [vp, vp, vp],
        "vlbm2d_kinetic_filter": [vp, vp],
        "vlbm2d_set_box_sides": [vp, vp, vp, vp],
        "vlbm2d_set_dynamic_traction": [vp, C.c_int32],
        "vlbm2d_set_graph_steps": [vp, C.c_int32],
        "vlbm2d_reduce": [vp, vp, _dp],
        "vlbm2d_batch_create": [vp, vp, vp, C.c_int32, vp, C.POINTER(vp)],
        "vlbm2d_batch_select": [vp, vp, vp, vp, C.c_int32],
        "vlbm2d_batch_set_graph_steps": [vp, C.c_int32],
        "vlbm2d_batch_parity": [vp],
        "vlbm2d_batch_step": [vp, _dp, C.c_int32],
        "vlbm2d_batch_destroy": [vp],
        "vlbm2d_batch_autotune": [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float)],
        "vlbm2d_batch_set_blocks": [vp, C.c_int32, C.c_int32],
        "vlbm2d_populations_to_f64": [vp, vp, vp],
        "vlbm2d_populations_from_f64": [vp, vp, vp],
        "vlbm2d_timing_enable": [vp, C.c_int32],
        "vlbm2d_timing_read": [vp, _dp, _ip, C.c_int32],
        "vlbm_geom3d_char_weights": [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64] + [vp] * 4 + [vp] * 7
        + [C.c_double, C.c_int32, vp, vp, vp, vp],
        "vlbm_geom2d_lsq_weights": [vp, C.c_int32, C.c_int32, vp, vp, C.c_int64] + [vp] * 6
        + [C.c_double, C.c_int32, C.c_int32] + [vp] * 7,
        "vlbm_geom3d_tangent_weights": [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64] + [vp] * 9
        + [C.c_double] * 3 + [C.c_int32, C.c_int32] + [vp] * 4,
        "vlbm_nn_conv_forward": [vp, vp, vp, vp, C.c_int32, vp, vp, vp, vp],
        "vlbm_nn_conv_backward_data": [vp, vp, vp, vp, vp, vp],
        "vlbm_nn_conv_backward_weight": [vp, vp, vp, vp, C.c_int32, vp, vp, vp],
        "vlbm_nn_loss": [vp, vp, vp, vp, C.c_int32, C.c_float, vp, vp, C.c_int32, vp, vp],
        "vlbm_nn_pack_weights": [vp, vp, vp, vp],
        "vlbm_nn_adam": [C.c_int32, vp, vp, vp, vp] + [C.c_float] * 6 + [vp],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    if lib.vlbm_abi_version() != ABI_VERSION:
        raise RuntimeError("libvlbm ABI version mismatch; rebuild the library")
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load_library().vlbm_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what}: {msg}")


def host_threads() -> int:
    """Host threads for the geometry builders: the CPUs this process may
    use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a.size else None


def geom3d_char_weights(active_id, li, lj, lk, lq, xb, yb, zb, nx, ny, nz, eta, dx):
    """Native characteristic stencils (see include/vlbm_geom.h); returns
    (indices (64, L), weights (4, 64, L), modes (L), status (L))."""
    lib = load_library()
    aid = np.ascontiguousarray(active_id, dtype=np.int32)
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (li, lj, lk, lq)]
    dbl = [np.ascontiguousarray(a, dtype=np.float64) for a in (xb, yb, zb, nx, ny, nz, eta)]
    L = int(ints[0].size)
    indices = np.full((64, L), -1, np.int32)
    weights = np.zeros((4, 64, L), np.float64)
    modes = np.zeros(L, np.int32)
    status = np.zeros(L, np.uint8)
    rc = lib.vlbm_geom3d_char_weights(_ptr(aid), *[int(v) for v in aid.shape], L, *[_ptr(a) for a in ints],
                                      *[_ptr(a) for a in dbl], float(dx), host_threads(), _ptr(indices),
                                      _ptr(weights), _ptr(modes), _ptr(status))
    if rc != 0:
        raise ValueError("vlbm_geom3d_char_weights: invalid arguments")
    return indices, weights, modes, status


def geom2d_lsq_weights(active, x_coords, y_coords, li, lj, xb, yb, nx, ny, dx, max_stencil):
    """Native 2D boundary stencils (see include/vlbm_geom.h); returns
    (out_i, out_j, w, wt, wx, wy, valid) with (L, max_stencil) arrays."""
    lib = load_library()
    act = np.ascontiguousarray(active, dtype=np.int32)
    xc = np.ascontiguousarray(x_coords, dtype=np.float64)
    yc = np.ascontiguousarray(y_coords, dtype=np.float64)
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (li, lj)]
    dbl = [np.ascontiguousarray(a, dtype=np.float64) for a in (xb, yb, nx, ny)]
    L = int(ints[0].size)
    S = int(max_stencil)
    oi = np.full((L, S), -1, np.int32)
    oj = np.full((L, S), -1, np.int32)
    w, wt, wx, wy = (np.zeros((L, S)) for _ in range(4))
    valid = np.zeros(L, np.int32)
    rc = lib.vlbm_geom2d_lsq_weights(_ptr(act), int(act.shape[0]), int(act.shape[1]), _ptr(xc), _ptr(yc), L,
                                     *[_ptr(a) for a in ints], *[_ptr(a) for a in dbl], float(dx), S, host_threads(),
                                     *[_ptr(a) for a in (oi, oj, w, wt, wx, wy, valid)])
    if rc != 0:
        raise ValueError("vlbm_geom2d_lsq_weights: invalid arguments")
    return oi, oj, w, wt, wx, wy, valid


def geom3d_tangent_weights(active_id, li, lj, lk, xb, yb, zb, nx, ny, nz, dx, dy, dz, max_stencil):
    """Native compat_bfl tangent stencils (see include/vlbm_geom.h); returns
    (lsq_p, lsq_wt1, lsq_wt2) as (L, max_stencil) arrays and lsq_valid (L)."""
    lib = load_library()
    aid = np.ascontiguousarray(active_id, dtype=np.int32)
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (li, lj, lk)]
    dbl = [np.ascontiguousarray(a, dtype=np.float64) for a in (xb, yb, zb, nx, ny, nz)]
    L, S = int(ints[0].size), int(max_stencil)
    out_p = np.full((L, S), -1, np.int32)
    w1, w2 = np.zeros((L, S)), np.zeros((L, S))
    valid = np.zeros(L, np.int32)
    rc = lib.vlbm_geom3d_tangent_weights(_ptr(aid), *[int(v) for v in aid.shape], L, *[_ptr(a) for a in ints],
                                         *[_ptr(a) for a in dbl], float(dx), float(dy), float(dz), S,
                                         host_threads(), _ptr(out_p), _ptr(w1), _ptr(w2), _ptr(valid))
    if rc != 0:
        raise ValueError("vlbm_geom3d_tangent_weights: invalid arguments")
    return out_p, w1, w2, valid
