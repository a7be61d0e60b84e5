"""ctypes binding of libbsamgpu.so (the C ABI of include/bsamgpu.h).

The product path always goes through this library; when the shared object is
missing the import of ``load()`` raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .problem import BsamError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "csrc", "libbsamgpu.so"))

c_int = ctypes.c_int
c_double = ctypes.c_double
c_ll = ctypes.c_longlong
P_int = ctypes.POINTER(c_int)
P_double = ctypes.POINTER(c_double)
P_ll = ctypes.POINTER(c_ll)

BSAM_Q, BSAM_F, BSAM_FTRUE, BSAM_QC = 0, 1, 2, 3
BSAM_RESTRICT, BSAM_RING, BSAM_CORRECT, BSAM_FILLDOWN = 0, 1, 2, 3
BSAM_FP64, BSAM_MIXED = 0, 1

ERROR_NAMES = {
    -1: "BSAM_E_ARG", -2: "BSAM_E_ALIGN", -3: "BSAM_E_NESTING", -4: "BSAM_E_RING",
    -5: "BSAM_E_OVERHANG", -6: "BSAM_E_CORR", -7: "BSAM_E_COVER", -8: "BSAM_E_BOX",
    -9: "BSAM_E_D", -10: "BSAM_E_NONFINITE", -11: "BSAM_E_SINGULAR", -20: "BSAM_E_HIP",
    -21: "BSAM_E_NOMEM",
}

# names every build must export (kept in sync with include/bsamgpu.h)
EXPORTS = [
    "bsam_default_opts", "bsam_last_error", "bsam_version", "bsam_ctx_create", "bsam_ctx_destroy",
    "bsam_hier_create", "bsam_hier_destroy", "bsam_solve_cycles", "bsam_comp_residual", "bsam_vcycle",
    "bsam_fill_ghosts", "bsam_relax_level", "bsam_transfer", "bsam_coarse_load", "bsam_coarsest_solve",
    "bsam_reset_force", "bsam_init_level_prolong", "bsam_composite_mean", "bsam_level_field_size",
    "bsam_get_level", "bsam_set_level", "bsam_export_size", "bsam_export", "bsam_hier_reset",
    "bsam_time_smoother", "bsam_smoother_bytes", "bsam_level_streams", "bsam_update_coeffs", "bsam_parabolic_init",
    "bsam_parabolic_set_history", "bsam_parabolic_get_history", "bsam_parabolic_set_source",
    "bsam_parabolic_rhs", "bsam_parabolic_set_shift", "bsam_parabolic_advance", "bsam_parabolic_step", "bsam_batch_create",
    "bsam_batch_destroy", "bsam_batch_step", "bsam_tag_level", "bsam_tag_level_static", "bsam_polygon_fields",
    "bsam_hier_inherit", "bsam_regrid_level", "bsam_parabolic_uo_from_q", "bsam_cluster_tags", "bsam_cluster_tags_per_parent",
    "bsam_build_create", "bsam_build_destroy", "bsam_build_add_level", "bsam_build_desc", "bsam_build_info",
    "bsam_build_allowed",
]

BSAM_BE, BSAM_BDF2, BSAM_CN = 0, 1, 2
BSAM_TAG_ULAP, BSAM_TAG_RTE = 0, 1


FIELD_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.c_longlong, ctypes.POINTER(ctypes.c_double))


class Problem(ctypes.Structure):
    _fields_ = [("domain", c_double * 4), ("D", FIELD_FN), ("C", FIELD_FN), ("f", FIELD_FN),
                ("bc_type", c_int * 4), ("g", FIELD_FN * 4), ("u0", FIELD_FN), ("user", ctypes.c_void_p)]


class BuildOpts(ctypes.Structure):
    _fields_ = [("base_cells", c_int * 2), ("coarsest_cells", c_int), ("source_quad", c_int)]


class HierDesc(ctypes.Structure):
    _fields_ = [("nlev", c_int), ("root_idx", c_int), ("domain", c_double * 4),
                ("n", P_int), ("np", P_int), ("boxes", P_int), ("bc_code", P_int),
                ("Q", P_double), ("FTrue", P_double), ("DeW", P_double), ("DnS", P_double),
                ("CC", P_double), ("InvDiagS", P_double), ("g", P_double)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("npre", c_int), ("npost", c_int), ("omega", c_double), ("rel_tol", c_double),
                ("max_vcycles", c_int), ("ghost_interp", c_int), ("mass_correction", c_int),
                ("precision", c_int), ("inner_cycles", c_int), ("use_graph", c_int),
                ("coarse_cells", c_int), ("smoother", c_int), ("stream_rows", c_int)]


class StageStats(ctypes.Structure):
    _fields_ = [("amr_levels", c_int), ("nlev", c_int), ("vcycles", c_int), ("final_rel", c_double),
                ("visible_cells", c_ll), ("total_cells", c_ll), ("avg_factor", c_double),
                ("res_hist", P_double), ("res_hist_cap", c_int), ("res_hist_len", c_int),
                ("gpu_ms", c_double)]


class Residual(ctypes.Structure):
    _fields_ = [("linf", c_double), ("l2", c_double), ("fnorm", c_double), ("rel", c_double),
                ("ncell", c_ll)]


_lib = None


def load(path=None):
    """Load libbsamgpu.so (raises BsamError when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise BsamError("bsamgpu:lib", f"{p} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    vp = ctypes.c_void_p
    sig = {
        "bsam_default_opts": (None, [ctypes.POINTER(SolveOpts)]),
        "bsam_last_error": (ctypes.c_char_p, []),
        "bsam_version": (ctypes.c_char_p, []),
        "bsam_ctx_create": (c_int, [c_int, ctypes.POINTER(vp)]),
        "bsam_ctx_destroy": (None, [vp]),
        "bsam_hier_create": (c_int, [vp, ctypes.POINTER(HierDesc), ctypes.POINTER(SolveOpts), ctypes.POINTER(vp)]),
        "bsam_hier_destroy": (None, [vp]),
        "bsam_solve_cycles": (c_int, [vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(StageStats)]),
        "bsam_comp_residual": (c_int, [vp, ctypes.POINTER(Residual)]),
        "bsam_vcycle": (c_int, [vp]),
        "bsam_fill_ghosts": (c_int, [vp, c_int]),
        "bsam_relax_level": (c_int, [vp, c_int, c_int]),
        "bsam_transfer": (c_int, [vp, c_int, c_int]),
        "bsam_coarse_load": (c_int, [vp, c_int]),
        "bsam_coarsest_solve": (c_int, [vp]),
        "bsam_reset_force": (c_int, [vp]),
        "bsam_init_level_prolong": (c_int, [vp, c_int]),
        "bsam_composite_mean": (c_int, [vp, P_double, P_double]),
        "bsam_level_field_size": (c_ll, [vp, c_int, c_int]),
        "bsam_get_level": (c_int, [vp, c_int, c_int, P_double]),
        "bsam_set_level": (c_int, [vp, c_int, c_int, P_double]),
        "bsam_export_size": (c_ll, [vp, c_int, c_int]),
        "bsam_export": (c_int, [vp, c_int, c_int, P_ll]),
        "bsam_update_coeffs": (c_int, [vp, ctypes.POINTER(HierDesc)]),
        "bsam_parabolic_init": (c_int, [vp]),
        "bsam_parabolic_set_history": (c_int, [vp, P_double, P_double]),
        "bsam_parabolic_get_history": (c_int, [vp, c_int, P_double]),
        "bsam_parabolic_set_source": (c_int, [vp, P_double]),
        "bsam_parabolic_rhs": (c_int, [vp, c_double, c_int]),
        "bsam_parabolic_set_shift": (c_int, [vp, c_double]),
        "bsam_parabolic_advance": (c_int, [vp]),
        "bsam_parabolic_step": (c_int, [vp, c_double, c_int, ctypes.POINTER(SolveOpts),
                                        ctypes.POINTER(StageStats), P_double]),
        "bsam_batch_create": (c_int, [vp, c_int, ctypes.POINTER(vp)]),
        "bsam_batch_destroy": (None, [vp]),
        "bsam_batch_step": (c_int, [vp, ctypes.POINTER(vp), P_int, c_int, ctypes.POINTER(SolveOpts),
                                    ctypes.POINTER(Residual)]),
        "bsam_hier_reset": (c_int, [vp]),
        "bsam_time_smoother": (c_int, [vp, c_int, c_int, c_int, P_double]),
        "bsam_smoother_bytes": (c_double, [vp, c_int, c_int]),
        "bsam_level_streams": (c_int, [vp, c_int]),
        "bsam_tag_level": (c_int, [vp, c_int, c_int, c_double, c_int, vp, vp, P_ll]),
        "bsam_tag_level_static": (c_int, [vp, c_int, c_int, c_double, c_int, vp, vp, P_ll]),
        "bsam_hier_inherit": (c_int, [vp, vp]),
        "bsam_cluster_tags": (c_int, [vp, vp, c_int, c_int, c_int, c_double, P_int, c_int, P_int, P_int]),
        "bsam_cluster_tags_per_parent": (c_int, [vp, vp, c_int, c_int, P_int, c_int, c_int, c_double, P_int,
                                                 c_int, P_int, P_int]),
        "bsam_regrid_level": (c_int, [vp, c_int, vp]),
        "bsam_parabolic_uo_from_q": (c_int, [vp]),
        "bsam_polygon_fields": (c_int, [c_int, P_double, P_double, c_int, P_double, P_double, c_ll, c_double,
                                        P_double, P_double, P_double]),
        "bsam_build_create": (c_int, [ctypes.POINTER(Problem), ctypes.POINTER(BuildOpts), ctypes.POINTER(vp)]),
        "bsam_build_destroy": (None, [vp]),
        "bsam_build_add_level": (c_int, [vp, P_int, c_int]),
        "bsam_build_desc": (c_int, [vp, ctypes.POINTER(HierDesc)]),
        "bsam_build_info": (c_int, [vp, P_int, P_int, P_int]),
        "bsam_build_allowed": (c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = load().bsam_last_error().decode()
        ident = msg.split(":", 2)
        ident = ":".join(ident[:2]) if len(ident) >= 2 else ERROR_NAMES.get(rc, "bsamgpu")
        raise BsamError(ident, f"{what} failed ({ERROR_NAMES.get(rc, rc)}): {msg}")
    return rc


def dptr(a):
    return a.ctypes.data_as(P_double)


def iptr(a):
    return a.ctypes.data_as(P_int)


def default_opts():
    o = SolveOpts()
    load().bsam_default_opts(ctypes.byref(o))
    return o
