"""ctypes binding of the C ABI in include/fornax_gpu.h.

The product path has no fallback: if ``libfornax_gpu.so`` is missing or no
GPU is visible, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os

from . import build as _build

_dp = C.POINTER(C.c_double)
_lib = None


class MeshDesc(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("xlim", C.c_double * 6), ("bc", C.c_int * 6),
                ("gamma", C.c_double), ("mhd", C.c_int), ("eta", C.c_double)]


class MeshView(C.Structure):
    _fields_ = [("W", C.c_void_p * 2), ("bf", C.c_void_p * 3), ("vstride", C.c_longlong),
                ("sx", C.c_longlong), ("sy", C.c_longlong), ("origin", C.c_longlong),
                ("n", C.c_int * 3), ("nh", C.c_int)]


class BlockIO(C.Structure):
    """fx_block_io: one block in the reference's own layouts (pointers)."""
    _fields_ = [("W", C.c_void_p), ("var_stride", C.c_longlong), ("eosvar", C.c_void_p),
                ("bf", C.c_void_p * 3), ("flux", C.c_void_p * 3), ("v_riemann", C.c_void_p * 3),
                ("Bv1", C.c_void_p * 3), ("Bv2", C.c_void_p * 3)]


class Bf2bcIO(C.Structure):
    """fx_bf2bc_io: the cell field of one block from its face fields."""
    _fields_ = [("W", C.c_void_p), ("var_stride", C.c_longlong), ("bf", C.c_void_p * 3),
                ("done", C.c_void_p)]


class CfBlock(C.Structure):
    """fx_cf_block: geometry of one block of a refined mesh and its device arrays."""
    _fields_ = [("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("dx", C.c_double * 3),
                ("local", C.c_int), ("flux", C.c_void_p * 3), ("emf", C.c_void_p * 3),
                ("edge_fidelity", C.c_void_p * 3)]


class CfLink(C.Structure):
    """fx_cf_link: one neighbour slot (coarser neighbour) of a local block."""
    _fields_ = [("block", C.c_int), ("neighbor", C.c_int), ("rel_level", C.c_int),
                ("type", C.c_int), ("neighbor_local", C.c_int), ("recv_start", C.c_int * 3),
                ("recv_end", C.c_int * 3), ("send_start", C.c_int * 3)]


class FxError(RuntimeError):
    pass


def lib_path():
    """The in-tree library, or the file named by FORNAX_GPU_LIB."""
    return os.environ.get("FORNAX_GPU_LIB") or _build.LIB


def load():
    """Load the gfx950 library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise FxError(f"{path} is missing: run fornaxamr_amd.build.build() (hipcc, gfx950) first")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.fx_last_error.restype = C.c_char_p
    lib.fx_status_message.restype = C.c_char_p
    lib.fx_status_message.argtypes = [C.c_int]
    lib.fx_version.restype = C.c_char_p
    lib.fx_mesh_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
    lib.fx_mesh_destroy.argtypes = [vp]
    lib.fx_mesh_set_stream.argtypes = [vp, vp]
    lib.fx_mesh_set_prim.argtypes = [vp, _dp]
    lib.fx_mesh_get_prim.argtypes = [vp, C.c_int, _dp]
    lib.fx_mesh_set_bf.argtypes = [vp, _dp, _dp, _dp]
    lib.fx_mesh_get_bf.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    lib.fx_calc_dt.argtypes = [vp, C.c_double, _dp]
    lib.fx_stage.argtypes = [vp, C.c_int, C.c_double]
    lib.fx_step_async.argtypes = [vp, C.c_double, C.c_double]
    lib.fx_step.argtypes = [vp, C.c_double, C.c_double, _dp]
    lib.fx_step_dt.argtypes = [vp, C.c_double]
    lib.fx_sync.argtypes = [vp]
    lib.fx_last_dt.argtypes = [vp, _dp]
    lib.fx_mesh_view_get.argtypes = [vp, C.POINTER(MeshView)]
    lib.fx_mesh_set_timing.argtypes = [vp, C.c_int]
    lib.fx_mesh_kernel_time.argtypes = [vp, C.c_int, _dp, C.POINTER(C.c_int)]
    lib.fx_ensemble_create.argtypes = [C.c_int, C.POINTER(MeshDesc), C.POINTER(vp)]
    lib.fx_ensemble_destroy.argtypes = [vp]
    lib.fx_ensemble_set_prim.argtypes = [vp, C.c_int, _dp]
    lib.fx_ensemble_get_prim.argtypes = [vp, C.c_int, C.c_int, _dp]
    lib.fx_ensemble_set_schedule.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.fx_ensemble_set_graph.argtypes = [vp, C.c_int]
    lib.fx_ensemble_step_async.argtypes = [vp, C.c_double, C.c_double]
    lib.fx_ensemble_sync.argtypes = [vp]
    lib.fx_ensemble_last_dt.argtypes = [vp, _dp]
    lib.fx_ensemble_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    lib.fx_ensemble_autotune.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int), _dp]
    lib.fx_adj_create.argtypes = [C.POINTER(C.c_int), _dp, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(vp)]
    lib.fx_adj_destroy.argtypes = [vp]
    lib.fx_adj_forward.argtypes = [vp, vp, vp, C.c_int, C.c_double, vp]
    lib.fx_adj_backward.argtypes = [vp, vp, vp, vp]
    lib.fx_flux_update_blocks.argtypes = [C.c_int, vp, C.c_int, vp, vp]
    lib.fx_flux_update_block.argtypes = [C.POINTER(BlockIO), C.c_int]
    lib.fx_flux_update_block_noalloc.argtypes = [C.c_int, C.POINTER(BlockIO), C.c_int, vp]
    lib.fx_bf2bc_blocks.argtypes = [C.c_int, vp, vp, vp]
    lib.fx_bf2bc_block.argtypes = [C.POINTER(Bf2bcIO)]
    lib.fx_flux_div_block.argtypes = [C.POINTER(_dp), _dp, C.c_double, C.c_int, _dp]
    lib.fx_cf_plan_create.argtypes = [C.c_int, C.POINTER(CfBlock), C.c_int, C.POINTER(CfLink),
                                      C.c_int, C.POINTER(vp)]
    lib.fx_cf_plan_info.argtypes = [vp, C.POINTER(C.c_int)]
    lib.fx_cf_plan_destroy.argtypes = [vp]
    lib.fx_cf_flux_send.argtypes = [vp, vp]
    lib.fx_cf_emf_send.argtypes = [vp, vp, vp]
    lib.fx_mesh_set_integrator.argtypes = [vp, C.c_int]
    lib.fx_mesh_set_riemann.argtypes = [vp, C.c_int]
    lib.fx_mesh_set_mhd_schedule.argtypes = [vp, C.c_int, C.c_int]
    lib.fx_reconstruct_mc_slope_batch.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    lib.fx_riemann_hllc_batch.argtypes = [C.c_int] + [vp] * 11 + [vp]
    lib.fx_riemann_hlld_batch.argtypes = [C.c_int] + [vp] * 12 + [vp]
    lib.fx_emf_upwind_batch.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    lib.fx_emf_upwind_block_batch.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.POINTER(vp), _dp,
                                              C.POINTER(vp), vp, vp, vp]
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise FxError(f"fornax_gpu error {rc}: {load().fx_last_error().decode()}")


# the exported symbols every build must provide (include/fornax_gpu.h)
EXPORTS = [
    "fx_mesh_create", "fx_mesh_destroy", "fx_mesh_set_stream", "fx_mesh_set_prim",
    "fx_mesh_get_prim", "fx_mesh_set_bf", "fx_mesh_get_bf", "fx_calc_dt", "fx_stage",
    "fx_step_async", "fx_step", "fx_step_dt", "fx_sync", "fx_last_dt", "fx_mesh_view_get",
    "fx_mesh_set_timing", "fx_mesh_kernel_time",
    "fx_reconstruct_mc_slope_batch", "fx_riemann_hllc_batch", "fx_riemann_hlld_batch",
    "fx_emf_upwind_batch",
    "fx_emf_upwind_block_batch",
    "fx_ensemble_create", "fx_ensemble_destroy", "fx_ensemble_set_prim", "fx_ensemble_get_prim",
    "fx_ensemble_set_schedule", "fx_ensemble_set_graph", "fx_ensemble_step_async",
    "fx_ensemble_sync", "fx_ensemble_last_dt", "fx_ensemble_info", "fx_ensemble_autotune",
    "fx_adj_create", "fx_adj_destroy", "fx_adj_forward", "fx_adj_backward",
    "fx_flux_update_blocks", "fx_flux_update_block", "fx_flux_update_block_noalloc",
    "fx_flux_div_block",
    "fx_bf2bc_blocks", "fx_bf2bc_block",
    "fx_mesh_set_integrator", "fx_mesh_set_riemann", "fx_mesh_set_mhd_schedule",
    "fx_cf_plan_create", "fx_cf_plan_info", "fx_cf_plan_destroy", "fx_cf_flux_send",
    "fx_cf_emf_send",
    "fx_last_error", "fx_status_message", "fx_version",
]
