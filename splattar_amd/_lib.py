"""ctypes binding of libgs_hip.so (the C ABI declared in include/gs_hip.h).

Loading is strict: if the library is missing or a symbol is absent, import
of the product API raises -- there is no CPU fallback path.
"""

from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libgs_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "gs_hip.h")

GS_OK = 0
STATUS_NAMES = {0: "GS_OK", 1: "GS_ERR_INVALID", 2: "GS_ERR_CAPACITY", 3: "GS_ERR_OOM",
                4: "GS_ERR_HIP", 5: "GS_ERR_OVERFLOW", 6: "GS_ERR_NO_DEVICE", 7: "GS_ERR_IO"}

GS_BWD_ACCUMULATE = 1
GS_BWD_COLOR_ONLY = 2
GS_BWD_NO_CAMERA = 4
GS_STEP_PLAIN, GS_STEP_BATCHED, GS_STEP_GRAPH = 0, 1, 2
GS_ACT_IDENTITY, GS_ACT_EXP, GS_ACT_SIGMOID = 0, 1, 2
GS_BINNING_SORT64, GS_BINNING_PRESORT, GS_BINNING_TIGHT = 0, 1, 2


class GsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class TrackConfig(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("lr_rot", C.c_float), ("lr_center", C.c_float),
                ("min_improvement", C.c_float), ("patience", C.c_uint32)]


class TrackReport(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("coverage", C.c_float), ("best", C.c_float),
                ("last_quat", C.c_float * 4), ("last_center", C.c_float * 3)]


class RasterCamera(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("quat", C.c_float * 4), ("focal", C.c_float),
                ("sh_degree", C.c_uint32)]


class View(C.Structure):
    _fields_ = [("camera", RasterCamera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("num_surfels", C.c_uint32)]


class SceneInput(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("scales", C.c_void_p), ("quats", C.c_void_p),
                ("opacities", C.c_void_p), ("sh", C.c_void_p), ("sh_coeffs", C.c_uint32),
                ("n", C.c_uint32)]


class RasterBuffers(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "positions", "scales", "quats", "opacities", "sh", "out_color", "out_aux", "out_normal",
        "out_ncontrib", "dl_dcolor", "dl_dnormal", "grad_pos", "grad_scales", "grad_quat",
        "grad_opacity", "grad_sh", "grad_cam")]


class RasterDebug(C.Structure):
    _fields_ = [("sorted_keys", C.c_void_p), ("sorted_payloads", C.c_void_p), ("ranges", C.c_void_p),
                ("rects", C.c_void_p), ("depth_keys", C.c_void_p), ("offsets", C.c_void_p),
                ("num_tiles", C.c_uint32)]


class TrainView(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("camera", RasterCamera),
                ("target_rgba", C.c_void_p), ("num_surfels", C.c_uint32)]


class TrainConfig(C.Structure):
    _fields_ = [("lambda_", C.c_float), ("dssim_scale", C.c_float), ("lr_pos", C.c_float),
                ("lr_scales", C.c_float), ("lr_quat", C.c_float), ("lr_opacity", C.c_float),
                ("lr_sh", C.c_float), ("max_views_per_step", C.c_uint32),
                ("entries_per_view", C.c_uint32), ("digit_bits", C.c_uint32), ("binning", C.c_uint32),
                ("pos_lr_final_factor", C.c_float), ("iters", C.c_uint32)]


_vp = C.c_void_p
_u32 = C.c_uint32
_PROTOS = {
    "gs_last_error": ([], C.c_char_p),
    "gs_abi_version": ([], C.c_int),
    "gs_build_info": ([], C.c_char_p),
    "gs_ctx_create": ([C.c_int, C.POINTER(_vp)], C.c_int),
    "gs_ctx_destroy": ([_vp], None),
    "gs_ctx_stream": ([_vp], _vp),
    "gs_ctx_synchronize": ([_vp], C.c_int),
    "gs_ctx_device_info": ([_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)], C.c_int),
    "gs_buffer_alloc": ([_vp, C.c_size_t, C.POINTER(_vp)], C.c_int),
    "gs_buffer_free": ([_vp, _vp], C.c_int),
    "gs_write_buffer": ([_vp, _vp, _vp, C.c_size_t], C.c_int),
    "gs_read_buffer": ([_vp, _vp, _vp, C.c_size_t], C.c_int),
    "gs_clear_buffer": ([_vp, _vp, _vp, C.c_size_t], C.c_int),
    "gs_scan_create": ([_vp, _u32, C.POINTER(_vp)], C.c_int),
    "gs_scan_destroy": ([_vp], None),
    "gs_scan_data": ([_vp], _vp),
    "gs_scan_total": ([_vp], _vp),
    "gs_scan_encode": ([_vp, _vp, _u32], C.c_int),
    "gs_scan_encode_noalloc": ([_vp, _vp, _vp, _u32, _vp], C.c_int),
    "gs_sorter_create": ([_vp, _u32, _u32, C.POINTER(_vp)], C.c_int),
    "gs_sorter_destroy": ([_vp], None),
    "gs_sorter_keys": ([_vp], _vp),
    "gs_sorter_payloads": ([_vp], _vp),
    "gs_sorter_counts": ([_vp], _vp),
    "gs_sorter_encode": ([_vp, _vp, _u32, _u32, _u32], C.c_int),
    "gs_binner_create": ([_vp, _u32, _u32, _u32, C.POINTER(_vp)], C.c_int),
    "gs_binner_destroy": ([_vp], None),
    "gs_binner_rects": ([_vp], _vp),
    "gs_binner_depths": ([_vp], _vp),
    "gs_binner_ranges": ([_vp], _vp),
    "gs_binner_sorted_entries": ([_vp], _vp),
    "gs_binner_sorted_keys": ([_vp], _vp),
    "gs_binner_entry_count": ([_vp], _vp),
    "gs_binner_encode": ([_vp, _vp, _u32, _u32, _u32], C.c_int),
    "gs_binner_overflowed": ([_vp, C.POINTER(C.c_int)], C.c_int),
    "gs_raster_create": ([_vp, _u32, _u32, _u32, _u32, C.POINTER(_vp)], C.c_int),
    "gs_raster_destroy": ([_vp], None),
    "gs_raster_get_buffers": ([_vp, C.POINTER(RasterBuffers)], C.c_int),
    "gs_raster_upload_scene": ([_vp, C.POINTER(SceneInput)], C.c_int),
    "gs_raster_set_lambda_dist": ([_vp, C.c_float], C.c_int),
    "gs_raster_set_digit_bits": ([_vp, _u32], C.c_int),
    "gs_raster_set_binning": ([_vp, _u32], C.c_int),
    "gs_raster_set_culling": ([_vp, C.c_int], C.c_int),
    "gs_raster_set_trace": ([_vp, _u32, _u32, _u32, _u32, _u32], C.c_int),
    "gs_raster_trace_buffers": ([_vp, C.POINTER(_vp), C.POINTER(_vp)], C.c_int),
    "gs_raster_forward": ([_vp, _vp, C.POINTER(RasterCamera), _u32], C.c_int),
    "gs_raster_backward": ([_vp, _vp, _u32], C.c_int),
    "gs_raster_backward_ex": ([_vp, _vp, _u32, _u32], C.c_int),
    "gs_raster_create_batch": ([_vp, _u32, _u32, _u32, _u32, _u32, C.POINTER(_vp)], C.c_int),
    "gs_raster_forward_batch": ([_vp, _vp, C.POINTER(View), _u32], C.c_int),
    "gs_raster_backward_batch": ([_vp, _vp, _u32], C.c_int),
    "gs_raster_entry_count": ([_vp, C.POINTER(_u32)], C.c_int),
    "gs_raster_debug_buffers": ([_vp, C.POINTER(RasterDebug)], C.c_int),
    "gs_loss_create": ([_vp, _u32, _u32, C.c_float, _vp, _vp, C.POINTER(_vp)], C.c_int),
    "gs_loss_destroy": ([_vp], None),
    "gs_loss_target": ([_vp], _vp),
    "gs_loss_set_dssim_scale": ([_vp, C.c_float], C.c_int),
    "gs_loss_encode": ([_vp, _vp], C.c_int),
    "gs_loss_read": ([_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "gs_mcmc_add_noise": ([_vp, _vp, _u32, _u32, C.c_float, C.c_float, _vp, _vp, _vp], C.c_int),
    "gs_mcmc_sample_pairs": ([_vp, _vp, _u32, C.c_float, C.c_float, C.POINTER(C.c_uint64), _vp,
                              C.POINTER(_u32)], C.c_int),
    "gs_trainer_set_mcmc_noise": ([_vp, C.c_float], C.c_int),
    "gs_trainer_mcmc_relocate": ([_vp, C.c_float, C.c_float, C.POINTER(C.c_uint64), C.POINTER(_u32)], C.c_int),
    "gs_trainer_relocate_pairs": ([_vp, _vp, _u32], C.c_int),
    "gs_write_3dgs_ply": ([_vp, C.c_char_p, _u32, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "gs_ply_pack_rows": ([_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "gs_trainer_write_ply": ([_vp, C.c_char_p], C.c_int),
    "gs_tracker_create": ([_vp, _vp, _u32, _u32, C.POINTER(_vp)], C.c_int),
    "gs_tracker_destroy": ([_vp], None),
    "gs_tracker_target": ([_vp], _vp),
    "gs_tracker_set_target": ([_vp, _vp], C.c_int),
    "gs_tracker_track": ([_vp, _u32, C.POINTER(RasterCamera), C.c_float, C.POINTER(TrackConfig), _vp,
                          C.POINTER(TrackReport)], C.c_int),
    "gs_tracker_photometric": ([_vp, _vp, C.c_float, C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_double)],
                               C.c_int),
    "gs_normal_loss_create": ([_vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp, C.POINTER(_vp)], C.c_int),
    "gs_normal_loss_destroy": ([_vp], None),
    "gs_normal_loss_loss_map": ([_vp], _vp),
    "gs_normal_loss_set_lambda": ([_vp, C.c_float, C.c_float], C.c_int),
    "gs_normal_loss_encode": ([_vp, _vp], C.c_int),
    "gs_normal_loss_read": ([_vp, C.POINTER(C.c_double)], C.c_int),
    "gs_optim_create": ([_vp, C.POINTER(_vp)], C.c_int),
    "gs_optim_destroy": ([_vp], None),
    "gs_optim_add_class": ([_vp, C.c_char_p, _u32, C.c_int, C.c_float, _vp, _vp], C.c_int),
    "gs_optim_class_raw": ([_vp, C.c_char_p], _vp),
    "gs_optim_set_lr": ([_vp, C.c_char_p, C.c_float], C.c_int),
    "gs_optim_set_reg": ([_vp, C.c_char_p, C.c_float], C.c_int),
    "gs_optim_set_lr_schedule": ([_vp, C.c_char_p, C.c_float, C.c_float, _u32], C.c_int),
    "gs_optim_current_lr": ([_vp, C.c_char_p, C.POINTER(C.c_float)], C.c_int),
    "gs_optim_state_bytes": ([_vp, C.POINTER(C.c_uint64)], C.c_int),
    "gs_optim_export_state": ([_vp, _vp, C.c_uint64], C.c_int),
    "gs_optim_import_state": ([_vp, _vp, C.c_uint64], C.c_int),
    "gs_optim_encode_step": ([_vp, _vp], C.c_int),
    "gs_optim_encode_activate": ([_vp, _vp], C.c_int),
    "gs_optim_step_count": ([_vp, C.POINTER(_u32)], C.c_int),
    "gs_optim_class_count": ([_vp, C.POINTER(_u32)], C.c_int),
    "gs_optim_class_info": ([_vp, _u32, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64)], C.c_int),
    "gs_optim_encode_snapshot": ([_vp, _vp, C.POINTER(_vp), _u32], C.c_int),
    "gs_optim_encode_restore": ([_vp, _vp, C.POINTER(_vp), _u32], C.c_int),
    "gs_trainer_create": ([_vp, C.POINTER(TrainConfig), C.POINTER(SceneInput), _u32,
                           C.POINTER(TrainView), C.POINTER(_vp)], C.c_int),
    "gs_trainer_destroy": ([_vp], None),
    "gs_trainer_step": ([_vp, _vp, C.POINTER(_u32), _u32, _u32], C.c_int),
    "gs_trainer_autotune": ([_vp, C.POINTER(_u32), _u32, _u32, C.POINTER(_u32)], C.c_int),
    "gs_trainer_set_schedule": ([_vp, _u32, _u32], C.c_int),
    "gs_trainer_set_tiling": ([_vp, _u32, _u32, _u32], C.c_int),
    "gs_trainer_set_overlap": ([_vp, C.c_int], C.c_int),
    "gs_trainer_set_keep_grads": ([_vp, C.c_int], C.c_int),
    "gs_trainer_optim": ([_vp, C.POINTER(_vp)], C.c_int),
    "gs_trainer_state_bytes": ([_vp, C.POINTER(C.c_uint64)], C.c_int),
    "gs_trainer_export_state": ([_vp, _vp, C.c_uint64], C.c_int),
    "gs_trainer_import_state": ([_vp, _vp, C.c_uint64], C.c_int),
    "gs_trainer_last_losses": ([_vp, C.POINTER(C.c_double), _u32], C.c_int),
    "gs_trainer_render": ([_vp, _u32, _vp], C.c_int),
    "gs_trainer_params": ([_vp, C.POINTER(RasterBuffers)], C.c_int),
    "gs_trainer_entry_count": ([_vp, C.POINTER(_u32)], C.c_int),
    "gs_trainer_digit_bits": ([_vp, C.POINTER(_u32)], C.c_int),
    "gs_stage_name": ([_u32], C.c_char_p),
    "gs_trainer_set_profiling": ([_vp, C.c_int], C.c_int),
    "gs_trainer_stage_times": ([_vp, C.POINTER(C.c_double), C.POINTER(_u32)], C.c_int),
}
GS_STAGE_COUNT = 11

_lib = None


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libgs_hip.so once; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status: int):
    if status != GS_OK:
        msg = lib().gs_last_error()
        raise GsError(status, msg.decode() if msg else "")
