"""ctypes binding of libyune_hip.so (include/yune_hip.h, include/yune_types.h).

The numpy dtypes below are byte-for-byte the reference's host/device
structs (include/CL_headers.h:57-107 of the reference); arrays of them are
handed to the library without conversion.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyune_hip.so")

# ---- data contract -------------------------------------------------------
TRI_DTYPE = np.dtype(
    [("v1", "<f4", 4), ("v2", "<f4", 4), ("v3", "<f4", 4), ("vn1", "<f4", 4), ("vn2", "<f4", 4),
     ("vn3", "<f4", 4), ("matID", "<i4"), ("pad", "<f4", 3)], align=True)
MAT_DTYPE = np.dtype(
    [("ke", "<f4", 4), ("kd", "<f4", 4), ("ks", "<f4", 4), ("n", "<f4"), ("k", "<f4"), ("px", "<f4"),
     ("py", "<f4"), ("alpha_x", "<f4"), ("alpha_y", "<f4"), ("is_specular", "<i4"),
     ("is_transmissive", "<i4")], align=True)
NODE_DTYPE = np.dtype(
    [("p_min", "<f4", 4), ("p_max", "<f4", 4), ("vert_list", "<i4", 10), ("child_idx", "<i4"),
     ("vert_len", "<i4")], align=True)
assert TRI_DTYPE.itemsize == 112 and MAT_DTYPE.itemsize == 80 and NODE_DTYPE.itemsize == 80

F4 = C.c_float * 4


class Cam(C.Structure):
    _fields_ = [("r1", F4), ("r2", F4), ("r3", F4), ("r4", F4), ("view_plane_dist", C.c_float),
                ("pad", C.c_float * 3)]


class LightQuad(C.Structure):
    _fields_ = [("pos", F4), ("normal", F4), ("ke", F4), ("edge_l", F4), ("edge_w", F4)]


class RenderParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("gi", C.c_int32), ("mis", C.c_int32), ("max_bounces", C.c_int32),
                ("rr_threshold", C.c_int32), ("bdpt_bounces", C.c_int32), ("oren_nayar", C.c_int32),
                ("fix_mis", C.c_int32), ("background", C.c_float * 4)]


class Config(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_paths", C.c_int32), ("mem_budget", C.c_size_t),
                ("bvh_width", C.c_int32), ("trace_block", C.c_int32), ("trace_chunk", C.c_int32),
                ("trace_grid", C.c_int32), ("hoist_frac", C.c_float), ("leaf_max", C.c_int32),
                ("lds_stack", C.c_int32), ("shade_waves", C.c_int32), ("batch_factor", C.c_int32),
                ("trace_overlap", C.c_int32), ("graph", C.c_int32), ("trace_xcd", C.c_int32),
                ("trace_refill", C.c_int32), ("bvh_build", C.c_int32), ("serve_mode", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("samples", C.c_uint64), ("rays_ext", C.c_uint64),
                ("rays_shadow", C.c_uint64), ("nodes_ext", C.c_uint64), ("tris_ext", C.c_uint64),
                ("nodes_shadow", C.c_uint64), ("tris_shadow", C.c_uint64),
                ("trace_launches_ext", C.c_uint64), ("trace_launches_shadow", C.c_uint64),
                ("iterations", C.c_uint64), ("ms_total", C.c_double), ("ms_trace_ext", C.c_double),
                ("ms_trace_shadow", C.c_double), ("ms_shade", C.c_double), ("ms_other", C.c_double),
                ("groups", C.c_uint64), ("paths_in_flight", C.c_uint64), ("refills", C.c_uint64),
                ("plan_stalls", C.c_uint64), ("ms_host", C.c_double)]


class SceneInfo(C.Structure):
    _fields_ = [("n_tri", C.c_int32), ("n_ref_nodes", C.c_int32), ("n_gpu_nodes", C.c_int32),
                ("width", C.c_int32), ("max_depth", C.c_int32), ("n_hoist", C.c_int32), ("node_bytes", C.c_size_t),
                ("tri_bytes", C.c_size_t), ("shade_bytes", C.c_size_t)]


class CamCtl(C.Structure):
    _fields_ = [("side", F4), ("up", F4), ("look_at", F4), ("eye", F4), ("view2world", F4 * 4), ("y_fov", C.c_float),
                ("rotation_speed", C.c_float), ("move_speed", C.c_float), ("view_plane_dist", C.c_float)]


class TuneResult(C.Structure):
    _fields_ = [("bvh_width", C.c_int32), ("trace_block", C.c_int32), ("trace_chunk", C.c_int32),
                ("lds_stack", C.c_int32), ("shade_waves", C.c_int32), ("n_tried", C.c_int32),
                ("batch_paths", C.c_int64), ("batch_factor", C.c_int32), ("trace_overlap", C.c_int32),
                ("graph", C.c_int32), ("trace_xcd", C.c_int32), ("trace_refill", C.c_int32), ("pad", C.c_int32),
                ("mrays_per_s", C.c_double)]


class JobStatus(C.Structure):
    _fields_ = [("frames_target", C.c_int32), ("frames_done", C.c_int32), ("admitted", C.c_int32),
                ("done", C.c_int32), ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double)]


class HostScene(C.Structure):
    # yn_light_quad is 16-byte aligned in C: explicit pad before it
    _fields_ = [("tris", C.c_void_p), ("n_tri", C.c_int32), ("mats", C.c_void_p), ("n_mat", C.c_int32),
                ("_pad0", C.c_int32), ("light", LightQuad), ("cam", Cam)]


assert C.sizeof(Cam) == 80 and C.sizeof(LightQuad) == 80
assert HostScene.light.offset == 32 and HostScene.cam.offset == 112 and C.sizeof(HostScene) == 192

YN_MODE_PT, YN_MODE_BDPT = 0, 1
TONEMAP_REINHARD, TONEMAP_HABLE = 0, 1
STATUS = {0: "YN_OK", -1: "YN_ERR_HIP", -2: "YN_ERR_INVALID", -3: "YN_ERR_NOT_FOUND", -4: "YN_ERR_NO_MEMORY",
          -5: "YN_ERR_IO", -6: "YN_ERR_NO_DEVICE"}

# name -> (restype, argtypes).  Every symbol include/yune_hip.h declares.
P = C.c_void_p
I32 = C.c_int32
PI32 = C.POINTER(C.c_int32)
PF = C.POINTER(C.c_float)
SIGNATURES = {
    "yn_params_reference": (None, [C.POINTER(RenderParams), I32]),
    "yn_create": (I32, [C.POINTER(Config), C.POINTER(P)]),
    "yn_destroy": (None, [P]),
    "yn_last_error": (C.c_char_p, [P]),
    "yn_device_info": (I32, [P, C.c_char_p, C.c_int, PI32, C.POINTER(C.c_size_t)]),
    "yn_bvh_build": (I32, [P, I32, I32, C.POINTER(P), PI32]),
    "yn_bvh_build_gpu": (I32, [P, I32, I32, C.POINTER(P), PI32]),
    "yn_free": (None, [P]),
    "yn_load_obj": (I32, [C.c_char_p, C.POINTER(HostScene)]),
    "yn_gen_cornell": (I32, [C.POINTER(HostScene)]),
    "yn_gen_icosphere_room": (I32, [C.c_uint32, I32, I32, C.POINTER(HostScene)]),
    "yn_host_scene_free": (None, [C.POINTER(HostScene)]),
    "yn_camera": (I32, [PF, PF, PF, PF, C.c_float, C.POINTER(Cam)]),
    "yn_scene_create": (I32, [P, P, I32, P, I32, P, I32, C.POINTER(LightQuad), PI32]),
    "yn_scene_release": (I32, [P, I32]),
    "yn_scene_info_get": (I32, [P, I32, C.POINTER(SceneInfo)]),
    "yn_pathtracer": (I32, [P, P, P, C.POINTER(Cam), I32, I32, I32, C.POINTER(RenderParams), I32, I32,
                            C.c_uint32, I32, I32, I32]),
    "yn_tonemap": (I32, [P, P, P, I32, I32]),
    "yn_tonemap_op": (I32, [P, P, P, I32, I32, I32]),
    "yn_job_set_tonemap": (I32, [P, I32, I32]),
    "yn_job_create": (I32, [P, I32, C.POINTER(Cam), I32, I32, C.POINTER(RenderParams), C.c_uint32, PI32]),
    "yn_job_render": (I32, [P, I32, I32]),
    "yn_job_read": (I32, [P, I32, P]),
    "yn_job_read_tonemapped": (I32, [P, I32, P]),
    "yn_job_save": (I32, [P, I32, C.c_char_p]),
    "yn_job_release": (I32, [P, I32]),
    "yn_job_stats": (I32, [P, I32, C.POINTER(Stats)]),
    "yn_job_set_counting": (I32, [P, I32, I32]),
    "yn_job_capture": (I32, [P, I32, P, P, P, P, P]),
    "yn_trace": (I32, [P, I32, P, I32, P, P, P]),
    "yn_trace_noalloc": (I32, [P, I32, P, I32, P, P, P]),
    "yn_submit": (I32, [P, I32, C.POINTER(Cam), I32, I32, C.POINTER(RenderParams), C.c_uint32, I32, PI32]),
    "yn_serve": (I32, [P, I32, PI32]),
    "yn_job_status_get": (I32, [P, I32, C.POINTER(JobStatus)]),
    "yn_serve_stats": (I32, [P, C.POINTER(Stats)]),
    "yn_clock_ms": (C.c_double, [P]),
    "yn_fmath_eval": (I32, [P, I32, P, P, I32, P]),
    "yn_autotune": (I32, [P, I32, C.POINTER(Cam), I32, I32, C.POINTER(RenderParams), I32, C.POINTER(TuneResult)]),
    "yn_config_get": (I32, [P, C.POINTER(Config), C.POINTER(C.c_int64)]),
    "yn_camctl_init": (I32, [C.c_float, C.c_float, C.c_float, C.POINTER(CamCtl)]),
    "yn_camctl_set_view": (I32, [C.POINTER(CamCtl), PF, PF, PF, PF]),
    "yn_camctl_orient": (I32, [C.POINTER(CamCtl), PF, C.c_float, C.c_float]),
    "yn_camctl_pack": (I32, [C.POINTER(CamCtl), C.POINTER(Cam)]),
}

_lib = None


class YuneError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libyune_hip.so (built in-tree by __graft_entry__.build()).

    There is no fallback: without the library every entry point raises.
    """
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C yune_amd/csrc)")
        # torch ships its own libamdhip64.so.7 and its libraries name that copy.
        # Loaded first, it also serves this library (same soname); loaded
        # after ROCm's copy, it would be a second HIP runtime in the process
        # and torch's device calls would fail.  One runtime for both.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int, ctx=None) -> None:
    if code != 0:
        msg = lib().yn_last_error(ctx)
        raise YuneError(code, msg.decode() if msg else "")


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)
