"""ctypes declarations of include/convector.h (libconvector_amd.so).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly when the built library is missing.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libconvector_amd.so")

CV_OK = 0
STATUS_NAMES = {
    0: "CV_OK", 1: "CV_ERR_INVALID_ARGUMENT", 2: "CV_ERR_IO", 3: "CV_ERR_PARSE",
    4: "CV_ERR_DEGENERATE_TRIANGLE", 5: "CV_ERR_UNKNOWN_MATERIAL",
    6: "CV_ERR_NOT_MULTIPLE_OF_16", 7: "CV_ERR_LIGHT_COUNT", 8: "CV_ERR_BVH_NO_SPLIT",
    9: "CV_ERR_NAN", 10: "CV_ERR_DEVICE", 11: "CV_ERR_OUT_OF_MEMORY", 12: "CV_ERR_NO_DEVICE",
    13: "CV_ERR_OUT_OF_RANGE",
}

CV_PTR_DEVICE = 1
CV_DENOISE_MATERIALIZE = 2
CV_NOALLOC_MAX_DEPTH = 44
CV_NOALLOC_SCRATCH_BYTES = 8192 + 16 * 1024 * 1024
CV_RENDER_COUNT = 2
CV_RENDER_TIMING = 4
CV_RENDER_GRAPH = 8
CV_RAY_SORT_NONE = 255
CV_TUNE_TOLERANCE = 1
CV_RNG_PIXEL = 0
CV_RNG_REF_PATCH = 1
CV_APPROX_PORTABLE = 0
CV_BUILDER_REFERENCE = 0
CV_BUILDER_BINS_32 = 32
CV_BUILDER_BINS_16 = 16
CV_NO_HIT = 0xFFFFFFFF

# every function the header declares
EXPORTS = [
    "cv_mesh_load_obj", "cv_mesh_from_triangles", "cv_mesh_num_triangles",
    "cv_mesh_copy_triangles", "cv_mesh_destroy", "cv_scene_create", "cv_scene_create_with_bvh",
    "cv_scene_set_camera", "cv_scene_set_tuning", "cv_scene_get_tuning", "cv_scene_get_bvh_info", "cv_scene_copy_bvh", "cv_scene_destroy",
    "cv_bvh_intersect_nearest", "cv_bvh_intersect_nearest_noalloc", "cv_renderer_create", "cv_renderer_render_patch_u8",
    "cv_renderer_accumulate_patch_f32", "cv_renderer_render_frame_u8",
    "cv_renderer_accumulate_frames_f32", "cv_renderer_buffer_f32_into_bitmap",
    "cv_renderer_get_stats", "cv_renderer_destroy", "cv_batch_create", "cv_batch_set_tuning",
    "cv_batch_get_tuning", "cv_batch_render", "cv_batch_autotune", "cv_batch_get_stats", "cv_batch_destroy",
    "cv_renderer_accumulate_aov_f32", "cv_denoiser_num_params", "cv_denoiser_create",
    "cv_denoiser_get_params", "cv_denoiser_set_params", "cv_denoiser_forward",
    "cv_denoiser_train_step", "cv_denoiser_prepare_input", "cv_denoiser_output_to_hdr",
    "cv_denoiser_destroy", "cv_server_create", "cv_server_add_scene", "cv_server_set_denoiser",
    "cv_server_set_tuning", "cv_server_submit", "cv_server_poll", "cv_server_wait",
    "cv_server_get_stats", "cv_server_reset_stats", "cv_server_destroy", "cv_trace_enable",
    "cv_trace_disable", "cv_trace_export",
    "cv_device_synchronize", "cv_last_error", "cv_status_string", "cv_version",
]


class CvError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        self.name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.name}: {message}")


class MaterialEntry(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("material", ctypes.c_uint32)]


class SceneOpts(ctypes.Structure):
    _fields_ = [("builder", ctypes.c_uint32), ("has_emission_override", ctypes.c_uint32),
                ("emission_override", ctypes.c_float * 3)]


class BvhInfo(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint32), ("num_triangles", ctypes.c_uint32),
                ("depth", ctypes.c_uint32), ("max_leaf", ctypes.c_uint32),
                ("num_direct", ctypes.c_uint32), ("avg_area_ratio", ctypes.c_float),
                ("avg_tris_per_leaf", ctypes.c_float), ("build_ms", ctypes.c_float),
                ("bins", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("stack_entries", ctypes.c_uint32), ("extend_waves_per_cu", ctypes.c_uint32),
                ("refill_threshold", ctypes.c_uint32), ("fetch_batches", ctypes.c_uint32),
                ("block_waves", ctypes.c_uint32), ("frames_per_step", ctypes.c_uint32),
                ("refill_interval", ctypes.c_uint32), ("leaf_batch", ctypes.c_uint32),
                ("bvh_width", ctypes.c_uint32), ("ray_sort_bits", ctypes.c_uint32),
                ("queue_parts", ctypes.c_uint32)]


class RenderOpts(ctypes.Structure):
    _fields_ = [("rng_mode", ctypes.c_uint32), ("approx_mode", ctypes.c_uint32),
                ("max_bounces", ctypes.c_uint32)]


class Job(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("first_frame", ctypes.c_uint32),
                ("num_frames", ctypes.c_uint32), ("hdr", ctypes.c_void_p),
                ("gbuffer", ctypes.c_void_p), ("features", ctypes.c_void_p)]


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("denoise", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("camera_pos", ctypes.c_float * 3), ("camera_q", ctypes.c_float * 4),
                ("screen_distance", ctypes.c_float), ("hdr", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("busy_ms", ctypes.c_double), ("peak_bytes", ctypes.c_uint64),
                ("resident_bytes", ctypes.c_uint64)]


class RenderStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("extend_rays", ctypes.c_uint64),
                ("extend_launches", ctypes.c_uint64), ("extend_ms", ctypes.c_double),
                ("shade_ms", ctypes.c_double), ("generate_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("sum_numi_aabb", ctypes.c_uint64),
                ("sum_numi_tri", ctypes.c_uint64), ("sort_ms", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__; __graft_entry__.build()')")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    P = ctypes.c_void_p
    U32 = ctypes.c_uint32
    SZ = ctypes.c_size_t
    I = ctypes.c_int
    F = ctypes.c_float
    PP = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "cv_mesh_load_obj": (I, [ctypes.c_char_p, P, SZ, PP]),
        "cv_mesh_from_triangles": (I, [P, P, P, SZ, PP]),
        "cv_mesh_num_triangles": (SZ, [P]),
        "cv_mesh_copy_triangles": (I, [P, P, P, P]),
        "cv_mesh_destroy": (None, [P]),
        "cv_scene_create": (I, [P, SZ, P, PP]),
        "cv_scene_create_with_bvh": (I, [P, SZ, P, P, SZ, P, PP]),
        "cv_scene_set_camera": (I, [P, P, P, P, P, F]),
        "cv_scene_set_tuning": (I, [P, P]),
        "cv_scene_get_tuning": (I, [P, P]),
        "cv_scene_get_bvh_info": (I, [P, P]),
        "cv_scene_copy_bvh": (I, [P, P, P]),
        "cv_scene_destroy": (None, [P]),
        "cv_bvh_intersect_nearest": (I, [P, P, SZ, P, P, U32]),
        "cv_bvh_intersect_nearest_noalloc": (I, [P, P, SZ, P, P]),
        "cv_renderer_create": (I, [P, U32, U32, P, PP]),
        "cv_renderer_render_patch_u8": (I, [P, P, P, U32, U32, U32, U32, U32]),
        "cv_renderer_accumulate_patch_f32": (I, [P, P, P, U32, U32, U32, U32, U32]),
        "cv_renderer_render_frame_u8": (I, [P, P, P, U32, U32]),
        "cv_renderer_accumulate_frames_f32": (I, [P, P, P, U32, U32, U32]),
        "cv_renderer_buffer_f32_into_bitmap": (I, [P, P, P, U32, U32]),
        "cv_renderer_get_stats": (I, [P, P]),
        "cv_renderer_destroy": (None, [P]),
        "cv_batch_create": (I, [P, ctypes.c_uint64, PP]),
        "cv_batch_set_tuning": (I, [P, P]),
        "cv_batch_get_tuning": (I, [P, P]),
        "cv_batch_render": (I, [P, P, SZ, U32]),
        "cv_batch_autotune": (I, [P, P, SZ, ctypes.c_uint64, U32, P, P]),
        "cv_batch_get_stats": (I, [P, P]),
        "cv_batch_destroy": (None, [P]),
        "cv_renderer_accumulate_aov_f32": (I, [P, P, P, U32, U32, U32]),
        "cv_denoiser_num_params": (SZ, []),
        "cv_denoiser_create": (I, [ctypes.c_uint64, PP]),
        "cv_denoiser_get_params": (I, [P, P, P]),
        "cv_denoiser_set_params": (I, [P, P]),
        "cv_denoiser_forward": (I, [P, P, P, U32, U32, U32, U32]),
        "cv_denoiser_train_step": (I, [P, P, P, U32, U32, U32, F, U32, P]),
        "cv_denoiser_prepare_input": (I, [P, P, U32, U32, U32, P, U32]),
        "cv_denoiser_output_to_hdr": (I, [P, U32, U32, U32, P, U32]),
        "cv_denoiser_destroy": (None, [P]),
        "cv_server_create": (I, [P, ctypes.c_uint64, ctypes.c_uint64, PP]),
        "cv_server_add_scene": (I, [P, P, P]),
        "cv_server_set_denoiser": (I, [P, P]),
        "cv_server_set_tuning": (I, [P, P]),
        "cv_server_submit": (I, [P, P, P]),
        "cv_server_poll": (I, [P, ctypes.c_uint64, P]),
        "cv_server_wait": (I, [P, ctypes.c_uint64, P]),
        "cv_server_get_stats": (I, [P, P]),
        "cv_server_reset_stats": (I, [P]),
        "cv_server_destroy": (None, [P]),
        "cv_trace_enable": (I, [SZ]),
        "cv_trace_disable": (I, []),
        "cv_trace_export": (I, [ctypes.c_char_p]),
        "cv_device_synchronize": (I, []),
        "cv_last_error": (ctypes.c_char_p, []),
        "cv_status_string": (ctypes.c_char_p, [I]),
        "cv_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def check(status: int):
    if status != CV_OK:
        msg = lib().cv_last_error().decode("utf-8", "replace")
        raise CvError(status, msg)
