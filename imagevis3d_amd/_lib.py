"""ctypes bindings of libtvkhip.so (the C ABI declared in include/tvk_hip.h).

The library is built in-tree by ``imagevis3d_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback for the renderer.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# IMAGEVIS3D_AMD_LIB selects another build of the same library (A/B runs)
LIB_PATH = os.environ.get("IMAGEVIS3D_AMD_LIB") or os.path.join(HERE, "libtvkhip.so")

TVK_OK = 0
ERRORS = {
    -1: "TVK_ERR_ARG", -2: "TVK_ERR_HIP", -3: "TVK_ERR_STATE", -4: "TVK_ERR_IO",
    -5: "TVK_ERR_OOM", -6: "TVK_ERR_FORMAT", -7: "TVK_ERR_CHECKSUM",
}


class TvkError(RuntimeError):
    """A failed call of the C ABI (Tuvok's bool-false / LuaError)."""

    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.func = func


class FrameStats(C.Structure):
    _fields_ = [
        ("evaluated_samples", C.c_uint64), ("nominal_samples", C.c_uint64),
        ("bricks_total", C.c_uint32), ("bricks_skipped", C.c_uint32),
        ("lod", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
        ("kernel_ms", C.c_float), ("paint_ms", C.c_float),
    ]


class MemStats(C.Structure):
    _fields_ = [
        ("gpu_mem", C.c_uint64), ("gpu_budget", C.c_uint64),
        ("allocated_gpu_mem", C.c_uint64), ("pool_slots", C.c_uint64),
        ("resident_bricks", C.c_uint64), ("slot_bytes", C.c_uint64),
        ("bricks_uploaded", C.c_uint64), ("bricks_evicted", C.c_uint64),
    ]


class RegionCheck(C.Structure):
    """tvk_region_check: the per-frame check of the last per-region frame."""
    _fields_ = [("rmse_estimate", C.c_float), ("fell_back", C.c_uint32), ("check_ms", C.c_float),
                ("reserved", C.c_uint32), ("check_samples", C.c_uint64)]


class MemManStats(C.Structure):
    """tvk_memman_stats: the device's memory manager as a whole."""
    _fields_ = [(n, C.c_uint64) for n in (
        "gpu_mem", "gpu_budget", "gpu_allocated", "cpu_mem", "cpu_budget", "cpu_allocated",
        "pool_bytes", "pool_slots", "pool_resident", "pool_evictions", "chunks_freed",
        "cache_bytes", "cache_hits", "cache_misses")]


class View(C.Structure):
    """tvk_view: one view of a batched launch."""
    _fields_ = [
        ("view_pos", C.c_float * 3), ("view_dir", C.c_float * 3), ("up_dir", C.c_float * 3),
        ("fov_deg", C.c_float), ("ortho", C.c_int32), ("ortho_half_height", C.c_float),
        ("width", C.c_uint32), ("height", C.c_uint32), ("tf_slot", C.c_uint32),
        ("rgba", C.c_void_p),
    ]


class ViewStats(C.Structure):
    _fields_ = [
        ("evaluated_samples", C.c_uint64), ("nominal_samples", C.c_uint64),
        ("lod", C.c_uint32), ("reserved", C.c_uint32),
        ("gathered_samples", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("brick_size", C.c_uint32), ("cell", C.c_uint32), ("samples_per_batch", C.c_uint32),
        ("tile_width", C.c_uint32), ("tile_group", C.c_uint32), ("xcd_split", C.c_uint32),
        ("kernel_ms", C.c_float), ("baseline_ms", C.c_float), ("candidates", C.c_uint32),
        ("wg_waves", C.c_uint32), ("brick_staging", C.c_uint32),
    ]


class ClearView(C.Structure):
    """tvk_clearview."""
    _fields_ = [("enabled", C.c_int32), ("iso_value", C.c_float), ("color", C.c_float * 3),
                ("size", C.c_float), ("context_scale", C.c_float), ("border_scale", C.c_float),
                ("focus", C.c_int32 * 2)]


class TocInfo(C.Structure):
    """tvk_toc_info."""
    _fields_ = [
        ("dims", C.c_uint64 * 3), ("bits", C.c_uint32), ("max_brick", C.c_uint32),
        ("overlap", C.c_uint32), ("n_lod", C.c_uint32), ("use_median", C.c_uint32),
        ("clamp_to_edge", C.c_uint32), ("compression", C.c_uint32), ("level", C.c_uint32),
        ("layout", C.c_uint32), ("reserved", C.c_uint32), ("n_bricks", C.c_uint64),
        ("stored_bytes", C.c_uint64), ("raw_bytes", C.c_uint64),
    ]


class Request(C.Structure):
    """tvk_request: one render request of the server."""
    _fields_ = [("volume", C.c_uint32), ("render_mode", C.c_int32), ("view", View)]


class RequestResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("launch_jobs", C.c_uint32),
        ("evaluated_samples", C.c_uint64), ("nominal_samples", C.c_uint64),
        ("queue_ms", C.c_double), ("latency_ms", C.c_double),
        ("parts", C.c_uint32), ("reserved_", C.c_uint32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
        ("max_jobs_per_launch", C.c_uint64), ("max_lanes_busy", C.c_uint64),
        ("nominal_samples", C.c_uint64), ("paged_launches", C.c_uint64),
        ("pool_waits", C.c_uint64), ("max_pool_reserved", C.c_uint64),
        ("split_requests", C.c_uint64), ("parts_launched", C.c_uint64),
    ]


H = C.c_void_p
# tvk_brick_source_fn
BRICK_SOURCE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                           C.c_void_p, C.c_uint64)
u32, u64, i32, f32, f64 = C.c_uint32, C.c_uint64, C.c_int, C.c_float, C.c_double
P = C.c_void_p
F3 = C.POINTER(C.c_float)

# name -> argtypes (every function returns int unless listed in _RESTYPE)
SIGNATURES = {
    "tvk_last_error": [],
    "tvk_version": [],
    "tvk_create": [i32, i32, i32, i32, i32, C.POINTER(H)],
    "tvk_destroy": [H],
    "tvk_initialize": [H, i32],
    "tvk_resize": [H, u32, u32],
    "tvk_set_renderer_target": [H, i32],
    "tvk_get_renderer_target": [H, C.POINTER(i32)],
    "tvk_load_dataset": [H, C.c_char_p, i32],
    "tvk_set_volume": [H, P, u64, u64, u64, u32, u32, u32, u32],
    "tvk_set_paged_volume": [H, u64, u64, u64, u32, u32, u32, u32, C.c_void_p, P, P, u64],
    "tvk_set_paged_flat_volume": [H, P, u64, u64, u64, u32, u32, u32, u32, u64],
    "tvk_get_lod_level_count": [H, C.POINTER(u32)],
    "tvk_get_domain_size": [H, u32, C.POINTER(u64)],
    "tvk_get_bit_width": [H, C.POINTER(u32)],
    "tvk_get_brick_layout": [H, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)],
    "tvk_read_lod": [H, u32, P, u64],
    "tvk_read_brick": [H, u32, u32, u32, u32, P, u64],
    "tvk_get_brick_minmax": [H, u32, P, P, u64],
    "tvk_compute_histogram_1d": [H, P, u64],
    "tvk_compute_histogram_2d": [H, P, u32, u32, C.POINTER(f32)],
    "tvk_set_render_mode": [H, i32],
    "tvk_set_1d_trans": [H, P, u32],
    "tvk_set_2d_trans": [H, P, u32, u32, f32],
    "tvk_set_iso_value": [H, f32],
    "tvk_set_use_mip": [H, i32],
    "tvk_get_use_mip": [H, C.POINTER(i32)],
    "tvk_set_mip_rotation_angle": [H, f32],
    "tvk_set_isosurface_color": [H, f32, f32, f32],
    "tvk_set_clearview_enabled": [H, i32],
    "tvk_set_cv_iso_value": [H, f32],
    "tvk_set_cv_color": [H, f32, f32, f32],
    "tvk_set_cv_size": [H, f32],
    "tvk_set_cv_context_scale": [H, f32],
    "tvk_set_cv_border_scale": [H, f32],
    "tvk_set_cv_focus_pos": [H, i32, i32],
    "tvk_get_clearview": [H, C.POINTER(ClearView)],
    "tvk_set_view_pos": [H, f32, f32, f32],
    "tvk_set_view_dir": [H, f32, f32, f32],
    "tvk_set_up_dir": [H, f32, f32, f32],
    "tvk_set_fov": [H, f32],
    "tvk_set_ortho_view": [H, i32, f32],
    "tvk_set_rotation_4x4": [H, P],
    "tvk_set_translation_4x4": [H, P],
    "tvk_set_rescale_factors": [H, f64, f64, f64],
    "tvk_set_sample_rate_modifier": [H, f32],
    "tvk_set_interpolant": [H, i32],
    "tvk_set_lod_limits": [H, u32, u32],
    "tvk_set_bg_colors": [H, P, P],
    "tvk_set_lighting_enabled": [H, i32],
    "tvk_set_light_colors": [H, P, P, P, P],
    "tvk_set_early_ray_termination": [H, f32],
    "tvk_set_empty_space_skipping": [H, i32],
    "tvk_render_batch_noalloc": [H, P, u32, P, P],
    "tvk_set_clip_plane": [H, P],
    "tvk_enable_clip_plane": [H, i32],
    "tvk_is_clip_plane_enabled": [H, C.POINTER(i32)],
    "tvk_get_clip_plane": [H, P],
    "tvk_set_blend_precision": [H, i32],
    "tvk_get_blend_precision": [H, C.POINTER(i32)],
    "tvk_set_perf_measures": [H, u32, i32, f32, f32, u32],
    "tvk_read_framebuffer": [H, P, u64],
    "tvk_set_kernel_tuning": [H, u32, u32, u32],
    "tvk_set_work_group_waves": [H, u32],
    "tvk_set_brick_staging": [H, i32],
    "tvk_set_tf_grouping": [H, i32],
    "tvk_set_ray_tile_shape": [H, u32],
    "tvk_set_empty_space_cell": [H, u32],
    "tvk_set_1d_trans_slot": [H, u32, P, u32],
    "tvk_set_2d_trans_slot": [H, u32, P, u32, u32, f32],
    "tvk_render_batch": [H, C.POINTER(View), u32, C.POINTER(ViewStats), C.POINTER(f32)],
    "tvk_get_batch_framebuffer": [H, u32, C.POINTER(C.c_void_p)],
    "tvk_set_launch_graphs": [H, i32],
    "tvk_set_region_lods": [H, P, u64],
    "tvk_brick_histograms": [H, u32, u32, P, u64],
    "tvk_region_detail": [H, P, u64],
    "tvk_lod_mlp_create": [i32, P, P, C.POINTER(C.c_void_p)],
    "tvk_lod_mlp_predict": [C.c_void_p, P, u64, P, C.POINTER(f32)],
    "tvk_lod_mlp_destroy": [C.c_void_p],
    "tvk_brick_slot_bytes": [u32, u32, u32],
    "tvk_toc_build": [i32, P, u64, u64, u64, u32, u32, u32, i32, i32, u32, u32, u32, C.c_char_p,
                      C.POINTER(TocInfo)],
    "tvk_toc_open": [C.c_char_p, C.POINTER(C.c_void_p)],
    "tvk_toc_close": [C.c_void_p],
    "tvk_toc_get_info": [C.c_void_p, C.POINTER(TocInfo)],
    "tvk_toc_get_lod": [C.c_void_p, u32, P, P],
    "tvk_toc_read_brick": [C.c_void_p, u32, u32, u32, u32, P, u64, C.POINTER(u32), C.POINTER(f32),
                           C.POINTER(f32)],
    "tvk_set_toc_volume": [H, C.c_void_p, u64],
    "tvk_toc_get_histogram_1d": [C.c_void_p, P, C.POINTER(u64)],
    "tvk_toc_get_histogram_2d": [C.c_void_p, P, C.POINTER(u64), C.POINTER(u64), C.POINTER(f32)],
    "tvk_toc_build_raw": [i32, C.c_char_p, u64, u64, u64, u64, u32, u32, u32, i32, i32, u32, u32, u32,
                          C.c_char_p, C.POINTER(TocInfo)],
    "tvk_lod_features": [C.POINTER(View), P, u32, u32, P, P, P, P, P, P],
    "tvk_lod_select_octree": [P, P, P, u32, C.c_double, u64, u64, P, C.POINTER(C.c_double),
                              C.POINTER(u64)],
    "tvk_set_region_output": [H, i32],
    "tvk_set_region_check": [H, u32, C.c_float],
    "tvk_get_region_check": [H, P],
    "tvk_read_region_output": [H, P, u64],
    "tvk_autotune": [H, C.POINTER(View), u32, u32, C.POINTER(Tuning)],
    "tvk_server_create": [i32, u64, u32, C.POINTER(C.c_void_p)],
    "tvk_server_destroy": [C.c_void_p],
    "tvk_server_add_volume": [C.c_void_p, H, C.POINTER(u32)],
    "tvk_server_submit": [C.c_void_p, C.POINTER(Request), C.POINTER(u64)],
    "tvk_server_wait": [C.c_void_p, u64, u32, C.POINTER(RequestResult)],
    "tvk_server_get_stats": [C.c_void_p, C.POINTER(ServerStats)],
    "tvk_server_set_part_cost": [C.c_void_p, u64],
    "tvk_paint": [H],
    "tvk_check_for_redraw": [H],
    "tvk_schedule_complete_redraw": [H],
    "tvk_capture_single_frame": [H, P, u64, i32],
    "tvk_get_progress": [H, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32), C.POINTER(u32),
                         C.POINTER(u64)],
    "tvk_get_frame_stats": [H, C.POINTER(FrameStats)],
    "tvk_get_device_framebuffer": [H, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    "tvk_synchronize": [H],
    "tvk_set_max_usable_gpu_mem": [H, u64],
    "tvk_get_mem_stats": [H, C.POINTER(MemStats)],
    "tvk_set_max_usable_cpu_mem": [H, u64],
    "tvk_get_memman_stats": [i32, C.POINTER(MemManStats)],
    "tvk_set_brick_source_threads": [H, u32],
    "tvk_get_paging_stats": [H, C.POINTER(u64), C.POINTER(C.c_double)],
    "tvk_generate_volume": [i32, u64, u64, u64, u64, u32, P, u64],
    "tvk_generate_volume_gpu": [i32, i32, u64, u64, u64, u64, u32, P, u64],
    "tvk_set_generated_volume": [H, i32, u64, u64, u64, u64, u32, u32, u32, u32],
}
_RESTYPE = {"tvk_last_error": C.c_char_p, "tvk_version": C.c_char_p,
            "tvk_brick_slot_bytes": C.c_uint64}

_lib = None


def load() -> C.CDLL:
    """Load libtvkhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python imagevis3d_amd/build.py` "
            "(the HIP backend has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def last_error() -> str:
    return (load().tvk_last_error() or b"").decode("utf-8", "replace")


def call(name: str, *args) -> int:
    """Call `name` and raise TvkError on a negative status."""
    rc = getattr(load(), name)(*args)
    if rc < 0:
        raise TvkError(name, rc, last_error())
    return rc
