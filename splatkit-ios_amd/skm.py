"""Python binding of the render path's C ABI (include/skm.h) and of the scene library
(include/skm_scene.h), over ctypes. The HIP library is required: loading fails loudly when it is
missing, and there is no CPU fallback.

The Renderer class follows the engine-side interface of the reference's SplatRenderer
(packages/splatkit-engine/include/splatkit/rendering/SplatRenderer.h:36-160): upload_world,
create_slab/upload_tile, draw(frame), finish, the stat getters and read_frame (captureNextFrame).
"""
import ctypes as C
import math
import time
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SKM_LIB names another build of the library (e.g. one with a different radix tile size).
LIB_PATH = os.environ.get("SKM_LIB") or os.path.join(HERE, "libskm.so")
SCENE_PATH = os.path.join(HERE, "libskm_scene.so")

SKM_OK, SKM_ERR_INVALID, SKM_ERR_DEVICE, SKM_ERR_MEMORY = 0, -1, -2, -3
SKM_ERR_STATE, SKM_ERR_UNSUPPORTED, SKM_ERR_CAPACITY = -4, -5, -6
PIXELS_BGRA8, PIXELS_RGBA32F = 0, 1
TILE = 16

_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


class Config(C.Structure):
    _fields_ = [("device", C.c_int32), ("tile_size", C.c_uint32), ("pair_factor", C.c_uint32),
                ("sort_kernel", C.c_uint32), ("raster_px", C.c_uint32), ("frames_in_flight", C.c_uint32)]


class Range(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("count", C.c_uint32)]


class Reveal(C.Structure):
    _fields_ = [("level", C.c_float), ("band", C.c_float), ("tint_amount", C.c_float),
                ("brightness_boost", C.c_float), ("tint", C.c_float * 4)]


class Frame(C.Structure):
    _fields_ = [("order_source", C.c_int32), ("order", _u32p), ("order_count", C.c_uint32),
                ("draw_count", C.c_uint32), ("ranges", C.POINTER(Range)), ("range_count", C.c_uint32),
                ("sh_degree", C.c_int32), ("view", C.c_float * 16), ("proj", C.c_float * 16),
                ("camera_position", C.c_float * 3), ("label_styles", C.c_void_p), ("reveal", Reveal)]


class WorldInfo(C.Structure):
    _fields_ = [("count", C.c_uint32), ("sh_degree", C.c_int32)]


class CloudView(C.Structure):
    _fields_ = [("count", C.c_uint32), ("sh_degree", C.c_int32), ("positions", _fp),
                ("covariances", _fp), ("colors", _fp), ("alphas", _fp), ("sh", _fp)]


class RenderPolicy(C.Structure):
    _fields_ = [("lod_error_pixels", C.c_float), ("lod_splat_limit", C.c_uint32),
                ("alpha_threshold", C.c_float), ("subpixel_threshold", C.c_float),
                ("enable_frustum_culling", C.c_uint8), ("enable_hiz_occlusion", C.c_uint8),
                ("enable_early_termination", C.c_uint8), ("reserved0", C.c_uint8),
                ("sort_depth", C.c_uint32), ("background_color", C.c_double * 3)]


class PolicySupport(C.Structure):
    _fields_ = [("lod_error_pixels", C.c_uint8), ("lod_splat_limit", C.c_uint8),
                ("alpha_threshold", C.c_uint8), ("subpixel_threshold", C.c_uint8),
                ("enable_frustum_culling", C.c_uint8), ("enable_hiz_occlusion", C.c_uint8),
                ("enable_early_termination", C.c_uint8), ("sort_depth", C.c_uint8),
                ("background_color", C.c_uint8), ("reserved", C.c_uint8 * 3),
                ("min_lod_error_pixels", C.c_float), ("max_lod_error_pixels", C.c_float),
                ("min_subpixel_threshold", C.c_float), ("max_subpixel_threshold", C.c_float),
                ("fallback", RenderPolicy)]


class DeviceCapabilities(C.Structure):
    _fields_ = [("max_lod_capacity_splats", C.c_uint32), ("min_residency_capacity_splats", C.c_uint32),
                ("max_residency_capacity_splats", C.c_uint32), ("supports_hiz_occlusion", C.c_uint8),
                ("supports_subgroups", C.c_uint8), ("reserved", C.c_uint8 * 2),
                ("max_texture_dimension", C.c_uint32), ("policy", PolicySupport)]


class LodThreshold(C.Structure):
    _fields_ = [("policy_pixels", C.c_float), ("pixels", C.c_float)]


class FrameTotals(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("stage_ms", C.c_double * 6), ("visible", C.c_uint64),
                ("pairs", C.c_uint64), ("dropped", C.c_uint64), ("raster_tests", C.c_uint64)]


class View(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("camera_position", C.c_float * 3),
                ("width", C.c_uint32), ("height", C.c_uint32), ("reserved", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_size", C.c_uint32), ("tile_radix_bits", C.c_uint32), ("use_graph", C.c_uint32),
                ("sort_kernel", C.c_uint32), ("raster_px", C.c_uint32), ("reserved", C.c_uint32)]


class TuningResult(C.Structure):
    _fields_ = [("tuning", BatchTuning), ("gpu_ms", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("views", C.c_uint32), ("graph", C.c_uint32), ("gpu_ms", C.c_double), ("visible", C.c_uint64),
                ("pairs", C.c_uint64), ("tile_passes", C.c_uint32), ("view_sort_passes", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_batch_views", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("sort_depth", C.c_uint32), ("max_in_flight", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("reserved", C.c_uint32), ("user_data", C.c_uint64), ("view", View)]


class Completion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("user_data", C.c_uint64), ("status", C.c_int32),
                ("batch_views", C.c_uint32), ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("launches", C.c_uint64), ("resident_bytes", C.c_uint64), ("scratch_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("mean_batch_views", C.c_double), ("frames_per_s", C.c_double)]


class SceneSpec(C.Structure):
    _fields_ = [("count", C.c_uint32), ("sh_degree", C.c_int32), ("seed", C.c_uint64),
                ("layout", C.c_int32), ("spread", C.c_float), ("area", C.c_float),
                ("surface_share", C.c_float), ("surface_sigma", C.c_float),
                ("log_scale_mean", C.c_float), ("log_scale_sigma", C.c_float),
                ("opacity_sigma", C.c_float), ("dc_sigma", C.c_float), ("sh_sigma", C.c_float)]


class LodNode(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("size", C.c_float), ("child_start", C.c_uint32),
                ("child_count", C.c_uint32)]


class LodCluster(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("extent", C.c_float * 3), ("error", C.c_float),
                ("color_variance", C.c_float), ("opacity", C.c_float), ("node", C.c_uint32),
                ("child_start", C.c_uint32), ("child_count", C.c_uint32), ("leaf_start", C.c_uint32),
                ("leaf_count", C.c_uint32), ("subtree_leaves", C.c_uint32)]


class LodFileHeader(C.Structure):
    _fields_ = [("version", C.c_uint32), ("node_count", C.c_uint32), ("leaf_count", C.c_uint32),
                ("sh_degree", C.c_int32), ("depth", C.c_uint32), ("cluster_count", C.c_uint32),
                ("leaf_ref_count", C.c_uint32), ("reserved", C.c_uint32), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3)]


class LodTreeView(C.Structure):
    _fields_ = [("nodes", CloudView), ("layout", C.POINTER(LodNode)), ("leaf_count", C.c_uint64),
                ("clusters", C.POINTER(LodCluster)), ("cluster_count", C.c_uint64),
                ("leaves", _u32p), ("leaf_ref_count", C.c_uint64)]


# name -> (restype, argtypes); every symbol include/skm.h declares.
_R = C.c_void_p
API = {
    "skm_renderer_create": (C.c_int, [C.POINTER(Config), C.POINTER(_R)]),
    "skm_renderer_destroy": (None, [_R]),
    "skm_last_error": (C.c_char_p, [_R]),
    "skm_set_render_scale": (None, [_R, C.c_float]),
    "skm_render_scale": (C.c_float, [_R]),
    "skm_set_linear_blending": (None, [_R, C.c_int]),
    "skm_count_raster_tests": (None, [_R, C.c_int]),
    "skm_set_lod_head": (None, [_R, C.c_uint32]),
    "skm_linear_blending": (C.c_int, [_R]),
    "skm_set_vsync": (None, [_R, C.c_int]),
    "skm_set_draw_extent": (C.c_int, [_R, C.c_uint32, C.c_uint32]),
    "skm_get_device_capabilities": (C.c_int, [_R, C.POINTER(DeviceCapabilities)]),
    "skm_apply_render_policy": (C.c_int, [_R, C.POINTER(RenderPolicy), C.c_char_p, C.c_size_t]),
    "skm_available_memory": (C.c_int, [_R, C.POINTER(C.c_uint64)]),
    "skm_ready": (C.c_int, [_R]),
    "skm_gpu_failed": (C.c_int, [_R]),
    "skm_has_completed_world_frame": (C.c_int, [_R]),
    "skm_draw_extent": (None, [_R, _u32p, _u32p]),
    "skm_generation": (C.c_uint32, [_R]),
    "skm_surface_rebuild_pending": (C.c_int, [_R]),
    "skm_upload_world": (C.c_int, [_R, C.POINTER(CloudView), C.c_int]),
    "skm_selects_lod_on_gpu": (C.c_int, [_R]),
    "skm_upload_lod_world": (C.c_int, [_R, C.POINTER(LodTreeView), C.c_int, C.c_uint32]),
    "skm_create_slab": (C.c_int, [_R, C.c_uint32, C.c_int]),
    "skm_upload_tile": (C.c_int, [_R, C.c_uint32, C.POINTER(CloudView)]),
    "skm_world": (C.c_int, [_R, C.POINTER(WorldInfo)]),
    "skm_sorts_on_gpu": (C.c_int, [_R]),
    "skm_draw": (C.c_int, [_R, C.POINTER(Frame)]),
    "skm_finish": (C.c_int, [_R]),
    "skm_submitted_frames": (C.c_uint64, [_R]),
    "skm_completed_frames": (C.c_uint64, [_R]),
    "skm_last_gpu_millis": (C.c_double, [_R]),
    "skm_last_sort_millis": (C.c_double, [_R]),
    "skm_last_cull_millis": (C.c_double, [_R]),
    "skm_last_draw_count": (C.c_uint32, [_R]),
    "skm_last_selected_count": (C.c_uint32, [_R]),
    "skm_last_select_millis": (C.c_double, [_R]),
    "skm_device_description": (C.c_char_p, [_R]),
    "skm_read_frame": (C.c_int, [_R, C.c_int, C.c_void_p, C.c_size_t]),
    "skm_read_visible": (C.c_int, [_R, _u32p, _u32p, _u32p, C.c_uint32]),
    "skm_read_order": (C.c_int, [_R, _u32p, C.c_uint32]),
    "skm_read_tiles": (C.c_int, [_R, C.POINTER(C.c_uint64), _u32p, _u32p, C.c_uint64, _u32p, C.c_uint32]),
    "skm_read_stage_millis": (C.c_int, [_R, C.POINTER(C.c_double)]),
    "skm_read_frame_totals": (C.c_int, [_R, C.POINTER(FrameTotals), C.c_int]),
    "skm_read_lod_cut": (C.c_int, [_R, _u32p, _u32p, C.c_uint32, _u32p]),
    "skm_lod_threshold_reset": (None, [C.POINTER(LodThreshold), C.c_float]),
    "skm_lod_threshold_update": (C.c_float, [C.POINTER(LodThreshold), C.c_uint32, C.c_uint32, C.c_uint32]),
    "skm_lod_threshold_pixels": (C.c_float, [_R]),
    "skm_lod_select": (C.c_int, [_R, C.POINTER(LodTreeView), _fp, _fp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                 C.c_float, C.c_int, _u32p, _u32p]),
    "skm_draw_views": (C.c_int, [_R, C.POINTER(View), C.c_uint32, C.POINTER(Range), C.c_uint32,
                                 C.POINTER(BatchTuning)]),
    "skm_read_batch_stats": (C.c_int, [_R, C.POINTER(BatchStats)]),
    "skm_set_tuning": (C.c_int, [_R, C.POINTER(BatchTuning)]),
    "skm_autotune_views": (C.c_int, [_R, C.POINTER(View), C.c_uint32, C.POINTER(Range), C.c_uint32, C.c_uint32,
                                     C.POINTER(TuningResult), C.c_uint32, _u32p, C.POINTER(BatchTuning)]),
    "skm_read_view_frame": (C.c_int, [_R, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t]),
    "skm_read_view_visible": (C.c_int, [_R, C.c_uint32, _u32p, _u32p, _u32p, C.c_uint32]),
    "skm_read_view_order": (C.c_int, [_R, C.c_uint32, _u32p, C.c_uint32]),
    "skm_read_view_tiles": (C.c_int, [_R, C.c_uint32, C.POINTER(C.c_uint64), _u32p, _u32p, C.c_uint64, _u32p,
                                      C.c_uint32]),
    "skm_argsort_u32": (C.c_int, [_R, _u32p, C.c_uint32, C.c_int, _u32p]),
    "skm_morton_code": (C.c_uint32, [_fp, _fp, _fp]),
    "skm_reorder_spatially": (C.c_int, [_R, _fp, _fp, _fp, _fp, _fp, C.c_uint32, C.c_uint32, _fp, _fp, _u32p]),
    "skm_argsort_u32_device": (C.c_int, [_R, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                         C.POINTER(C.c_double)]),
    "skm_argsort_u32_device_static": (C.c_int, [_R, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
}

_S = C.c_void_p
# include/skm_server.h
SERVER_API = {
    "skm_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_S)]),
    "skm_server_destroy": (None, [_S]),
    "skm_server_last_error": (C.c_char_p, [_S]),
    "skm_server_add_scene": (C.c_int, [_S, C.POINTER(CloudView), C.c_int, _u32p]),
    "skm_server_submit": (C.c_int, [_S, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "skm_server_poll": (C.c_int, [_S, C.POINTER(Completion), C.c_uint32, _u32p, C.c_double]),
    "skm_server_image": (C.c_int, [_S, C.c_uint64, C.POINTER(C.POINTER(C.c_uint8)), _u32p, _u32p]),
    "skm_server_release": (C.c_int, [_S, C.c_uint64]),
    "skm_server_stats_read": (C.c_int, [_S, C.POINTER(ServerStats), C.c_int]),
}

# include/skm_lod.h
LOD_API = {
    "skm_lod_build": (C.c_void_p, [C.POINTER(CloudView), C.c_uint32, C.c_uint32, C.c_uint32]),
    "skm_lod_from_layout": (C.c_void_p, [C.POINTER(CloudView), C.POINTER(LodNode), C.c_uint64]),
    "skm_lod_free": (None, [C.c_void_p]),
    "skm_lod_view": (None, [C.c_void_p, C.POINTER(LodTreeView)]),
    "skm_lod_rounds": (C.c_uint32, [C.POINTER(LodTreeView)]),
    "skm_lod_bounds": (None, [C.c_void_p, _fp]),
    "skm_lod_read_header": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(LodFileHeader), C.c_char_p, C.c_size_t]),
    "skm_lod_decode": (C.c_void_p, [C.c_void_p, C.c_size_t, C.c_int, C.c_char_p, C.c_size_t]),
    "skm_lod_read": (C.c_void_p, [C.c_char_p, C.c_int, C.c_char_p, C.c_size_t]),
    "skm_lod_validate": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_char_p, C.c_size_t]),
    "skm_lod_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]),
}

_lib = None
_scene = None


def lib():
    """The HIP library; raises when it is not built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in {**API, **LOD_API, **SERVER_API}.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def scene_lib():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_PATH):
            raise RuntimeError(f"{SCENE_PATH} is missing: run __graft_entry__.build() first")
        h = C.CDLL(SCENE_PATH)
        h.skm_scene_fill.restype = C.c_int
        h.skm_scene_fill.argtypes = [C.POINTER(SceneSpec), _fp, _fp, _fp, _fp, _fp, C.c_uint32]
        h.skm_look_at.argtypes = [_fp, _fp, _fp, _fp]
        h.skm_perspective.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, _fp]
        _scene = h
    return _scene


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(_u32p) if a is not None else None


# ---------------------------------------------------------------------------------------------
# Scenes and cameras.

def read_lod_header(data):
    """readLodSplatHeader: (header dict, None) or (None, reason); `data` is the file's first bytes and
    its length the file's (pass a bytes-like of the file, or (prefix, length))."""
    if isinstance(data, tuple):
        prefix, size = data
    else:
        prefix, size = data, len(data)
    buf = np.frombuffer(bytes(prefix), np.uint8)
    h = LodFileHeader()
    err = C.create_string_buffer(256)
    if lib().skm_lod_read_header(buf.ctypes.data_as(C.c_void_p), size, C.byref(h), err, 256) != 0:
        return None, err.value.decode()
    return {k: (list(getattr(h, k)) if k.startswith("bounds") else getattr(h, k)) for k, _ in LodFileHeader._fields_}, None


def sh_floats(degree):
    return 3 * ((degree + 1) ** 2 - 1)


def make_cloud(count, sh_degree=0, seed=0x5EED0000, layout=0, spread=2.0, area=10.0, surface_share=0.7,
               surface_sigma=0.05, log_scale_mean=-4.5, log_scale_sigma=0.5, opacity_sigma=1.5,
               dc_sigma=0.5, sh_sigma=0.05, threads=8):
    """A seeded synthetic SplatCloud as a dict of float32 arrays (see include/skm_scene.h)."""
    spec = SceneSpec(count, sh_degree, seed, layout, spread, area, surface_share, surface_sigma,
                     log_scale_mean, log_scale_sigma, opacity_sigma, dc_sigma, sh_sigma)
    cloud = {
        "positions": np.empty((count, 3), np.float32),
        "covariances": np.empty((count, 6), np.float32),
        "colors": np.empty((count, 3), np.float32),
        "alphas": np.empty(count, np.float32),
        "sh": np.empty((count, sh_floats(sh_degree)), np.float32) if sh_degree > 0 else None,
        "sh_degree": sh_degree,
    }
    rc = scene_lib().skm_scene_fill(C.byref(spec), fptr(cloud["positions"]), fptr(cloud["covariances"]),
                                    fptr(cloud["colors"]), fptr(cloud["alphas"]), fptr(cloud["sh"]), threads)
    if rc != 0:
        raise ValueError("invalid scene spec")
    return cloud


def empty_cloud(count, sh_degree=0):
    return {
        "positions": np.zeros((count, 3), np.float32), "covariances": np.zeros((count, 6), np.float32),
        "colors": np.zeros((count, 3), np.float32), "alphas": np.zeros(count, np.float32),
        "sh": np.zeros((count, sh_floats(sh_degree)), np.float32) if sh_degree > 0 else None,
        "sh_degree": sh_degree,
    }


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    out = np.zeros(16, np.float32)
    e, t, u = (np.asarray(v, np.float32) for v in (eye, target, up))
    scene_lib().skm_look_at(fptr(e), fptr(t), fptr(u), fptr(out))
    return out


def perspective(fovy, aspect, near, far):
    out = np.zeros(16, np.float32)
    scene_lib().skm_perspective(fovy, aspect, near, far, fptr(out))
    return out


def orbit_camera(frame, width, height, radius=8.0, height_m=1.5, step_deg=6.0, fovy=math.pi / 3,
                 near=0.2, far=100.0):
    a = math.radians(step_deg * frame)
    eye = (radius * math.sin(a), height_m, radius * math.cos(a))
    return look_at(eye), perspective(fovy, width / height, near, far), np.asarray(eye, np.float32)


# The workloads (seeded; scene k uses seed 0x5EED0000 + k).
def scene_c0(count=100_000):
    return make_cloud(count, 0, seed=0x5EED0000, layout=0, spread=2.0)


def scene_c1(count=3_000_000, sh_degree=3):
    return make_cloud(count, sh_degree, seed=0x5EED0001, layout=1, spread=5.0, area=10.0)


RESOLUTIONS_C2 = ((1280, 720), (1920, 1080), (2560, 1440))


def views_c2(count=64, seed=0x5EED0002, radius=(4.0, 12.0), resolutions=RESOLUTIONS_C2, fovy=math.pi / 3,
             near=0.2, far=100.0):
    """Batch views of the 3M-splat scene: eyes uniform on spheres of radius U(4, 12) m looking at the
    origin, each at a resolution drawn uniformly from 720p / 1080p / 1440p. Returns a list of
    (view, proj, eye, width, height)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = (d * rng.uniform(*radius)).astype(np.float32)
        if abs(d[1]) > 0.999:
            eye[0] += 1e-3
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        out.append((look_at(eye), perspective(fovy, w / h, near, far), eye, w, h))
    return out


def camera_c0(width=1280, height=720):
    return look_at((0.0, 0.0, 6.0)), perspective(math.pi / 3, width / height, 0.2, 100.0), \
        np.array([0.0, 0.0, 6.0], np.float32)


# ---------------------------------------------------------------------------------------------

class SkmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"skm status {status}: {message}")
        self.status = status


class DeviceBuffer:
    """Device memory from the HIP runtime libskm.so itself links (not a second runtime such as the one
    a framework may bundle), for the device-pointer entry points in tests and benchmarks."""

    def __init__(self, nbytes):
        lib()  # loads libskm.so, and with it its HIP runtime
        self._hip = C.CDLL("libamdhip64.so.7")
        self._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self._hip.hipFree.argtypes = [C.c_void_p]
        self._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        if self._hip.hipMalloc(C.byref(self.ptr), max(1, nbytes)) != 0:
            raise MemoryError(f"hipMalloc of {nbytes} bytes failed")

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        if self._hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to device failed")

    def copy_from(self, other, nbytes):
        if self._hip.hipMemcpy(self.ptr, other.ptr, nbytes, 3) != 0:  # device to device
            raise RuntimeError("hipMemcpy on device failed")

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        if self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2) != 0:  # device to host
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def close(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    def __init__(self, device=0, pair_factor=0, sort_kernel=0, tile=TILE, raster_px=0, frames_in_flight=0):
        self._lib = lib()
        self._h = _R()
        self.tile = tile
        cfg = Config(device, tile, pair_factor, sort_kernel, raster_px, frames_in_flight)
        rc = self._lib.skm_renderer_create(C.byref(cfg), C.byref(self._h))
        if rc != SKM_OK:
            raise SkmError(rc, "renderer creation failed (no HIP device?)")
        self.world_count = 0

    def close(self):
        if self._h:
            self._lib.skm_renderer_destroy(self._h)
            self._h = _R()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != SKM_OK:
            raise SkmError(rc, self.error())
        return rc

    def error(self):
        return (self._lib.skm_last_error(self._h) or b"").decode()

    # -- configuration
    def set_extent(self, width, height):
        self._check(self._lib.skm_set_draw_extent(self._h, width, height))

    def extent(self):
        w, h = C.c_uint32(), C.c_uint32()
        self._lib.skm_draw_extent(self._h, C.byref(w), C.byref(h))
        return w.value, h.value

    def apply_policy(self, sort_depth=32, subpixel=1.0, background=(0.05, 0.05, 0.08), lod_error=1.0,
                     lod_limit=0):
        p = RenderPolicy(lod_error, lod_limit, 1.0 / 255.0, subpixel, 1, 0, 1, 0, sort_depth,
                         (C.c_double * 3)(*background))
        reason = C.create_string_buffer(256)
        self._check(self._lib.skm_apply_render_policy(self._h, C.byref(p), reason, 256))

    def capabilities(self):
        caps = DeviceCapabilities()
        self._check(self._lib.skm_get_device_capabilities(self._h, C.byref(caps)))
        return caps

    def description(self):
        return self._lib.skm_device_description(self._h).decode()

    # -- worlds
    @staticmethod
    def _view(cloud):
        n = len(cloud["alphas"])
        keep = {k: np.ascontiguousarray(cloud[k], np.float32) for k in ("positions", "covariances", "colors", "alphas")}
        sh = cloud.get("sh")
        keep["sh"] = np.ascontiguousarray(sh, np.float32) if sh is not None else None
        view = CloudView(n, cloud.get("sh_degree", 0) if sh is not None else 0, fptr(keep["positions"]),
                         fptr(keep["covariances"]), fptr(keep["colors"]), fptr(keep["alphas"]), fptr(keep["sh"]))
        return view, keep

    def upload_world(self, cloud, max_sh_degree=3):
        view, keep = self._view(cloud)
        self._check(self._lib.skm_upload_world(self._h, C.byref(view), max_sh_degree))
        self.world_count = view.count
        del keep

    def create_slab(self, capacity, sh_degree):
        self._check(self._lib.skm_create_slab(self._h, capacity, sh_degree))
        self.world_count = capacity

    def upload_tile(self, offset, cloud):
        view, keep = self._view(cloud)
        self._check(self._lib.skm_upload_tile(self._h, offset, C.byref(view)))
        del keep

    def upload_lod_world(self, tree, max_sh_degree=3, budget=0):
        self._check(self._lib.skm_upload_lod_world(self._h, C.byref(tree.view), max_sh_degree, budget))
        self.world_count = tree.node_count
        self._lod_tree = tree  # the view borrows the tree's arrays only during the call; keep it alive anyway

    def read_lod_cut(self):
        count = C.c_uint32()
        stats = np.zeros(6, np.uint32)
        self._check(self._lib.skm_read_lod_cut(self._h, C.byref(count), None, 0, uptr(stats)))
        idx = np.zeros(max(1, count.value), np.uint32)
        self._check(self._lib.skm_read_lod_cut(self._h, C.byref(count), uptr(idx), count.value, uptr(stats)))
        return idx[:count.value].copy(), stats

    def lod_select(self, tree, view, proj, width, height, capacity, pixel_limit=1.0, color_weight=4.0, cull=True):
        """One cut on its own (MetalLodSelection): node indices and stats6."""
        cap = max(1, min(capacity or tree.leaf_count, tree.leaf_count))
        idx = np.zeros(cap, np.uint32)
        stats = np.zeros(6, np.uint32)
        v = np.ascontiguousarray(view, np.float32)
        p = np.ascontiguousarray(proj, np.float32)
        self._check(self._lib.skm_lod_select(self._h, C.byref(tree.view), fptr(v), fptr(p), width, height, cap,
                                             pixel_limit, color_weight, 1 if cull else 0, uptr(idx), uptr(stats)))
        return idx[:stats[0]].copy(), stats

    def world(self):
        info = WorldInfo()
        return (info.count, info.sh_degree) if self._lib.skm_world(self._h, C.byref(info)) else None

    # -- frames
    def make_frame(self, view, proj, cam_pos, ranges=None, sh_degree=3, reveal=None):
        if ranges is None:
            ranges = [(0, self.world_count)]
        rarr = (Range * max(1, len(ranges)))(*[Range(o, c) for o, c in ranges])
        f = Frame()
        f.order_source = 1
        f.ranges = C.cast(rarr, C.POINTER(Range)) if ranges else None
        f.range_count = len(ranges)
        f.sh_degree = sh_degree
        f.view[:] = [float(x) for x in np.asarray(view, np.float32)]
        f.proj[:] = [float(x) for x in np.asarray(proj, np.float32)]
        f.camera_position[:] = [float(x) for x in np.asarray(cam_pos, np.float32)[:3]]
        if reveal is None:
            f.reveal = Reveal(math.inf, 0.0, 0.0, 0.0, (C.c_float * 4)(1, 1, 1, 0))
        else:
            f.reveal = Reveal(*reveal[:4], (C.c_float * 4)(*reveal[4:8]))
        f._ranges_keep = rarr
        return f

    def draw(self, frame):
        return self._check(self._lib.skm_draw(self._h, C.byref(frame)))

    def render(self, view, proj, cam_pos, **kw):
        """Draws one frame and waits for it. A frame whose tile pairs outgrew the capacity is
        dropped and the capacity grows; it is then drawn once more."""
        frame = self.make_frame(view, proj, cam_pos, **kw)
        self.draw(frame)
        try:
            self.finish()
        except SkmError as e:
            if e.status != SKM_ERR_CAPACITY:
                raise
            self.draw(frame)
            self.finish()

    def finish(self):
        return self._check(self._lib.skm_finish(self._h))

    # -- view batches
    @staticmethod
    def make_views(views):
        arr = (View * len(views))()
        for k, (view, proj, eye, w, h) in enumerate(views):
            arr[k].view[:] = [float(x) for x in np.asarray(view, np.float32)]
            arr[k].proj[:] = [float(x) for x in np.asarray(proj, np.float32)]
            arr[k].camera_position[:] = [float(x) for x in np.asarray(eye, np.float32)[:3]]
            arr[k].width, arr[k].height = int(w), int(h)
        return arr

    def draw_views(self, views, ranges=None, tile_size=16, radix_bits=8, graph=True, sort_kernel=0, raster_px=0):
        """Enqueues one batch; `views` is a list of (view, proj, eye, width, height) or a View array."""
        arr = views if isinstance(views, C.Array) else self.make_views(views)
        rarr = None
        if ranges is not None:
            rarr = (Range * max(1, len(ranges)))(*[Range(o, c) for o, c in ranges])
        t = BatchTuning(tile_size, radix_bits, 1 if graph else 0, sort_kernel, raster_px, 0)
        self._views = arr
        self._view_tile = tile_size
        return self._check(self._lib.skm_draw_views(self._h, arr, len(arr), rarr, 0 if ranges is None else len(ranges),
                                                    C.byref(t)))

    def render_views(self, views, **kw):
        """Draws one batch and waits; a batch whose tile pairs outgrew the capacity is drawn again."""
        arr = views if isinstance(views, C.Array) else self.make_views(views)
        self.draw_views(arr, **kw)
        try:
            self.finish()
        except SkmError as e:
            if e.status != SKM_ERR_CAPACITY:
                raise
            self.draw_views(arr, **kw)
            self.finish()

    def autotune_views(self, views, ranges=None, repeats=2):
        """Times every batch configuration on these views and keeps the fastest; returns
        (best, [(tile, raster_px, sort_kernel, radix_bits, gpu_ms), ...])."""
        arr = views if isinstance(views, C.Array) else self.make_views(views)
        rarr = None
        if ranges is not None:
            rarr = (Range * max(1, len(ranges)))(*[Range(o, c) for o, c in ranges])
        res = (TuningResult * 64)()
        n = C.c_uint32()
        best = BatchTuning()
        self._check(self._lib.skm_autotune_views(self._h, arr, len(arr), rarr, 0 if ranges is None else len(ranges),
                                                 repeats, res, 64, C.byref(n), C.byref(best)))
        table = [(r.tuning.tile_size, r.tuning.raster_px, r.tuning.sort_kernel, r.tuning.tile_radix_bits, r.gpu_ms)
                 for r in res[:n.value]]
        self._views = arr
        best = {"tile_size": best.tile_size, "raster_px": best.raster_px, "sort_kernel": best.sort_kernel,
                "radix_bits": best.tile_radix_bits}
        self._view_tile = best["tile_size"]
        return best, table

    def set_tuning(self, tile_size=16, raster_px=0, sort_kernel=0, radix_bits=8):
        """Tile size, raster pixels per lane, radix pass variant and tile-sort digit width for single
        draws and batches."""
        t = BatchTuning(tile_size, radix_bits, 1, sort_kernel, raster_px, 0)
        self._check(self._lib.skm_set_tuning(self._h, C.byref(t)))
        self.tile = tile_size

    def batch_stats(self):
        s = BatchStats()
        self._check(self._lib.skm_read_batch_stats(self._h, C.byref(s)))
        return {"views": s.views, "graph": bool(s.graph), "gpu_ms": s.gpu_ms, "visible": s.visible, "pairs": s.pairs,
                "tile_passes": s.tile_passes, "view_sort_passes": s.view_sort_passes}

    def read_view_frame(self, v, fmt=PIXELS_RGBA32F):
        w, h = self._views[v].width, self._views[v].height
        out = np.empty((h, w, 4), np.float32 if fmt == PIXELS_RGBA32F else np.uint8)
        self._check(self._lib.skm_read_view_frame(self._h, v, fmt, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def read_view_visible(self, v, capacity=None):
        capacity = capacity or max(1, self.world_count)
        count = C.c_uint32()
        keys = np.empty(capacity, np.uint32)
        proj = np.empty((capacity, 8), np.uint32)
        self._check(self._lib.skm_read_view_visible(self._h, v, C.byref(count), uptr(keys), uptr(proj), capacity))
        n = min(count.value, capacity)
        return count.value, keys[:n].copy(), proj[:n].copy()

    def read_view_order(self, v, n):
        order = np.empty(max(1, n), np.uint32)
        self._check(self._lib.skm_read_view_order(self._h, v, uptr(order), n))
        return order[:n].copy()

    def read_view_tiles(self, v):
        w, h = self._views[v].width, self._views[v].height
        t = self._view_tile
        tiles = ((w + t - 1) // t) * ((h + t - 1) // t)
        pairs = C.c_uint64()
        self._check(self._lib.skm_read_view_tiles(self._h, v, C.byref(pairs), None, None, 0, None, 0))
        cap = int(pairs.value)
        pt = np.empty(max(1, cap), np.uint32)
        pv = np.empty(max(1, cap), np.uint32)
        ranges = np.empty((max(1, tiles), 2), np.uint32)
        self._check(self._lib.skm_read_view_tiles(self._h, v, C.byref(pairs), uptr(pt), uptr(pv), cap, uptr(ranges),
                                                  tiles))
        return pairs.value, pt[:cap].copy(), pv[:cap].copy(), ranges[:tiles].copy()

    def set_linear_blending(self, linear):
        self._lib.skm_set_linear_blending(self._h, 1 if linear else 0)

    def count_raster_tests(self, enable=True):
        """Counts the raster's coverage tests into frame_totals()['raster_tests'] (off by default)."""
        self._lib.skm_count_raster_tests(self._h, 1 if enable else 0)

    def set_lod_head(self, max_frontier):
        """Leading cut rounds with frontiers of at most max_frontier clusters in one block (0: none)."""
        self._lib.skm_set_lod_head(self._h, int(max_frontier))

    def lod_threshold(self):
        """The screen-space error the next tree-world frame's cut runs at (LodThreshold::pixels)."""
        return float(self._lib.skm_lod_threshold_pixels(self._h))

    def stage_millis(self):
        """GPU ms of the last frame: LOD cut, cull, depth sort, bin + tile sort + ranges, raster, total."""
        ms = (C.c_double * 6)()
        self._check(self._lib.skm_read_stage_millis(self._h, ms))
        return list(ms)

    def frame_totals(self, reset=True):
        t = FrameTotals()
        self._check(self._lib.skm_read_frame_totals(self._h, C.byref(t), 1 if reset else 0))
        return {"frames": t.frames, "stage_ms": list(t.stage_ms), "visible": t.visible, "pairs": t.pairs,
                "dropped": t.dropped, "raster_tests": t.raster_tests}

    def read_frame(self, fmt=PIXELS_RGBA32F):
        w, h = self.extent()
        if fmt == PIXELS_RGBA32F:
            out = np.empty((h, w, 4), np.float32)
        else:
            out = np.empty((h, w, 4), np.uint8)
        self._check(self._lib.skm_read_frame(self._h, fmt, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def read_visible(self, capacity=None):
        if capacity is None:
            capacity = max(1, self.world_count)
        count = C.c_uint32()
        keys = np.empty(capacity, np.uint32)
        proj = np.empty((capacity, 8), np.uint32)
        self._check(self._lib.skm_read_visible(self._h, C.byref(count), uptr(keys), uptr(proj), capacity))
        n = min(count.value, capacity)
        return count.value, keys[:n].copy(), proj[:n].copy()

    def read_order(self, n):
        order = np.empty(max(1, n), np.uint32)
        self._check(self._lib.skm_read_order(self._h, uptr(order), n))
        return order[:n].copy()

    def read_tiles(self, pair_capacity=None, tiles=None):
        """Tile pairs and ranges of the last frame; capacities default to what the frame holds."""
        pairs = C.c_uint64()
        if tiles is None:
            w, h = self.extent()
            tiles = ((w + self.tile - 1) // self.tile) * ((h + self.tile - 1) // self.tile)
        if pair_capacity is None:
            self._check(self._lib.skm_read_tiles(self._h, C.byref(pairs), None, None, 0, None, 0))
            pair_capacity = int(pairs.value)
        pt = np.empty(max(1, pair_capacity), np.uint32)
        pv = np.empty(max(1, pair_capacity), np.uint32)
        ranges = np.empty((max(1, tiles), 2), np.uint32)
        self._check(self._lib.skm_read_tiles(self._h, C.byref(pairs), uptr(pt), uptr(pv), pair_capacity,
                                             uptr(ranges), tiles))
        n = min(pairs.value, pair_capacity)
        return pairs.value, pt[:n].copy(), pv[:n].copy(), ranges[:tiles].copy()

    def argsort(self, keys, key_bits=32):
        keys = np.ascontiguousarray(keys, np.uint32)
        order = np.empty(max(1, len(keys)), np.uint32)
        self._check(self._lib.skm_argsort_u32(self._h, uptr(keys), len(keys), key_bits, uptr(order)))
        return order[:len(keys)]

    def reorder_spatially(self, cloud, lo=None, hi=None):
        """The loader's spatial order (reorderSpatially, SpatialOrder.cpp:48-92) on the GPU: permutes the
        cloud's arrays in place along the Morton curve over [lo, hi] (default: the positions' bounds)
        and returns the source index of each splat."""
        pos = cloud["positions"]
        n = len(pos)
        arrays = [cloud[k] for k in ("positions", "covariances", "colors", "alphas")]
        sh = cloud.get("sh")
        for a in arrays + ([sh] if sh is not None else []):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous and len(a) == n):
                raise ValueError("reorder_spatially needs contiguous float32 arrays of one length")
        lo = np.asarray(pos.min(0) if lo is None else lo, np.float32) if n else np.zeros(3, np.float32)
        hi = np.asarray(pos.max(0) if hi is None else hi, np.float32) if n else np.zeros(3, np.float32)
        order = np.zeros(max(n, 1), np.uint32)
        sh_floats = sh.shape[1] if sh is not None and n else 0
        ptr = lambda a: a.ctypes.data_as(_fp)
        self._check(self._lib.skm_reorder_spatially(self._h, ptr(arrays[0]), ptr(arrays[1]), ptr(arrays[2]),
                                                     ptr(arrays[3]), ptr(sh) if sh_floats else None, sh_floats, n,
                                                     ptr(lo), ptr(hi), order.ctypes.data_as(_u32p)))
        return order[:n]

    def argsort_device(self, keys_ptr, n, key_bits, order_ptr):
        ms = C.c_double()
        self._check(self._lib.skm_argsort_u32_device(self._h, C.c_void_p(keys_ptr), n, key_bits,
                                                     C.c_void_p(order_ptr), C.byref(ms)))
        return ms.value

    def argsort_device_static(self, keys_ptr, n, key_bits, order_ptr):
        """skm_argsort_u32_device_static: the same sort with its scratch in the library's static device
        array (no allocation); returns the status rather than raising, for the tests of its limits."""
        return self._lib.skm_argsort_u32_device_static(self._h, C.c_void_p(keys_ptr), n, key_bits,
                                                       C.c_void_p(order_ptr))

    def stats(self):
        L, h = self._lib, self._h
        return {
            "submitted": L.skm_submitted_frames(h), "completed": L.skm_completed_frames(h),
            "gpu_ms": L.skm_last_gpu_millis(h), "sort_ms": L.skm_last_sort_millis(h),
            "cull_ms": L.skm_last_cull_millis(h), "draw_count": L.skm_last_draw_count(h),
            "selected": L.skm_last_selected_count(h), "select_ms": L.skm_last_select_millis(h),
            "gpu_failed": L.skm_gpu_failed(h),
        }


class Server:
    """The request loop of include/skm_server.h: resident scenes, continuous batching under an HBM
    budget, BGRA8 outputs in host memory."""

    def __init__(self, device=0, max_batch_views=64, budget_bytes=0, sort_depth=16, max_in_flight=0):
        self._lib = lib()
        self._h = _S()
        cfg = ServerConfig(device, max_batch_views, budget_bytes, sort_depth, max_in_flight)
        rc = self._lib.skm_server_create(C.byref(cfg), C.byref(self._h))
        if rc != 0:
            raise SkmError(rc, "server creation failed")

    def close(self):
        if self._h:
            self._lib.skm_server_destroy(self._h)
            self._h = _S()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise SkmError(rc, self._lib.skm_server_last_error(self._h).decode())
        return rc

    def add_scene(self, cloud, max_sh_degree=3):
        view, keep = Renderer._view(cloud)
        sid = C.c_uint32()
        self._check(self._lib.skm_server_add_scene(self._h, C.byref(view), max_sh_degree, C.byref(sid)))
        return sid.value

    def submit(self, scene, view, proj, eye, width, height, user_data=0):
        q = Request()
        q.scene, q.user_data = scene, user_data
        q.view.view[:] = [float(x) for x in np.asarray(view, np.float32)]
        q.view.proj[:] = [float(x) for x in np.asarray(proj, np.float32)]
        q.view.camera_position[:] = [float(x) for x in np.asarray(eye, np.float32)[:3]]
        q.view.width, q.view.height = int(width), int(height)
        t = C.c_uint64()
        self._check(self._lib.skm_server_submit(self._h, C.byref(q), C.byref(t)))
        return t.value

    def poll(self, capacity=256, timeout_ms=0.0):
        out = (Completion * capacity)()
        n = C.c_uint32()
        self._check(self._lib.skm_server_poll(self._h, out, capacity, C.byref(n), timeout_ms))
        return [{"ticket": c.ticket, "user_data": c.user_data, "status": c.status, "batch_views": c.batch_views,
                 "queue_ms": c.queue_ms, "latency_ms": c.latency_ms} for c in out[:n.value]]

    def image(self, ticket):
        ptr = C.POINTER(C.c_uint8)()
        w, h = C.c_uint32(), C.c_uint32()
        self._check(self._lib.skm_server_image(self._h, ticket, C.byref(ptr), C.byref(w), C.byref(h)))
        return np.ctypeslib.as_array(ptr, shape=(h.value, w.value, 4)).copy()

    def release(self, ticket):
        self._check(self._lib.skm_server_release(self._h, ticket))

    def stats(self, reset=False):
        s = ServerStats()
        self._check(self._lib.skm_server_stats_read(self._h, C.byref(s), 1 if reset else 0))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def wait_all(self, tickets, timeout_s=120.0):
        """Polls until every ticket in `tickets` completed; returns their completions by ticket."""
        want, got = set(tickets), {}
        t0 = time.time()
        while want - set(got):
            for c in self.poll(timeout_ms=50.0):
                got[c["ticket"]] = c
            if time.time() - t0 > timeout_s:
                raise TimeoutError(f"{len(want - set(got))} requests still pending")
        return got


class LodTree:
    """A level-of-detail tree (include/skm_lod.h) built by the library; arrays are numpy copies."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("invalid LOD input")
        self._h = C.c_void_p(handle)
        self.view = LodTreeView()
        lib().skm_lod_view(self._h, C.byref(self.view))
        v = self.view
        n = v.nodes.count
        self.node_count = n
        self.leaf_count = v.leaf_count
        self.sh_degree = v.nodes.sh_degree
        arr = lambda p, shape, dt=np.float32: np.ctypeslib.as_array(p, shape=shape).astype(dt, copy=True) if n else np.zeros(shape, dt)
        self.nodes = {"positions": arr(v.nodes.positions, (n, 3)), "covariances": arr(v.nodes.covariances, (n, 6)),
                      "colors": arr(v.nodes.colors, (n, 3)), "alphas": arr(v.nodes.alphas, (n,)),
                      "sh_degree": self.sh_degree,
                      "sh": arr(v.nodes.sh, (n, sh_floats(self.sh_degree))) if self.sh_degree else None}
        self.layout = np.ctypeslib.as_array(C.cast(v.layout, _u32p), shape=(n, 6)).copy() if n else np.zeros((0, 6), np.uint32)
        k = v.cluster_count
        self.clusters = np.ctypeslib.as_array(C.cast(v.clusters, _u32p), shape=(k, 16)).copy() if k else np.zeros((0, 16), np.uint32)
        m = v.leaf_ref_count
        self.leaves = np.ctypeslib.as_array(v.leaves, shape=(m,)).copy() if m else np.zeros(0, np.uint32)
        self.rounds = lib().skm_lod_rounds(C.byref(self.view))

    @classmethod
    def build(cls, cloud, depth=21, cluster_leaves=32, threads=8):
        view, keep = Renderer._view(cloud)
        h = lib().skm_lod_build(C.byref(view), depth, cluster_leaves, threads)
        del keep
        return cls(h)

    @classmethod
    def from_layout(cls, nodes, layout, leaf_count):
        view, keep = Renderer._view(nodes)
        lay = np.ascontiguousarray(layout, np.uint32).reshape(-1, 6)
        h = lib().skm_lod_from_layout(C.byref(view), lay.ctypes.data_as(C.POINTER(LodNode)), leaf_count)
        del keep
        return cls(h)

    @classmethod
    def read(cls, path, max_sh_degree=3):
        """A tree from a LODSPLAT file (skm_lod_read); raises ValueError with the decoder's reason."""
        err = C.create_string_buffer(256)
        h = lib().skm_lod_read(os.fsencode(path), max_sh_degree, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    @classmethod
    def decode(cls, data, max_sh_degree=3):
        """A tree from LODSPLAT bytes (skm_lod_decode)."""
        buf = np.frombuffer(data, np.uint8)
        err = C.create_string_buffer(256)
        h = lib().skm_lod_decode(buf.ctypes.data_as(C.c_void_p), buf.nbytes, max_sh_degree, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    def write(self, path):
        err = C.create_string_buffer(256)
        rc = lib().skm_lod_write(self._h, os.fsencode(path), err, 256)
        if rc != 0:
            raise SkmError(rc, err.value.decode())

    def validate(self):
        """validateLodTree: the refinement rounds, or ValueError with the first failure."""
        err = C.create_string_buffer(256)
        depth = C.c_uint32()
        if lib().skm_lod_validate(self._h, C.byref(depth), err, 256) != 0:
            raise ValueError(err.value.decode())
        return depth.value

    def bounds(self):
        b = np.zeros(6, np.float32)
        lib().skm_lod_bounds(self._h, fptr(b))
        return b[:3].copy(), b[3:].copy()

    def close(self):
        if self._h:
            lib().skm_lod_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
