"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and, where built, of the reference's own
CPU sources (oracle/_ref/libref.so). Test infrastructure only: tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")
TILE = 16
T_STOP = 1.0 / 4096.0

_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("focal", C.c_float * 2),
                ("tan_half_fov", C.c_float * 2), ("screen_size", C.c_float * 2), ("output_linear", C.c_uint32),
                ("min_pixel_radius", C.c_float), ("camera_position", C.c_float * 4), ("reveal", C.c_float * 4),
                ("reveal_tint", C.c_float * 4)]


assert C.sizeof(Camera) == 208

_API = {
    "orc_to_half": (C.c_uint16, [C.c_float]),
    "orc_half_rne": (C.c_uint16, [C.c_float]),
    "orc_from_half": (C.c_float, [C.c_uint16]),
    "orc_pack_splats": (None, [_fp, _fp, _fp, _fp, C.c_uint32, _u32p]),
    "orc_sh_stride": (C.c_uint32, [C.c_int]),
    "orc_pack_sh": (None, [_fp, C.c_uint32, C.c_int, C.c_int, _u32p]),
    "orc_make_camera": (None, [_fp, _fp, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _fp, _fp,
                               C.POINTER(Camera)]),
    "orc_cull": (C.c_uint32, [C.POINTER(Camera), _u32p, _u32p, C.c_int, _u32p, C.c_uint32, _u32p, C.c_uint32,
                              C.c_int, _u32p, _u32p]),
    "orc_project": (C.c_int, [C.POINTER(Camera), _u32p, _u32p, C.c_int, C.c_uint32, _u32p]),
    "orc_depth_key16": (C.c_uint32, [C.POINTER(Camera), C.c_float]),
    "orc_argsort": (None, [_u32p, C.c_uint32, _u32p]),
    "orc_morton_code": (C.c_uint32, [_fp, _fp, _fp]),
    "orc_spatial_order": (None, [_fp, C.c_uint32, _fp, _fp, _u32p]),
    "orc_bin": (C.c_uint64, [_u32p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _u32p,
                             C.c_uint64, _u32p]),
    "orc_tile_rect": (C.c_uint32, [_u32p, C.c_uint32, C.c_uint32, C.c_uint32, _i32p]),
    "orc_raster64": (None, [_u32p, _u32p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, _fp, _fp]),
    "orc_raster32": (None, [_u32p, _u32p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, _fp, C.c_float,
                            C.c_uint32, _fp]),
    "orc_decide": (C.c_int, [_u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _fp]),
    "orc_render_frame": (C.c_int, [C.POINTER(Camera), _u32p, _u32p, C.c_int, C.c_uint32, C.c_int, C.c_uint32,
                                   _fp, C.c_uint32, _fp, C.POINTER(C.c_double), _u32p, C.POINTER(C.c_uint64)]),
    "orc_lod_select": (C.c_uint32, [C.POINTER(Camera), _u32p, C.c_uint32, _u32p, C.c_uint32, C.c_uint32, C.c_float,
                                    C.c_float, C.c_int, _u32p, _u32p]),
    "orc_lod_cluster": (C.c_int, [C.POINTER(Camera), _u32p, C.c_float, _fp]),
    "orc_missed_coverage": (C.c_uint64, [_u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _u32p,
                                         C.POINTER(C.c_uint64)]),
    "orc_region_keep": (C.c_int, [_u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_region_misses": (C.c_uint64, [_u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_uint64)]),
    "orc_lod_threshold": (None, [C.c_float, C.c_uint32, _u32p, _u32p, _u32p, _fp]),
    "orc_log_d": (C.c_double, [C.c_double]),
    "orc_log_f": (C.c_float, [C.c_float]),
    "orc_exp_d": (C.c_double, [C.c_double]),
}

_REF_API = {
    "ref_to_half": (C.c_uint16, [C.c_float]),
    "ref_morton_code": (C.c_uint32, [_fp, _fp, _fp]),
    "ref_spatial_order": (None, [_fp, C.c_uint32, _fp, _fp, _u32p]),
    "ref_from_half": (C.c_float, [C.c_uint16]),
    "ref_pack_splats": (None, [C.c_uint32, _fp, _fp, _fp, _fp, _u32p]),
    "ref_sh_stride": (C.c_uint32, [C.c_int]),
    "ref_pack_sh": (None, [C.c_uint32, C.c_int, _fp, _fp, C.c_int, _u32p]),
    "ref_covariance": (None, [_fp, _fp, _fp]),
    "ref_perspective": (None, [C.c_float, C.c_float, C.c_float, C.c_float, _fp]),
    "ref_lod_threshold": (None, [C.c_float, C.c_uint32, _u32p, _u32p, _u32p, _fp]),
    "ref_lod_build": (C.c_void_p, [C.c_uint32, C.c_int, _fp, _fp, _fp, _fp, _fp, C.c_uint32, C.c_uint32]),
    "ref_lod_free": (None, [C.c_void_p]),
    "ref_lod_counts": (None, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "ref_lod_read": (None, [C.c_void_p, _fp, _fp, _fp, _fp, _fp, _u32p, _u32p, _u32p]),
    "ref_lod_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]),
    "ref_lod_decode": (C.c_void_p, [C.c_void_p, C.c_size_t, C.c_int, C.c_char_p, C.c_size_t]),
    "ref_lod_header": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]),
    "ref_lod_validate": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_char_p, C.c_size_t]),
    "ref_lod_bounds": (None, [C.c_void_p, _fp]),
}

_lib = None
_ref = None


def build():
    """Compiles liboracle.so (and the reference library when the reference tree is present)."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            build()
        h = C.CDLL(ORACLE_PATH)
        for name, (res, args) in _API.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def ref():
    """The reference's own CPU code, or None where it was not built (e.g. no reference tree)."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        h = C.CDLL(REF_PATH)
        for name, (res, args) in _REF_API.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _ref = h
    return _ref


def fp(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def up(a):
    return a.ctypes.data_as(_u32p) if a is not None else None


def pack_splats(cloud):
    n = len(cloud["alphas"])
    out = np.zeros((n, 8), np.uint32)
    lib().orc_pack_splats(fp(np.ascontiguousarray(cloud["positions"], np.float32)),
                          fp(np.ascontiguousarray(cloud["covariances"], np.float32)),
                          fp(np.ascontiguousarray(cloud["colors"], np.float32)),
                          fp(np.ascontiguousarray(cloud["alphas"], np.float32)), n, up(out))
    return out


def pack_sh(cloud, degree):
    n = len(cloud["alphas"])
    stride = lib().orc_sh_stride(degree)
    out = np.zeros((n, max(stride, 1)), np.uint32)
    if degree > 0:
        lib().orc_pack_sh(fp(np.ascontiguousarray(cloud["sh"], np.float32)), n, cloud["sh_degree"], degree, up(out))
    return out


def make_camera(view, proj, width, height, cam_pos, output_linear=0, min_pixel_radius=1.0, reveal=None):
    cam = Camera()
    v = np.ascontiguousarray(view, np.float32)
    p = np.ascontiguousarray(proj, np.float32)
    c = np.ascontiguousarray(np.asarray(cam_pos, np.float32)[:3])
    r = np.ascontiguousarray(reveal, np.float32) if reveal is not None else None
    lib().orc_make_camera(fp(v), fp(p), width, height, output_linear, min_pixel_radius, fp(c), fp(r), C.byref(cam))
    return cam


def cull(cam, splats, sh, degree, ranges=None, active=None, key16=False):
    n = len(splats)
    splats = np.ascontiguousarray(splats, np.uint32)
    sh = np.ascontiguousarray(sh, np.uint32) if sh is not None else np.zeros((n, 1), np.uint32)
    if active is not None:
        active = np.ascontiguousarray(active, np.uint32)
        lanes = len(active)
        rr, rc = None, 0
    else:
        if ranges is None:
            ranges = [(0, n)]
        rr = np.ascontiguousarray(np.asarray(ranges, np.uint32).reshape(-1, 2))
        rc = len(rr)
        lanes = int(rr[:, 1].sum()) if rc else 0
    keys = np.zeros(max(lanes, 1), np.uint32)
    proj = np.zeros((max(lanes, 1), 8), np.uint32)
    count = lib().orc_cull(C.byref(cam), up(splats), up(sh), degree, up(rr), rc, up(active),
                           len(active) if active is not None else 0, 1 if key16 else 0, up(keys), up(proj))
    return count, keys[:count].copy(), proj[:count].copy()


def argsort(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    order = np.zeros(max(len(keys), 1), np.uint32)
    lib().orc_argsort(up(keys), len(keys), up(order))
    return order[:len(keys)]


def morton_code(p, lo, hi, f=None):
    """mortonCode (SpatialOrder.cpp:39-44); f: another library with the same entry (the reference's)."""
    args = [np.ascontiguousarray(x, np.float32) for x in (p, lo, hi)]
    return int((f or lib().orc_morton_code)(*[fp(a) for a in args]))


def spatial_order(positions, lo, hi, f=None):
    """reorderSpatially's permutation (SpatialOrder.cpp:48-84): source index of each splat."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    lo, hi = np.ascontiguousarray(lo, np.float32), np.ascontiguousarray(hi, np.float32)
    order = np.zeros(max(len(pos), 1), np.uint32)
    (f or lib().orc_spatial_order)(fp(pos), len(pos), fp(lo), fp(hi), up(order))
    return order[:len(pos)]


def bin_tiles(projected, order, width, height, tile=TILE):
    projected = np.ascontiguousarray(projected, np.uint32).reshape(-1, 8)
    order = np.ascontiguousarray(order, np.uint32)
    n = len(order)
    pairs = lib().orc_bin(up(projected), up(order), n, width, height, tile, None, None, 0, None)
    tiles = ((width + tile - 1) // tile) * ((height + tile - 1) // tile)
    pt = np.zeros(max(pairs, 1), np.uint32)
    pv = np.zeros(max(pairs, 1), np.uint32)
    ranges = np.zeros((tiles, 2), np.uint32)
    lib().orc_bin(up(projected), up(order), n, width, height, tile, up(pt), up(pv), pairs, up(ranges))
    return pairs, pt[:pairs].copy(), pv[:pairs].copy(), ranges


def tile_rect(record, width, height, tile=TILE):
    rect = np.zeros(4, np.int32)
    n = lib().orc_tile_rect(up(np.ascontiguousarray(record, np.uint32)), width, height, tile,
                            rect.ctypes.data_as(_i32p))
    return n, rect


def raster64(projected, pair_values, ranges, width, height, background=(0.05, 0.05, 0.08), tile=TILE):
    out = np.zeros((height, width, 4), np.float32)
    bg = np.asarray(background, np.float32)
    lib().orc_raster64(up(np.ascontiguousarray(projected, np.uint32)), up(np.ascontiguousarray(pair_values, np.uint32)),
                       up(np.ascontiguousarray(ranges, np.uint32)), width, height, tile, fp(bg), fp(out))
    return out


def raster32(projected, pair_values, ranges, width, height, background=(0.05, 0.05, 0.08), tile=TILE, threads=1):
    out = np.zeros((height, width, 4), np.float32)
    bg = np.asarray(background, np.float32)
    lib().orc_raster32(up(np.ascontiguousarray(projected, np.uint32)), up(np.ascontiguousarray(pair_values, np.uint32)),
                       up(np.ascontiguousarray(ranges, np.uint32)), width, height, tile, fp(bg), T_STOP, threads,
                       fp(out))
    return out


def missed_coverage(projected, count, pair_values, ranges, width, height, tile=TILE):
    """(missed, accepted): accepted (pixel, record) pairs of the coverage rule, and how many of
    them are missing from their pixel's tile list."""
    acc = C.c_uint64()
    pv = np.ascontiguousarray(pair_values, np.uint32)
    if pv.size == 0:
        pv = np.zeros(1, np.uint32)
    missed = lib().orc_missed_coverage(up(np.ascontiguousarray(projected, np.uint32)), count, width, height, tile,
                                       up(pv), up(np.ascontiguousarray(ranges, np.uint32)), C.byref(acc))
    return int(missed), int(acc.value)


def region_misses(projected, count, width, height, region):
    acc = C.c_uint64()
    missed = lib().orc_region_misses(up(np.ascontiguousarray(projected, np.uint32)), count, width, height, region,
                                     C.byref(acc))
    return int(missed), int(acc.value)


def decide(record, width, height, px, py):
    power = C.c_float()
    take = lib().orc_decide(up(np.ascontiguousarray(record, np.uint32)), width, height, px, py, C.byref(power))
    return bool(take), power.value


def frame(cloud, view, proj, cam_pos, width, height, key16=False, min_pixel_radius=1.0,
          background=(0.05, 0.05, 0.08), ranges=None, reveal=None, output_linear=0, image=True, active=None):
    """The whole oracle frame: packing, cull, depth order, tile pairs, ranges and the FP64 image."""
    degree = cloud.get("sh_degree", 0) if cloud.get("sh") is not None else 0
    splats = pack_splats(cloud)
    sh = pack_sh(cloud, degree)
    cam = make_camera(view, proj, width, height, cam_pos, output_linear, min_pixel_radius, reveal)
    count, keys, projected = cull(cam, splats, sh, degree, ranges=ranges, active=active, key16=key16)
    order = argsort(keys)
    pairs, pt, pv, tranges = bin_tiles(projected, order, width, height)
    out = {"count": count, "keys": keys, "projected": projected, "order": order, "pairs": pairs,
           "pair_tiles": pt, "pair_values": pv, "ranges": tranges, "camera": cam}
    if image:
        out["image"] = raster64(projected, pv, tranges, width, height, background)
    return out


def render_frame_timed(cam, splats, sh, degree, key16, background, threads):
    """The CPU baseline: one whole frame (cull, sort, bin, raster32); returns (seconds[4], visible, pairs)."""
    n = len(splats)
    w = int(cam.screen_size[0])
    h = int(cam.screen_size[1])
    out = np.zeros((h, w, 4), np.float32)
    secs = (C.c_double * 4)()
    vis = C.c_uint32()
    pairs = C.c_uint64()
    bg = np.asarray(background, np.float32)
    lib().orc_render_frame(C.byref(cam), up(np.ascontiguousarray(splats, np.uint32)),
                           up(np.ascontiguousarray(sh, np.uint32)), degree, n, 1 if key16 else 0, TILE, fp(bg),
                           threads, fp(out), secs, C.byref(vis), C.byref(pairs))
    return list(secs), vis.value, pairs.value, out


def lod_select(cam, clusters, leaves, rounds, capacity, pixel_limit=1.0, color_weight=4.0, cull=True):
    """The oracle's LOD cut: (node indices, stats6)."""
    clusters = np.ascontiguousarray(clusters, np.uint32).reshape(-1, 16)
    leaves = np.ascontiguousarray(leaves, np.uint32)
    out = np.zeros(max(1, capacity), np.uint32)
    stats = np.zeros(6, np.uint32)
    n = lib().orc_lod_select(C.byref(cam), up(clusters), len(clusters), up(leaves if len(leaves) else np.zeros(1, np.uint32)),
                             rounds, capacity, pixel_limit, color_weight, 1 if cull else 0, up(out), up(stats))
    return out[:n].copy(), stats


def lod_threshold(policy, steps, f=None):
    """LodThreshold over (selected, limited, limit) steps: the threshold after each. f: another library
    with the same entry (the reference's ref_lod_threshold)."""
    sel = np.ascontiguousarray([s[0] for s in steps], np.uint32)
    lim = np.ascontiguousarray([s[1] for s in steps], np.uint32)
    cap = np.ascontiguousarray([s[2] for s in steps], np.uint32)
    out = np.zeros(max(1, len(steps)), np.float32)
    (f or lib().orc_lod_threshold)(policy, len(steps), up(sel), up(lim), up(cap), fp(out))
    return out[:len(steps)]
