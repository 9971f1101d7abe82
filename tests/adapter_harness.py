"""ctypes wrapper of splatkit-ios_amd/adapter/_build/libhsr.so: HipSplatRenderer, the engine-side
adapter, compiled against the reference engine's headers, and the engine's renderer-facing steps
over it (splatkit-ios_amd/adapter/engine_harness.cpp)."""
import ctypes as C
import os

import numpy as np

import skm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADAPTER = os.path.join(ROOT, "splatkit-ios_amd", "adapter")
LIB_PATH = os.path.join(ADAPTER, "_build", "libhsr.so")
REFERENCE = "/root/reference"

_fp = C.POINTER(C.c_float)
_API = {
    "hsr_checked_layouts": (C.c_int, []),
    "hsr_translate": (C.c_int, [C.POINTER(skm.Range), C.c_uint32, _fp, _fp, _fp, C.POINTER(skm.Reveal), C.c_int,
                                C.POINTER(skm.Frame)]),
    "hsr_session_create": (C.c_void_p, [C.c_int, C.c_uint32, C.c_int, C.c_int]),
    "hsr_session_destroy": (None, [C.c_void_p]),
    "hsr_session_error": (C.c_char_p, [C.c_void_p]),
    "hsr_session_probe": (None, [C.c_void_p, C.POINTER(C.c_int32)]),
    "hsr_session_policy": (C.c_int, [C.c_void_p, C.c_uint32, C.c_float, C.c_uint32, C.c_float]),
    "hsr_session_extent": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "hsr_session_load_cloud": (C.c_int, [C.c_void_p, C.POINTER(skm.CloudView), C.c_int]),
    "hsr_session_load_tree": (C.c_int, [C.c_void_p, C.POINTER(skm.LodTreeView), C.c_int, C.c_uint32]),
    "hsr_session_load_slab": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_int]),
    "hsr_session_arrive": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(skm.CloudView)]),
    "hsr_session_render": (C.c_int, [C.c_void_p, _fp, _fp, _fp, C.c_int, C.POINTER(skm.Range), C.c_uint32,
                                     C.POINTER(C.c_uint32)]),
    "hsr_session_wait": (C.c_int, [C.c_void_p, C.c_double]),
    "hsr_session_counters": (None, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "hsr_session_capture": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "hsr_session_renderer": (C.c_void_p, [C.c_void_p]),
}
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _API.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def f32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(_fp)


class Session:
    """The engine loop over one HipSplatRenderer."""

    def __init__(self, device=0, frames_in_flight=2, gpu_sort=True, gpu_lod=True):
        self.L = lib()
        self.h = self.L.hsr_session_create(device, frames_in_flight, int(gpu_sort), int(gpu_lod))
        self._keep = []

    def close(self):
        if self.h:
            self.L.hsr_session_destroy(self.h)
            self.h = None

    def error(self):
        return (self.L.hsr_session_error(self.h) or b"").decode()

    def probe(self):
        f = (C.c_int32 * 8)()
        self.L.hsr_session_probe(self.h, f)
        keys = ("sorts_on_gpu", "selects_lod_on_gpu", "ready", "gpu_failed", "world", "policy_sort_depth",
                "max_texture_dimension", "described")
        return dict(zip(keys, list(f)))

    def policy(self, sort_depth=32, lod_error=1.0, lod_limit=0, subpixel=1.0):
        return bool(self.L.hsr_session_policy(self.h, sort_depth, lod_error, lod_limit, subpixel))

    def extent(self, w, h):
        return bool(self.L.hsr_session_extent(self.h, w, h))

    def load_cloud(self, cloud, max_sh=3):
        view, keep = skm.Renderer._view(cloud)
        return bool(self.L.hsr_session_load_cloud(self.h, C.byref(view), max_sh))

    def load_tree(self, tree, max_sh=3, budget=0):
        return bool(self.L.hsr_session_load_tree(self.h, C.byref(tree.view), max_sh, budget))

    def load_slab(self, residency, sh_degree, max_sh=3):
        return bool(self.L.hsr_session_load_slab(self.h, residency, sh_degree, max_sh))

    def arrive(self, offset, cloud):
        view, keep = skm.Renderer._view(cloud)
        return bool(self.L.hsr_session_arrive(self.h, offset, C.byref(view)))

    def render(self, view, proj, eye, sh_degree=3):
        _, v = f32(view)
        _, p = f32(proj)
        _, e = f32(eye)
        out = (skm.Range * 65536)()
        n = C.c_uint32()
        ok = self.L.hsr_session_render(self.h, v, p, e, sh_degree, out, 65536, C.byref(n))
        return bool(ok), [(out[i].offset, out[i].count) for i in range(n.value)]

    def wait(self, timeout_s=60.0):
        return bool(self.L.hsr_session_wait(self.h, timeout_s))

    def counters(self):
        c = (C.c_uint64 * 4)()
        self.L.hsr_session_counters(self.h, c)
        return dict(zip(("submitted", "completed", "draw_count", "selected"), list(c)))

    def capture(self, w, h):
        out = np.zeros((h, w, 4), np.uint8)
        ok = self.L.hsr_session_capture(self.h, out.ctypes.data_as(C.c_void_p), out.nbytes)
        return out if ok else None

    def renderer_handle(self):
        return C.c_void_p(self.L.hsr_session_renderer(self.h))
