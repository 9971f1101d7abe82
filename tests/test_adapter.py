"""The engine-side adapter (splatkit-ios_amd/adapter/HipSplatRenderer.h) compiled against the
reference engine's own headers (packages/splatkit-engine/include), on a machine without a GPU:

* every ABI mirror has its engine type's size and field offsets (static_assert in
  engine_harness.cpp, which the compile below runs);
* a world frame as configureWorldFrame fills it (SplatEngine.cpp:620-638) reaches skm_draw field for
  field: GPU order source, the ranges (reinterpreted in place), matrices, eye, reveal;
* the engine's load paths over the adapter (FrameOrderTest.cpp:112-158): a renderer without GPU
  sort fails the load with the engine's message, a tree on a renderer without GPU level of detail
  fails with its own, and HipSplatRenderer itself reports both, so the engine hands it ranges and
  hierarchies; with no HIP device it is not ready, reports a GPU failure and uploads nothing.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import adapter_harness as ah
import cases

pytestmark = pytest.mark.skipif(not os.path.isdir(ah.REFERENCE + "/packages"),
                                reason="no reference tree on this machine (the adapter compiles against its headers)")


def test_adapter_compiles_against_the_engine_headers():
    """A fresh compile of the adapter and its layout checks against the engine's headers."""
    inc = [f"-I{ah.REFERENCE}/packages/splat-core/include", f"-I{ah.REFERENCE}/packages/splatkit-engine/include",
           f"-I{ah.ROOT}/include"]
    out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-invalid-offsetof"]
                         + inc + [os.path.join(ah.ADAPTER, "engine_harness.cpp")],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert out.returncode == 0, out.stdout
    assert ah.available(), "build() makes splatkit-ios_amd/adapter/_build/libhsr.so where the reference is present"
    assert ah.lib().hsr_checked_layouts() == 9


def test_world_frame_reaches_the_abi_field_for_field(skm_mod):
    ranges = (skm_mod.Range * 3)(skm_mod.Range(0, 10), skm_mod.Range(40, 5), skm_mod.Range(7, 1))
    view = np.arange(16, dtype=np.float32) * np.float32(0.5)
    proj = cases.perspective(1.0, 1.5, 0.1, 100.0)
    eye = np.array([1.5, -2.0, 3.25], np.float32)
    rev = skm_mod.Reveal(0.5, 0.25, 0.125, 0.75, (C.c_float * 4)(0.1, 0.2, 0.3, 0.0))
    out = skm_mod.Frame()
    _, v = ah.f32(view)
    _, p = ah.f32(proj)
    _, e = ah.f32(eye)
    assert ah.lib().hsr_translate(ranges, 3, v, p, e, C.byref(rev), 2, C.byref(out)) == 1
    assert out.order_source == 1 and not out.order and out.order_count == 0
    assert C.addressof(out.ranges.contents) == C.addressof(ranges) and out.range_count == 3
    assert [(out.ranges[i].offset, out.ranges[i].count) for i in range(3)] == [(0, 10), (40, 5), (7, 1)]
    assert out.sh_degree == 2
    assert np.array_equal(np.array(out.view[:], np.float32), view)
    assert np.array_equal(np.array(out.proj[:], np.float32), proj)
    assert np.array_equal(np.array(out.camera_position[:], np.float32), eye)
    assert not out.label_styles
    assert (out.reveal.level, out.reveal.band, out.reveal.tint_amount, out.reveal.brightness_boost) == (0.5, 0.25, 0.125, 0.75)
    assert np.allclose(out.reveal.tint[:], [0.1, 0.2, 0.3, 0.0])


def one_splat():
    return cases.iso_cloud([(0, 0, 2)])


def test_renderer_without_gpu_sort_fails_the_load(skm_mod):
    """FrameOrderTest.cpp:112-126: the load fails with "GPU sort" and no world is uploaded."""
    s = ah.Session(gpu_sort=False)
    try:
        assert not s.probe()["sorts_on_gpu"]
        assert not s.load_cloud(one_splat())
        assert "GPU sort" in s.error()
        assert not s.probe()["world"]
    finally:
        s.close()


def test_lod_world_without_gpu_selection_fails_with_its_own_message(skm_mod):
    """FrameOrderTest.cpp:128-145: sorts on the GPU but selects no level of detail there."""
    cloud = skm_mod.make_cloud(2000, 0, seed=3, layout=0, spread=1.0)
    tree = skm_mod.LodTree.build(cloud, 21, 32, threads=1)
    s = ah.Session(gpu_lod=False)
    try:
        p = s.probe()
        assert p["sorts_on_gpu"] and not p["selects_lod_on_gpu"]
        assert not s.load_tree(tree, budget=1)
        assert "level of detail" in s.error()
        assert not s.probe()["world"]
    finally:
        s.close()
        tree.close()


def test_hip_renderer_takes_the_gpu_paths_and_fails_loudly_without_a_device(skm_mod):
    """HipSplatRenderer reports GPU sort and GPU level of detail (so the engine sends ranges and trees,
    FrameOrderTest.cpp:79-95,147-158); on this machine, with no HIP device, it is not ready, latches a
    GPU failure, refuses uploads and draws, and the engine's policy stays the backend fallback."""
    if skm_mod.lib().skm_renderer_create(None, C.byref(C.c_void_p())) == 0:
        pytest.skip("a HIP device is present")
    s = ah.Session()
    try:
        p = s.probe()
        assert p["sorts_on_gpu"] and p["selects_lod_on_gpu"]
        assert not p["ready"] and p["gpu_failed"] and not p["world"]
        assert not s.load_cloud(one_splat()) and "GPU upload failed" in s.error()
        ok, ranges = s.render(cases.identity(), cases.PIX_PROJ, np.zeros(3, np.float32))
        assert not ok and ranges == []
        assert s.counters()["submitted"] == 0
    finally:
        s.close()
