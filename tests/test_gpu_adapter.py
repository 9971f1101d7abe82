"""The engine's renderer-facing steps (SplatEngine.cpp:29-56, 226-267, 429-449, 559-638) driving the
engine-side adapter HipSplatRenderer (compiled against the reference engine's headers) on one
MI355X, against the CPU oracle: a flat world drawn with its whole range every frame
(FrameOrderTest.cpp:79-95), a tree world whose budget reaches the renderer
(FrameOrderTest.cpp:147-158), and a tiled world whose tiles arrive into a slab."""
import sys

import numpy as np
import pytest

import adapter_harness as ah

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ah.available(), reason="libhsr.so is built only where the reference tree is")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def close_to_oracle(img, o, label):
    """BGRA8 capture against the oracle's FP64 image: within 1/255 plus the 8-bit rounding."""
    want = np.clip(o["image"][..., [2, 1, 0]], 0.0, 1.0) * 255.0
    err = np.abs(img[..., :3].astype(np.float64) - want).max()
    say(f"{label}: visible {o['count']} max |capture - 255 fp64| = {err:.3f}")
    assert err <= 1.5, label


def test_flat_world_frames_through_the_adapter(skm_mod, oracle):
    say("case: engine loop over HipSplatRenderer, flat world, two frames")
    cloud = skm_mod.make_cloud(80_000, 2, seed=121, layout=0, spread=1.5, log_scale_mean=-3.2)
    s = ah.Session(frames_in_flight=2)
    try:
        p = s.probe()
        assert p["ready"] == 0 and p["policy_sort_depth"] == 16  # no extent yet; the fallback policy's keys
        assert s.policy(sort_depth=32), s.error()
        assert s.probe()["policy_sort_depth"] == 32
        assert s.extent(640, 360) and s.probe()["ready"]
        assert s.load_cloud(cloud), s.error()
        frames = [skm_mod.orbit_camera(k, 640, 360, radius=4.0, height_m=0.6) for k in (3, 11)]
        for view, proj, eye in frames:
            ok, ranges = s.render(view, proj, eye)
            assert ok, s.error()
            assert ranges == [(0, 80_000)]  # the world's range, every frame
        assert s.wait() and s.counters()["submitted"] == 2 and s.counters()["completed"] == 2
        o = oracle.frame(cloud, *frames[1], 640, 360)
        assert s.counters()["draw_count"] == o["count"]
        close_to_oracle(s.capture(640, 360), o, "flat world")
    finally:
        s.close()


def test_tree_world_budget_reaches_the_renderer(skm_mod, oracle):
    say("case: engine loop over HipSplatRenderer, tree world, budget 3000")
    cloud = skm_mod.make_cloud(60_000, 1, seed=122, layout=1, spread=6.0, area=20.0, log_scale_mean=-3.5)
    tree = skm_mod.LodTree.build(cloud, 21, 32, threads=16)
    s = ah.Session(frames_in_flight=1)
    try:
        assert s.policy(sort_depth=16, lod_error=1.0)
        assert s.extent(800, 450)
        assert s.load_tree(tree, budget=3000), s.error()
        eye = np.array([0, 2.0, 14.0], np.float32)
        view = skm_mod.look_at(eye)
        proj = skm_mod.perspective(np.pi / 3, 800 / 450, 0.2, 500.0)
        ok, ranges = s.render(view, proj, eye)
        assert ok and s.wait()
        cnt = s.counters()
        assert 0 < cnt["selected"] <= 3000
        cam = oracle.make_camera(view, proj, 800, 450, eye)
        cut, _ = oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, 3000, 1.0, 4.0, True)
        assert cnt["selected"] == len(cut)
        o = oracle.frame(tree.nodes, view, proj, eye, 800, 450, active=cut, key16=True)
        close_to_oracle(s.capture(800, 450), o, "tree world")
    finally:
        s.close()
        tree.close()


def test_tiled_world_draws_the_arrived_tiles(skm_mod, oracle):
    say("case: engine loop over HipSplatRenderer, slab of 3 tiles arriving between frames")
    cloud = skm_mod.make_cloud(60_000, 1, seed=123, layout=0, spread=1.2, log_scale_mean=-3.0)
    part = lambda lo, hi: {k: (v[lo:hi] if isinstance(v, np.ndarray) else v) for k, v in cloud.items()}
    s = ah.Session(frames_in_flight=2)
    try:
        assert s.policy(sort_depth=32) and s.extent(480, 320)
        assert s.load_slab(60_000, 1), s.error()
        view, proj, eye = skm_mod.orbit_camera(2, 480, 320, radius=3.0, height_m=0.4)
        assert s.arrive(0, part(0, 20_000)) and s.arrive(40_000, part(40_000, 60_000))
        ok, ranges = s.render(view, proj, eye)
        assert ok and ranges == [(0, 20_000), (40_000, 20_000)]
        assert s.arrive(20_000, part(20_000, 40_000))  # the frame in flight draws the other two tiles
        ok, ranges = s.render(view, proj, eye)
        assert ok and ranges == [(0, 20_000), (40_000, 20_000), (20_000, 20_000)]
        assert s.wait()
        o = oracle.frame(cloud, view, proj, eye, 480, 320, ranges=ranges)
        close_to_oracle(s.capture(480, 320), o, "tiled world")
    finally:
        s.close()
