"""Grids sized from earlier frames. After the first frame of a world, the renderer sizes the grids of
the radix passes, binning and ranges from the latest finished frame's counts (cull lanes, survivors,
tile pairs); a frame with much more work than that loops their blocks over the rest. Such a frame,
drawn right after a light one, must still match the oracle bit for bit."""
import sys

import numpy as np
import pytest

from test_gpu_parity import assert_parity, gpu_frame

pytestmark = pytest.mark.gpu

SURVIVOR_SLACK = 8 * 3072  # the renderer's estimate: 1.5 x the last count + 8 sort tiles
LANE_SLACK = 8 * 512


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_heavy_frame_after_a_light_one(renderer, skm_mod, oracle):
    say("case: 300k SH1 world, a frame looking away, then one facing the scene")
    cloud = skm_mod.make_cloud(300_000, 1, seed=71, layout=1, spread=3.0, area=6.0)
    w, h = 1280, 720
    proj = skm_mod.perspective(np.pi / 3, w / h, 0.2, 100.0)
    light_eye = np.array([0.0, 2.0, 14.0], np.float32)
    light = gpu_frame(renderer, skm_mod, cloud, skm_mod.look_at(light_eye, (0.0, 2.0, 40.0)), proj, light_eye, w, h,
                      sh_degree=1)
    eye = np.array([0.0, 2.0, 7.0], np.float32)
    view = skm_mod.look_at(eye)
    heavy = gpu_frame(renderer, skm_mod, None, view, proj, eye, w, h, upload=False, sh_degree=1)
    say(f"light frame: {light['count']} survivors, {light['pairs']} pairs; heavy: {heavy['count']}, {heavy['pairs']}")
    # The heavy frame's survivors and pairs exceed the estimates made from the light one: its sort,
    # binning and range kernels loop.
    assert heavy["count"] > 1.5 * light["count"] + 2 * SURVIVOR_SLACK
    assert heavy["pairs"] > 1.5 * light["pairs"] + 2 * SURVIVOR_SLACK
    o = oracle.frame(cloud, view, proj, eye, w, h)
    assert_parity(heavy, o, label="heavy frame after a light one")


def test_large_cut_after_a_small_one(renderer, skm_mod, oracle):
    say("case: LOD world, a frame with a small cut, then one with a large cut")
    cloud = skm_mod.make_cloud(150_000, 1, seed=72, layout=1, spread=10.0, area=30.0, log_scale_mean=-3.5)
    tree = skm_mod.LodTree.build(cloud, 21, 256, threads=16)
    budget = 60_000
    renderer.upload_lod_world(tree, budget=budget)
    w, h = 960, 540
    proj = skm_mod.perspective(np.pi / 3, w / h, 0.2, 500.0)
    renderer.set_extent(w, h)
    renderer.apply_policy(sort_depth=32, lod_error=1.5)
    away_eye = np.array([0.0, 2.0, 18.0], np.float32)  # facing away from the scene: a small cut
    renderer.render(skm_mod.look_at(away_eye, (0.0, 2.0, 60.0)), proj, away_eye)
    small = len(renderer.read_lod_cut()[0])
    small_pairs = renderer.read_tiles()[0]
    eye = np.array([0.0, 2.0, 18.0], np.float32)
    view = skm_mod.look_at(eye)
    g = gpu_frame(renderer, skm_mod, None, view, proj, eye, w, h, upload=False, lod_error=1.5)
    cut, _ = renderer.read_lod_cut()
    say(f"small cut: {small} indices, {small_pairs} pairs; large cut: {len(cut)} indices, {g['count']} survivors, "
        f"{g['pairs']} pairs")
    # The large cut's first depth pass (over the cut's lanes) and its tile pairs exceed the estimates
    # made from the small cut: those kernels loop.
    assert len(cut) > 1.5 * (small + 512) + 2 * LANE_SLACK
    assert g["pairs"] > 1.5 * small_pairs + 2 * SURVIVOR_SLACK
    cam = oracle.make_camera(view, proj, w, h, eye)
    o_cut, _ = oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, budget, 1.5, 4.0, True)
    assert np.array_equal(cut, o_cut)
    o = oracle.frame(tree.nodes, view, proj, eye, w, h, active=o_cut)
    assert_parity(g, o, label="large cut after a small one")
