"""The GPU LOD tree cut against the CPU oracle, bit-exact in node indices and counters, on the
reference's selection-test trees and on built trees, then a whole LOD frame (cut, indexed cull,
sort, tiles, raster) against the oracle."""
import sys

import numpy as np
import pytest

import lod_cases as lc
from test_gpu_parity import assert_parity, gpu_frame

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def both(renderer, oracle, tree, view, proj, eye, w, h, capacity, pixel, cw=4.0, cull=True):
    cap = min(capacity or tree.leaf_count, tree.leaf_count)
    g_cut, g_stats = renderer.lod_select(tree, view, proj, w, h, cap, pixel, cw, cull)
    cam = oracle.make_camera(view, proj, w, h, eye)
    o_cut, o_stats = oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, cap, pixel, cw, cull)
    assert np.array_equal(g_cut, o_cut), (len(g_cut), len(o_cut))
    # count, packets, denied refinements and evaluated clusters (the final frontier and the last
    # round's accepted extra are bookkeeping of the last round).
    assert list(g_stats[[0, 2, 4, 5]]) == list(o_stats[[0, 2, 4, 5]]), (g_stats, o_stats)
    return g_cut, g_stats


def test_reference_selection_cases(renderer, skm_mod, oracle):
    say("case: MetalLodSelectionTest trees")
    tree = lc.hierarchy(skm_mod)
    view, proj, eye = lc.camera(0.0)
    for budget in (1, 7, 31, 32, 33, 64, 257, 400, 515):
        cut, _ = both(renderer, oracle, tree, view, proj, eye, lc.W, lc.H, budget, 0.0)
        assert 0 < len(cut) <= budget and lc.covers_every_leaf_once(tree, cut)
    far = lc.camera(1_000_000)
    assert len(both(renderer, oracle, tree, *far, lc.W, lc.H, 400, 1.0)[0]) == 1
    behind = lc.camera(-100)
    assert len(both(renderer, oracle, tree, *behind, lc.W, lc.H, 515, 0.0)[0]) == 0
    cut, stats = both(renderer, oracle, lc.quality_tree(skm_mod), *lc.camera(100), lc.W, lc.H, 4, 1.0, cull=False)
    assert set(cut.tolist()) == {1, 5, 6} and stats[4] == 0 and stats[5] == 3
    cut, stats = both(renderer, oracle, lc.rounds_tree(skm_mod), view, proj, eye, lc.W, lc.H, 3, 0.0, cull=False)
    assert list(cut) == [1, 3, 4] and stats[5] == 2
    packet = lc.packet_tree(skm_mod)
    cut, stats = both(renderer, oracle, packet, view, proj, eye, lc.W, lc.H, 513, 0.0)
    assert list(cut) == list(range(1, 514)) and stats[4] == 0
    cut, stats = both(renderer, oracle, packet, view, proj, eye, lc.W, lc.H, 512, 0.0)
    assert list(cut) == [0] and stats[4] == 1


@pytest.mark.parametrize("capacity", [20000, 32768])
def test_binary_tree_crosses_blocks(renderer, skm_mod, oracle, capacity):
    say(f"case: binary tree of 32768 leaves, capacity {capacity}")
    tree = lc.binary_tree(skm_mod)
    view, proj, eye = lc.camera(0.0)
    cut, stats = both(renderer, oracle, tree, view, proj, eye, lc.W, lc.H, capacity, 0.0, cull=False)
    assert len(cut) == capacity and lc.covers_every_leaf_once(tree, cut)


@pytest.mark.parametrize("head", [0, 64, 1 << 30])
def test_cut_launch_shapes_give_the_same_cut(renderer, skm_mod, oracle, head):
    """The single-block head of the cut (the leading rounds whose frontier can hold at most `head`
    clusters) against the oracle: none (every round grid-wide), a few rounds, and the whole cut in one
    block; a budget that denies refinements, so the budget search runs in both forms."""
    say(f"case: cut with head frontier {head}")
    renderer.set_lod_head(head)
    try:
        tree = lc.binary_tree(skm_mod)
        view, proj, eye = lc.camera(0.0)
        for capacity in (20000, 32768):
            cut, _ = both(renderer, oracle, tree, view, proj, eye, lc.W, lc.H, capacity, 0.0, cull=False)
            assert len(cut) == capacity and lc.covers_every_leaf_once(tree, cut)
        cloud = skm_mod.make_cloud(200_000, 1, seed=56, layout=1, spread=20.0, area=60.0)
        built = skm_mod.LodTree.build(cloud, 21, 32, threads=16)
        proj = skm_mod.perspective(np.pi / 3, 16 / 9, 0.2, 1000.0)
        for eye, pixel, budget in (((0, 3, 40), 1.5, 0), ((5, 2, 5), 0.5, 3_000)):
            both(renderer, oracle, built, skm_mod.look_at(eye), proj, np.asarray(eye, np.float32), 1920, 1080, budget,
                 pixel)
    finally:
        renderer.set_lod_head(4096)  # the default


def test_built_trees_at_many_cameras(renderer, skm_mod, oracle):
    say("case: built trees, 200k splats, cameras x thresholds x budgets")
    cloud = skm_mod.make_cloud(200_000, 1, seed=55, layout=1, spread=20.0, area=60.0)
    tree = skm_mod.LodTree.build(cloud, 21, 64, threads=16)
    proj = skm_mod.perspective(np.pi / 3, 16 / 9, 0.2, 1000.0)
    for k, (eye, pixel, budget) in enumerate([((0, 3, 40), 1.5, 0), ((10, 1, 5), 1.5, 0), ((0, 50, 1), 0.5, 50_000),
                                              ((-30, 2, -30), 4.0, 0), ((5, 2, 5), 1.0, 20_000)]):
        view = skm_mod.look_at(eye)
        both(renderer, oracle, tree, view, proj, np.asarray(eye, np.float32), 1920, 1080, budget, pixel)


@pytest.mark.parametrize("key16", [False, True])
def test_lod_world_frame_matches_oracle(renderer, skm_mod, oracle, key16):
    """A tree world's frame with 32-bit and with 16-bit depth keys (the two variants of the LOD
    fly-through; quantizedDepthKey, SplatVisibility.metal:21-28, on the cut's indexed cull)."""
    say(f"case: LOD world frame (cut -> indexed cull -> sort -> tiles -> raster), key16={key16}")
    cloud = skm_mod.make_cloud(150_000, 1, seed=66, layout=1, spread=10.0, area=30.0, log_scale_mean=-3.5)
    tree = skm_mod.LodTree.build(cloud, 21, 256, threads=16)
    renderer.upload_lod_world(tree, budget=60_000)
    assert renderer.world() == (tree.node_count, 1)
    eye = np.array([0, 2.0, 18.0], np.float32)
    view = skm_mod.look_at(eye)
    proj = skm_mod.perspective(np.pi / 3, 960 / 540, 0.2, 500.0)
    renderer.set_extent(960, 540)
    renderer.apply_policy(sort_depth=16 if key16 else 32, lod_error=1.5)
    renderer.render(view, proj, eye)
    cut, stats = renderer.read_lod_cut()
    cam = oracle.make_camera(view, proj, 960, 540, eye)
    o_cut, o_stats = oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, 60_000, 1.5, 4.0, True)
    assert np.array_equal(cut, o_cut)
    assert renderer.stats()["selected"] == len(o_cut)
    g = gpu_frame(renderer, skm_mod, None, view, proj, eye, 960, 540, upload=False, lod_error=1.5, key16=key16)
    o = oracle.frame(tree.nodes, view, proj, eye, 960, 540, active=o_cut, key16=key16)
    assert_parity(g, o, label=f"LOD frame key16={key16}")


def test_lod_budget_denials_raise_then_relax_the_threshold(skm_mod, oracle):
    """adaptLodThreshold (MetalFrameCompute.mm:166-172, LodThreshold.h:27-36) on the GPU path: under a
    budget the close-up cut cannot meet, every frame's cut equals the oracle's cut at the threshold
    the renderer reported before it, denials raise the threshold 10% a frame, and once the camera
    turns away (the cut sits under 90% of its limit) it relaxes 3% a frame back to the policy's."""
    say("case: LOD budget 6000 at 1 px, denials then relaxation, one frame in flight")
    cloud = skm_mod.make_cloud(150_000, 1, seed=68, layout=1, spread=10.0, area=30.0, log_scale_mean=-3.5)
    tree = skm_mod.LodTree.build(cloud, 21, 32, threads=16)
    budget = 6000
    w, h = 960, 540
    proj = skm_mod.perspective(np.pi / 3, w / h, 0.2, 500.0)
    eye = np.array([0, 2.0, 14.0], np.float32)
    r = skm_mod.Renderer(frames_in_flight=1)
    try:
        r.upload_lod_world(tree, budget=budget)
        r.set_extent(w, h)
        r.apply_policy(sort_depth=16, lod_error=1.0)
        policy = 1.0
        steps, seen = [], []
        for k in range(14):
            # Eight frames at the scene, then six looking away from it (most clusters culled).
            view = skm_mod.look_at(eye) if k < 8 else skm_mod.look_at(eye, (0.0, 2.0, 100.0))
            before = r.lod_threshold()
            r.render(view, proj, eye)
            cut, stats = r.read_lod_cut()
            cam = oracle.make_camera(view, proj, w, h, eye)
            o_cut, o_stats = oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, budget, before, 4.0, True)
            assert np.array_equal(cut, o_cut), (k, len(cut), len(o_cut))
            assert stats[4] == o_stats[4], (k, stats, o_stats)
            steps.append((int(stats[0]), int(stats[4]), budget))
            want = oracle.lod_threshold(policy, steps)[-1]
            after = r.lod_threshold()
            assert np.float32(after) == want, (k, after, want)
            seen.append((before, int(stats[0]), int(stats[4]), after))
        say(f"threshold trace (before, selected, denied, after): {seen}")
        assert all(s[2] > 0 for s in seen[:8]) and seen[7][3] > 1.8  # raised by denials
        assert all(s[2] == 0 and s[1] < budget // 10 * 9 for s in seen[8:])
        assert all(seen[k][3] < seen[k][0] for k in range(8, 14))  # relaxing
    finally:
        r.close()
        tree.close()


def test_lod_frames_in_flight_match_one_at_a_time(skm_mod):
    """Each frame slot has its own cut state: tree-world frames drawn with three in flight give every
    frame's counts and the last cut and image exactly as frames drawn one at a time (the budget is
    never reached, so the threshold stays at the policy's and does not depend on completion order)."""
    say("case: LOD world, 6 orbit frames, frames_in_flight 1 and 3")
    cloud = skm_mod.make_cloud(150_000, 1, seed=67, layout=1, spread=10.0, area=30.0, log_scale_mean=-3.5)
    tree = skm_mod.LodTree.build(cloud, 21, 256, threads=16)
    proj = skm_mod.perspective(np.pi / 3, 960 / 540, 0.2, 500.0)
    eyes = [np.array([6.0 * np.sin(k), 2.0, 18.0 * np.cos(0.3 * k)], np.float32) for k in range(6)]
    results = []
    for fif in (1, 3):
        r = skm_mod.Renderer(frames_in_flight=fif)
        try:
            r.upload_lod_world(tree)
            r.set_extent(960, 540)
            r.apply_policy(sort_depth=32, lod_error=1.5)
            r.frame_totals(reset=True)
            for eye in eyes:
                r.draw(r.make_frame(skm_mod.look_at(eye), proj, eye))
            r.finish()
            tot = r.frame_totals(reset=True)
            cut, stats = r.read_lod_cut()
            results.append((tot["frames"], tot["visible"], tot["pairs"], cut, r.read_frame()))
        finally:
            r.close()
    assert results[0][:3] == results[1][:3]
    assert np.array_equal(results[0][3], results[1][3])
    assert np.array_equal(results[0][4], results[1][4])
    assert results[0][0] == 6


def test_lod_world_from_a_lodsplat_file_matches_oracle(renderer, skm_mod, oracle, tmp_path):
    """The C3 input path from a file: a built tree written as LODSPLAT (version 2), read back by the
    decoder (decodeLodSplat's checks, node cap lifted), uploaded (uploadLodWorld) and drawn: the cut and
    the frame match the oracle's on the decoded tree."""
    say("case: LOD world read from a LODSPLAT file")
    cloud = skm_mod.make_cloud(120_000, 1, seed=69, layout=1, spread=10.0, area=30.0, log_scale_mean=-3.5)
    built = skm_mod.LodTree.build(cloud, 21, 256, threads=16)
    path = tmp_path / "world.lodsplat"
    built.write(str(path))
    tree = skm_mod.LodTree.read(str(path))
    assert tree.node_count == built.node_count and np.array_equal(tree.clusters, built.clusters)
    renderer.upload_lod_world(tree, budget=50_000)
    eye = np.array([3.0, 2.5, 16.0], np.float32)
    view = skm_mod.look_at(eye)
    proj = skm_mod.perspective(np.pi / 3, 960 / 540, 0.2, 500.0)
    renderer.set_extent(960, 540)
    renderer.apply_policy(sort_depth=16, lod_error=1.5)
    renderer.render(view, proj, eye)
    cut, _ = renderer.read_lod_cut()
    cam = oracle.make_camera(view, proj, 960, 540, eye)
    o_cut, _ = oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, 50_000, 1.5, 4.0, True)
    assert np.array_equal(cut, o_cut)
    g = gpu_frame(renderer, skm_mod, None, view, proj, eye, 960, 540, upload=False, lod_error=1.5, key16=True)
    o = oracle.frame(tree.nodes, view, proj, eye, 960, 540, active=o_cut, key16=True)
    assert_parity(g, o, label="LOD world from file")
    tree.close()
    built.close()
