"""View batches on one MI355X, through the C ABI: every view of a batch (one graph launch for all
views) must give the single-view frame's survivors, depth keys, projected records, depth order,
(tile | depth) pairs, tile ranges and pixels, bit for bit, and the single-view frames are in turn
checked against the CPU oracle (test_gpu_parity.py, and here for a sample view)."""
import math
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single(r, view, key16):
    v, p, eye, w, h = view
    r.set_extent(w, h)
    r.render(v, p, eye)
    count, keys, proj = r.read_visible()
    order = r.read_order(count)
    pairs, pt, pv, ranges = r.read_tiles()
    return {"count": count, "keys": keys, "projected": proj, "order": order, "pairs": pairs, "pair_tiles": pt,
            "pair_values": pv, "ranges": ranges, "image": r.read_frame()}


def batched(r, k):
    count, keys, proj = r.read_view_visible(k)
    order = r.read_view_order(k, count)
    pairs, pt, pv, ranges = r.read_view_tiles(k)
    return {"count": count, "keys": keys, "projected": proj, "order": order, "pairs": pairs, "pair_tiles": pt,
            "pair_values": pv, "ranges": ranges, "image": r.read_view_frame(k)}


def assert_same(b, s, label):
    for key in ("count", "pairs"):
        assert b[key] == s[key], f"{label}: {key} {b[key]} vs {s[key]}"
    for key in ("keys", "projected", "order", "pair_tiles", "pair_values", "ranges"):
        assert np.array_equal(b[key], s[key]), f"{label}: {key} differ"
    assert np.array_equal(b["image"].view(np.uint32), s["image"].view(np.uint32)), f"{label}: pixels differ"


def small_views(skm_mod, count, seed, resolutions=((333, 217), (640, 360), (256, 256), (1280, 720), (97, 131))):
    return skm_mod.views_c2(count, seed=seed, radius=(3.0, 9.0), resolutions=resolutions)


@pytest.mark.parametrize("key16", [False, True])
def test_batch_views_equal_single_draws(renderer, skm_mod, key16):
    say(f"case: 7 views of a 200k SH2 scene, key16={key16}")
    cloud = skm_mod.make_cloud(200_000, 2, seed=41, layout=1, spread=3.0, area=6.0)
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=16 if key16 else 32)
    views = small_views(skm_mod, 7, seed=5 + key16)
    renderer.render_views(views)
    st = renderer.batch_stats()
    assert st["views"] == 7 and st["graph"]
    got = [batched(renderer, k) for k in range(7)]
    for k, view in enumerate(views):
        assert_same(got[k], single(renderer, view, key16), f"view {k} {view[3]}x{view[4]}")
    assert sum(g["count"] for g in got) == st["visible"] and sum(g["pairs"] for g in got) == st["pairs"]


def test_batch_view_matches_oracle(renderer, skm_mod, oracle):
    say("case: batch of 3 views of a 60k SH3 scene against the oracle")
    cloud = skm_mod.make_cloud(60_000, 3, seed=43, layout=0, spread=1.5)
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=32)
    views = small_views(skm_mod, 3, seed=9, resolutions=((320, 200), (200, 320), (256, 144)))
    renderer.render_views(views)
    for k, (v, p, eye, w, h) in enumerate(views):
        b = batched(renderer, k)
        o = oracle.frame(cloud, v, p, eye, w, h)
        assert b["count"] == o["count"] and np.array_equal(b["keys"], o["keys"])
        assert np.array_equal(b["projected"], o["projected"]) and np.array_equal(b["order"], o["order"])
        assert b["pairs"] == o["pairs"] and np.array_equal(b["pair_values"], o["pair_values"])
        assert np.array_equal(b["ranges"], o["ranges"])
        err = np.abs(b["image"] - o["image"]).max()
        say(f"view {k}: visible {o['count']} pairs {o['pairs']} max |GPU - fp64| {err:.3e}")
        assert err <= 1 / 255


def test_graph_replays_and_eager_launches_agree(renderer, skm_mod):
    say("case: graph replays interleaved with single draws, and the eager batch")
    cloud = skm_mod.make_cloud(150_000, 1, seed=44, layout=1, spread=3.0, area=6.0)
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=32)
    views = small_views(skm_mod, 5, seed=11)
    renderer.render_views(views, graph=False)
    ref = [renderer.read_view_frame(k).copy() for k in range(5)]
    for rep in range(3):
        renderer.render_views(views, graph=True)
        assert renderer.batch_stats()["graph"]
        for k in range(5):
            assert np.array_equal(renderer.read_view_frame(k).view(np.uint32), ref[k].view(np.uint32)), (rep, k)
        v, p, eye, w, h = views[rep]
        renderer.set_extent(w, h)
        renderer.render(v, p, eye)  # eager work between the replays


def test_batch_ranges_and_reordered_views(renderer, skm_mod):
    say("case: ranges shared by the views, views in another order")
    cloud = skm_mod.make_cloud(120_000, 0, seed=45, layout=0, spread=2.0)
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=16)
    ranges = [(1000, 30_000), (50_000, 2), (60_000, 40_000)]
    views = small_views(skm_mod, 4, seed=13)
    renderer.render_views(views, ranges=ranges)
    first = [batched(renderer, k) for k in range(4)]
    renderer.render_views(views[::-1], ranges=ranges)
    for k in range(4):
        assert_same(batched(renderer, 3 - k), first[k], f"reordered view {k}")
    v, p, eye, w, h = views[2]
    renderer.set_extent(w, h)
    renderer.render(v, p, eye, ranges=ranges)
    count, keys, proj = renderer.read_visible()
    assert count == first[2]["count"] and np.array_equal(proj, first[2]["projected"])


@pytest.mark.parametrize("count,res,radix,view_passes,sort_kernel", [
    (7, None, 8, 0, 0),              # global tile ids: the view-local ids would take as many passes
    (24, ((1280, 720),), 8, 1, 0),   # 86,400 tiles (17 bits: 3 passes) vs 3,600 per view (12 bits: 2 passes)
    (24, ((1280, 720),), 8, 1, 1),   # the same with the look-back passes (the pass on the view is reduce-then-scan)
    (24, ((1280, 720),), 6, 1, 0),   # 3 passes (6, 6, 5 bits) vs 2 of 6
    (300, ((256, 256),), 8, 2, 0),   # 76,800 tiles (3 passes) vs 256 per view (1 pass); 300 views: 2 view passes
])
def test_view_local_tile_keys(renderer, skm_mod, count, res, radix, view_passes, sort_kernel):
    say(f"case: {count} views, resolutions {res}, radix {radix}, sort kernel {sort_kernel}: {view_passes} passes "
        "on the view expected")
    cloud = skm_mod.make_cloud(80_000, 1, seed=47, layout=1, spread=3.0, area=6.0)
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=32)
    views = small_views(skm_mod, count, seed=17, **({"resolutions": res} if res else {}))
    renderer.render_views(views, radix_bits=radix, sort_kernel=sort_kernel)
    st = renderer.batch_stats()
    say(f"batch: {st}")
    assert st["view_sort_passes"] == view_passes
    got = {k: batched(renderer, k) for k in sorted({0, 1, count // 2, count - 1})}
    renderer.set_tuning(radix_bits=radix)
    try:
        for k, b in got.items():
            assert_same(b, single(renderer, views[k], False), f"view {k}")
    finally:
        renderer.set_tuning()
    counts = [renderer.read_view_visible(k, capacity=1)[0] for k in range(count)]
    assert sum(counts) == st["visible"]


def test_eager_batch_after_a_light_one(renderer, skm_mod):
    """An eager batch (no graph: the serving loop's launches) sizes its grids from the latest launch,
    scaled by lanes and pixels; a batch with far more survivors and pairs per pixel than that loops its
    blocks past the estimate and must still equal single draws."""
    say("case: eager batch of 2 small far views, then 6 close views at 1280x720")
    cloud = skm_mod.make_cloud(250_000, 1, seed=48, layout=1, spread=3.0, area=6.0)
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=32)
    proj_small = skm_mod.perspective(np.pi / 3, 1.0, 0.2, 100.0)
    far = [(skm_mod.look_at(np.array([0.0, 2.0, z], np.float32), (0.0, 2.0, z + 40.0)), proj_small,
            np.array([0.0, 2.0, z], np.float32), 256, 256) for z in (20.0, 25.0)]
    renderer.render_views(far, graph=False)
    light = renderer.batch_stats()
    views = small_views(skm_mod, 6, seed=19, resolutions=((1280, 720),))
    renderer.render_views(views, graph=False)
    st = renderer.batch_stats()
    say(f"light batch: {light}; heavy batch: {st}")
    assert not st["graph"] and st["visible"] > 10 * max(1, light["visible"]) and st["pairs"] > 10 * max(1, light["pairs"])
    got = {k: batched(renderer, k) for k in (0, 3, 5)}
    for k, b in got.items():
        assert_same(b, single(renderer, views[k], False), f"view {k}")


def test_64_view_batch_of_the_3m_scene(renderer, skm_mod):
    say("case: 64 views (720p/1080p/1440p) of the 3M SH3 scene, one graph launch")
    cloud = skm_mod.scene_c1()
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=32)
    views = skm_mod.views_c2(64)
    renderer.render_views(views)
    st = renderer.batch_stats()
    say(f"batch: {st}")
    counts = [renderer.read_view_visible(k, capacity=1)[0] for k in range(64)]
    assert sum(counts) == st["visible"]
    for k in (0, 17, 42, 63):
        b = batched(renderer, k)
        s = single(renderer, views[k], False)
        assert_same(b, s, f"view {k}")
