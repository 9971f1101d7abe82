"""Parity of the HIP render path with the CPU oracle, through the C ABI, on one MI355X.

Bit-exact: survivor count, depth keys and 32-byte projected records (compaction order), the depth
sort's permutation, the (tile | depth) pairs and the per-tile ranges. Pixels: within 1/255 per
channel of the oracle's FP64 composite (the kernel composites in fp32 and stops a pixel at
transmittance 1/4096; the bound of that stop is 4/4096 for colours up to the packColor cap of 4).
"""
import math
import sys

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
PIXEL_TOL = 1.0 / 255.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_frame(r, skm_mod, cloud, view, proj, eye, w, h, key16=False, min_px=1.0, ranges=None, upload=True,
              background=cases.BACKGROUND, reveal=None, sh_degree=3, lod_error=1.0):
    if upload:
        r.upload_world(cloud)
    r.set_extent(w, h)
    r.apply_policy(sort_depth=16 if key16 else 32, subpixel=min_px, background=background, lod_error=lod_error)
    r.render(view, proj, eye, ranges=ranges, reveal=reveal, sh_degree=sh_degree)
    count, keys, proj_rec = r.read_visible()
    order = r.read_order(count)
    pairs, pt, pv, tranges = r.read_tiles()
    img = r.read_frame()
    return {"count": count, "keys": keys, "projected": proj_rec, "order": order, "pairs": pairs,
            "pair_tiles": pt, "pair_values": pv, "ranges": tranges, "image": img}


def assert_parity(g, o, image=True, label=""):
    assert g["count"] == o["count"], label
    assert np.array_equal(g["keys"], o["keys"]), label
    if not np.array_equal(g["projected"], o["projected"]):
        bad = np.nonzero((g["projected"] != o["projected"]).any(1))[0]
        raise AssertionError(f"{label}: {len(bad)} projected records differ, first {bad[:5]}: "
                             f"{g['projected'][bad[0]]} vs {o['projected'][bad[0]]}")
    assert np.array_equal(g["order"], o["order"]), label
    assert g["pairs"] == o["pairs"], label
    assert np.array_equal(g["pair_tiles"], o["pair_tiles"]), label
    assert np.array_equal(g["pair_values"], o["pair_values"]), label
    assert np.array_equal(g["ranges"], o["ranges"]), label
    if image:
        err = np.abs(g["image"] - o["image"]).max()
        say(f"{label}: visible {o['count']} pairs {o['pairs']} max |fp32 GPU - fp64 oracle| = {err:.3e}")
        assert err <= PIXEL_TOL, label


def test_c0_frame_matches_oracle(renderer, skm_mod, oracle):
    say("case: C0 100k SH0 1280x720, 32-bit keys")
    cloud = skm_mod.scene_c0()
    view, proj, eye = skm_mod.camera_c0()
    g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 1280, 720)
    o = oracle.frame(cloud, view, proj, eye, 1280, 720)
    assert_parity(g, o, label="C0")


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
@pytest.mark.parametrize("key16", [False, True])
def test_random_scenes_match_oracle(renderer, skm_mod, oracle, deg, key16):
    say(f"case: random scene SH{deg} key16={key16}")
    cloud = skm_mod.make_cloud(60_000, deg, seed=100 + deg, layout=1, spread=3.0, area=6.0, log_scale_mean=-3.5)
    for k, frame in enumerate((0, 17, 33)):
        view, proj, eye = skm_mod.orbit_camera(frame, 640, 360, radius=5.0, height_m=1.0)
        g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 640, 360, key16=key16, upload=(k == 0))
        o = oracle.frame(cloud, view, proj, eye, 640, 360, key16=key16)
        assert_parity(g, o, label=f"SH{deg} key16={key16} frame {frame}")


def test_ranges_reveal_and_subpixel(renderer, skm_mod, oracle):
    say("case: three ranges, reveal band, sub-pixel cutoff 2 px, odd extent")
    cloud = skm_mod.make_cloud(50_000, 2, seed=7, layout=0, spread=1.5, log_scale_mean=-3.0)
    view, proj, eye = skm_mod.orbit_camera(4, 333, 217, radius=4.0, height_m=0.5)
    ranges = [(100, 5000), (20000, 7), (30000, 15000)]
    reveal = (0.4, 0.6, 0.5, 0.8, 1.0, 0.2, 0.1, 0.0)
    g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 333, 217, min_px=2.0, ranges=ranges, reveal=reveal)
    o = oracle.frame(cloud, view, proj, eye, 333, 217, min_pixel_radius=2.0, ranges=ranges, reveal=reveal)
    assert_parity(g, o, label="ranges+reveal")


def run_case(r, skm_mod, cloud, min_px=0.0, ranges=None, view=None, proj=None, eye=(0, 0, 0), key16=False):
    view = cases.identity() if view is None else view
    proj = cases.VIS_PROJ if proj is None else proj
    r.upload_world(cloud)
    r.set_extent(cases.VIS_W, cases.VIS_H)
    r.apply_policy(sort_depth=16 if key16 else 32, subpixel=min_px)
    r.render(view, proj, np.asarray(eye, np.float32), ranges=ranges)
    count, keys, proj_rec = r.read_visible()
    return count, keys, proj_rec, r.read_order(count)


def test_reference_visibility_cases(renderer, skm_mod):
    say("case: MetalVisibilityPassTest cases")
    cloud, ranges, expected = cases.case_ranges_front_to_back()
    count, _, pr, order = run_case(renderer, skm_mod, cloud, ranges=ranges)
    assert count == 3 and list(pr[order, 7]) == expected
    cloud, cutoff, kept = cases.case_subpixel_dropped()
    assert run_case(renderer, skm_mod, cloud, min_px=cutoff)[0] == kept
    cloud, rows = cases.case_camera_cutoff()
    for cutoff, kept in rows:
        assert run_case(renderer, skm_mod, cloud, min_px=cutoff)[0] == kept
    cloud, cutoff, kept = cases.case_edge_and_transparent()
    count, _, pr, _ = run_case(renderer, skm_mod, cloud, min_px=cutoff)
    assert sorted(pr[:, 7]) == kept
    cloud, cutoff, kept, radius = cases.case_footprint_planes_radius()
    count, _, pr, _ = run_case(renderer, skm_mod, cloud, min_px=cutoff)
    assert sorted(pr[:, 7]) == kept
    assert abs(pr[list(pr[:, 7]).index(5), 6].view(np.float32) - radius) <= 0.02
    cloud, view, proj, eye, depths, expected = cases.case_depth_keys16()
    count, keys, pr, order = run_case(renderer, skm_mod, cloud, view=view, proj=proj, eye=eye, key16=True)
    assert list(pr[order, 7]) == expected
    sk = np.sort(keys)
    assert sk[1] < sk[2] and sk[3] == sk[4] and sk[-1] < 65535
    cloud, sizes = cases.case_simd_tails()
    renderer.upload_world(cloud)
    for n in sizes:
        renderer.apply_policy(sort_depth=32, subpixel=0.0)
        renderer.render(cases.identity(), cases.VIS_PROJ, np.zeros(3, np.float32), ranges=[(0, n)] if n else [])
        count, _, pr = renderer.read_visible()
        assert count == n - (n + 2) // 3
        ids = pr[:, 7]
        assert list(ids) == sorted(set(ids.tolist())) and all(i % 3 != 0 and i < n for i in ids)


def test_two_million_splats_are_all_kept(renderer, skm_mod):
    """MetalVisibilityPassTest.mm:129-147."""
    say("case: 2000003 splats at depth 2")
    n = 2_000_003
    cloud = cases.iso_cloud(np.tile(np.array([[0, 0, -2]], np.float32), (n, 1)))
    renderer.upload_world(cloud)
    renderer.set_extent(16, 16)  # one tile: every splat covers the whole target
    renderer.apply_policy(sort_depth=32, subpixel=0.0)
    renderer.render(cases.identity(), cases.VIS_PROJ, np.zeros(3, np.float32))
    count, _, pr = renderer.read_visible()
    order = renderer.read_order(count)
    assert count == n
    assert np.array_equal(pr[:, 7], np.arange(n, dtype=np.uint32))
    assert np.array_equal(order, np.arange(n, dtype=np.uint32))  # equal keys keep compaction order


def pixel(r, skm_mod, cloud):
    r.upload_world(cloud)
    r.set_extent(cases.PIX, cases.PIX)
    r.apply_policy(sort_depth=32, subpixel=1.0)
    r.render(cases.identity(), cases.PIX_PROJ, np.zeros(3, np.float32))
    return r.read_frame(skm_mod.PIXELS_BGRA8)[32, 32], r.read_frame()[32, 32]


def test_reference_pixel_cases(renderer, skm_mod):
    """MetalSplatPipelineTest.mm:96-232 through the GPU: background, depth order, coverage."""
    say("case: MetalSplatPipelineTest pixel cases")
    bgra, _ = pixel(renderer, skm_mod, cases.iso_cloud(np.zeros((1, 3), np.float32) + [0, 0, 5]))
    assert abs(int(bgra[2]) - 13) <= 1 and abs(int(bgra[1]) - 13) <= 1 and abs(int(bgra[0]) - 20) <= 1
    assert bgra[3] == 255
    bgra, _ = pixel(renderer, skm_mod, cases.case_overlap_red_green())
    assert bgra[2] > 150 and bgra[1] < 80
    for alpha in (0.5, 0.1, 4.0):
        cloud, cov = cases.case_single_colour(alpha)
        bgra, rgba = pixel(renderer, skm_mod, cloud)
        for c in range(3):
            expect = cov * cloud["colors"][0, c] + (1 - cov) * cases.BACKGROUND[c]
            assert abs(int(bgra[2 - c]) - expect * 255) <= 1.0
            assert abs(rgba[c] - expect) * 255 <= 1.0


@pytest.mark.parametrize("n,mask,seed,bits", [
    (1000003, 0xFFFFFFFF, 1, 32), (70000, 0x7, 2, 32), (70000, 0xFFFF, 6, 16), (17, 0xFFFFFFFF, 3, 32),
    (4096, 0xFFFFFFFF, 4, 32), (0, 0xFFFF, 16, 16), (1, 0xFFFF, 16, 16), (31, 0xFFFF, 16, 16),
    (33, 0xFFFF, 16, 32), (4095, 0xFFFF, 16, 16), (4097, 0xFFFF, 16, 32), (1000003, 0xFFFF, 16, 16),
    (70000, 0x7, 17, 16), (5000000, 0xFFFFFFFF, 5, 32)])
def test_radix_sort_matches_stable_sort(renderer, n, mask, seed, bits):
    """MetalRadixSortTest.mm:81-165: sizes, masks and seeds of the reference's sort tests."""
    say(f"case: argsort n={n} mask={mask:#x} bits={bits}")
    keys = np.random.default_rng(seed).integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) & np.uint32(mask)
    if n > 1 and mask == 0xFFFF:
        keys[0], keys[-1] = 65535, 0
    assert np.array_equal(renderer.argsort(keys, bits), np.argsort(keys, kind="stable"))


@pytest.mark.parametrize("n,bits,seed", [(7_400_000, 13, 21), (300_001, 20, 22), (65_537, 5, 23), (4097, 24, 24),
                                         (3, 1, 25)])
def test_radix_sort_with_one_to_four_passes(renderer, n, bits, seed):
    """Keys below 2^bits take ceil(bits / 8) passes (1, 2, 3); the order lands in the caller's buffer for
    odd and even pass counts. 7.4M 13-bit keys are the tile sort of a 1080p frame's pairs."""
    say(f"case: argsort n={n} bits={bits}")
    keys = np.random.default_rng(seed).integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(renderer.argsort(keys, bits), np.argsort(keys, kind="stable"))


def test_slab_tiles_and_world_info(renderer, skm_mod, oracle):
    """createSlab/uploadTile (MetalWorld.mm:59-100): a slab filled by tiles renders like the world."""
    say("case: slab of 3 tiles")
    cloud = skm_mod.make_cloud(30_000, 1, seed=21, layout=0, spread=1.0, log_scale_mean=-3.0)
    renderer.create_slab(30_000, 1)
    assert renderer.world() == (30_000, 1)
    for lo, hi in ((0, 10_000), (10_000, 25_000), (25_000, 30_000)):
        part = {k: (v[lo:hi] if isinstance(v, np.ndarray) else v) for k, v in cloud.items()}
        renderer.upload_tile(lo, part)
    view, proj, eye = skm_mod.orbit_camera(2, 320, 240, radius=3.0, height_m=0.4)
    g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 320, 240, upload=False)
    o = oracle.frame(cloud, view, proj, eye, 320, 240)
    assert_parity(g, o, label="slab")
    with pytest.raises(skm_mod.SkmError):
        renderer.upload_tile(29_999, {k: (v[:2] if isinstance(v, np.ndarray) else v) for k, v in cloud.items()})


def test_invalid_ranges_are_rejected_before_any_work(renderer, skm_mod):
    say("case: invalid ranges")
    cloud = cases.iso_cloud([(0, 0, -2)] * 4)
    renderer.upload_world(cloud)
    renderer.set_extent(64, 64)
    with pytest.raises(skm_mod.SkmError) as e:
        renderer.draw(renderer.make_frame(cases.identity(), cases.PIX_PROJ, np.zeros(3, np.float32), ranges=[(3, 2)]))
    assert e.value.status == skm_mod.SKM_ERR_INVALID
    assert not renderer.stats()["gpu_failed"]


def test_frames_in_flight_give_the_same_image(renderer, skm_mod):
    say("case: 6 frames in flight, alternating cameras")
    cloud = skm_mod.make_cloud(200_000, 3, seed=31, layout=1, spread=4.0)
    renderer.upload_world(cloud)
    renderer.set_extent(800, 600)
    renderer.apply_policy(sort_depth=32)
    frames = [skm_mod.orbit_camera(k % 2, 800, 600) for k in range(6)]
    for view, proj, eye in frames:
        renderer.draw(renderer.make_frame(view, proj, eye))
    renderer.finish()
    last = renderer.read_frame()
    renderer.render(*frames[1])
    assert np.array_equal(last, renderer.read_frame())
    st = renderer.stats()
    assert st["completed"] == st["submitted"] and st["gpu_ms"] > 0


def test_one_two_and_three_frames_in_flight_agree(skm_mod):
    """Each frame slot has its own stream and scratch: overlapping frames give every frame's counts and
    the last image exactly as frames drawn one at a time."""
    say("case: frames_in_flight 1, 2, 3 over 8 orbit frames")
    cloud = skm_mod.make_cloud(300_000, 3, seed=37, layout=1, spread=4.0)
    frames = [skm_mod.orbit_camera(k * 7, 640, 480) for k in range(8)]
    results = []
    for fif in (1, 2, 3):
        r = skm_mod.Renderer(frames_in_flight=fif)
        try:
            r.upload_world(cloud)
            r.set_extent(640, 480)
            r.apply_policy(sort_depth=32)
            r.frame_totals(reset=True)
            for view, proj, eye in frames:
                r.draw(r.make_frame(view, proj, eye))
            r.finish()
            tot = r.frame_totals(reset=True)
            results.append((tot["frames"], tot["visible"], tot["pairs"], r.read_frame()))
        finally:
            r.close()
    for fif, res in zip((2, 3), results[1:]):
        assert res[:3] == results[0][:3], fif
        assert np.array_equal(res[3], results[0][3]), fif
    assert results[0][0] == 8
    with pytest.raises(skm_mod.SkmError):
        skm_mod.Renderer(frames_in_flight=4)


def test_c1_frame_matches_oracle(renderer, skm_mod, oracle):
    """The 3M-splat SH3 scene at 1920x1080, in the loader's spatial order as the benchmark draws it:
    order bit-exact, pixels within 1/255 of FP64."""
    say("case: C1 3M SH3 1920x1080 frame 0, spatial order")
    cloud = skm_mod.scene_c1()
    order = renderer.reorder_spatially(cloud)
    assert np.array_equal(order, oracle.spatial_order(cloud["positions"][np.argsort(order)],
                                                      cloud["positions"].min(0), cloud["positions"].max(0)))
    view, proj, eye = skm_mod.orbit_camera(0, 1920, 1080)
    g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 1920, 1080)
    o = oracle.frame(cloud, view, proj, eye, 1920, 1080)
    assert_parity(g, o, label="C1")


def test_c1_orbit_frames_with_one_slot_match_oracle(renderer, skm_mod, oracle):
    """Three more orbit frames of the 3M-splat SH3 scene at 1920x1080, drawn with one frame slot, where
    the raster draws the heaviest tiles as four quadrant blocks: records, order, pairs and ranges
    bit-exact against the oracle, pixels within 1/255 of FP64 and bit for bit the image that the
    several-slot renderer (whole-tile raster) draws of the same frame."""
    cloud = skm_mod.scene_c1()
    renderer.reorder_spatially(cloud)
    renderer.upload_world(cloud)
    one = skm_mod.Renderer(frames_in_flight=1)
    try:
        one.upload_world(cloud)
        for k in (15, 30, 45):
            say(f"case: C1 3M SH3 1920x1080 orbit frame {k}, one frame slot")
            view, proj, eye = skm_mod.orbit_camera(k, 1920, 1080)
            g = gpu_frame(one, skm_mod, cloud, view, proj, eye, 1920, 1080, upload=False)
            per = g["ranges"][:, 1].astype(np.int64) - g["ranges"][:, 0]
            say(f"heaviest tile {per.max()} pairs, tiles over 1024 pairs {(per > 1024).sum()}")
            o = oracle.frame(cloud, view, proj, eye, 1920, 1080)
            assert_parity(g, o, label=f"C1 frame {k}")
            whole = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 1920, 1080, upload=False)
            assert np.array_equal(whole["image"].view(np.uint32), g["image"].view(np.uint32)), k
    finally:
        one.close()
