"""The renderer's frame contract on one MI355X, through the C ABI, against the CPU oracle:
anisotropic off-axis splats under rotated non-square cameras, linear output, dropped frames when
tile pairs overflow (reported until skm_finish, then redrawn at the grown capacity), the 65,536
ranges a frame may draw (MetalVisibilityPass.h:72), and tile uploads into a slab while frames
drawing its other ranges are in flight (SplatRenderer.h:94-96)."""
import sys

import numpy as np
import pytest

import cases
import projection_cases as pc
from test_gpu_parity import assert_parity, gpu_frame

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("key16", [False, True])
def test_anisotropic_off_axis_scenes_match_oracle(renderer, skm_mod, oracle, key16):
    """The projection pins' scenes (landscape, portrait and 2:1 cameras under random rotations, splats
    past the 1.3 tan(half FOV) clamp on each axis): bit-exact survivors, keys, records, order and
    tiles, pixels within 1/255 of FP64."""
    for seed in (11, 12):
        for name, w, h, view, proj, eye, cloud in pc.scene(seed):
            say(f"case: anisotropic {name} {w}x{h} seed {seed} key16={key16}")
            for min_px in (0.0, 1.0):
                g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, w, h, key16=key16, min_px=min_px)
                o = oracle.frame(cloud, view, proj, eye, w, h, key16=key16, min_pixel_radius=min_px)
                assert_parity(g, o, label=f"{name} seed {seed} min_px {min_px}")


def test_linear_output_matches_oracle(renderer, skm_mod, oracle):
    """outputLinear (SplatProjection.metalh:131, and the reveal glow's :184-186): colours raised to
    2.2 after SH, bit-exact records, pixels within 1/255."""
    say("case: linear output, SH2, reveal band")
    cloud = skm_mod.make_cloud(40_000, 2, seed=17, layout=0, spread=1.5, log_scale_mean=-3.0)
    view, proj, eye = skm_mod.orbit_camera(9, 480, 320, radius=4.0, height_m=0.6)
    reveal = (0.3, 0.5, 0.6, 0.7, 0.9, 0.3, 0.2, 0.0)
    renderer.set_linear_blending(True)
    try:
        for rv in (None, reveal):
            g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 480, 320, reveal=rv)
            o = oracle.frame(cloud, view, proj, eye, 480, 320, reveal=rv, output_linear=1)
            assert_parity(g, o, label=f"linear reveal={rv is not None}")
            plain = oracle.frame(cloud, view, proj, eye, 480, 320, reveal=rv, image=False)
            assert not np.array_equal(plain["projected"], o["projected"])  # the exponent changed colours
    finally:
        renderer.set_linear_blending(False)


def overflow_scene(skm_mod):
    """7.8M tile pairs at 1280x720: past the 2^22 pairs a renderer with pair_factor 1 starts with."""
    cloud = skm_mod.make_cloud(400_000, 0, seed=91, layout=0, spread=1.5, log_scale_mean=-3.0)
    near = skm_mod.orbit_camera(0, 1280, 720, radius=4.0, height_m=0.5)
    far = skm_mod.orbit_camera(0, 1280, 720, radius=60.0, height_m=0.5)
    return cloud, near, far


def test_pair_overflow_drops_the_frame_then_regrows(skm_mod, oracle):
    """A frame whose pairs exceed the capacity is dropped: skm_finish and skm_read_frame report
    SKM_ERR_CAPACITY, the totals count it as dropped and not drawn, and the same frame drawn again
    (the capacity has grown) matches the oracle bit for bit."""
    say("case: tile-pair overflow at pair_factor 1, then the regrown frame")
    cloud, near, _ = overflow_scene(skm_mod)
    r = skm_mod.Renderer(pair_factor=1, frames_in_flight=1)
    try:
        r.upload_world(cloud)
        r.set_extent(1280, 720)
        r.apply_policy(sort_depth=32)
        r.frame_totals(reset=True)
        frame = r.make_frame(*near)
        r.draw(frame)
        with pytest.raises(skm_mod.SkmError) as e:
            r.finish()
        assert e.value.status == skm_mod.SKM_ERR_CAPACITY
        with pytest.raises(skm_mod.SkmError) as e:
            r.read_frame()
        assert e.value.status == skm_mod.SKM_ERR_CAPACITY
        tot = r.frame_totals(reset=True)
        assert tot["dropped"] == 1 and tot["frames"] == 0, tot
        r.finish()  # the report was taken
        r.draw(frame)
        r.finish()
        tot = r.frame_totals(reset=True)
        assert tot["dropped"] == 0 and tot["frames"] == 1
        count, keys, projected = r.read_visible()
        pairs, pt, pv, ranges = r.read_tiles()
        o = oracle.frame(cloud, *near, 1280, 720)
        assert pairs == o["pairs"] > (1 << 22)
        g = {"count": count, "keys": keys, "projected": projected, "order": r.read_order(count), "pairs": pairs,
             "pair_tiles": pt, "pair_values": pv, "ranges": ranges, "image": r.read_frame()}
        assert_parity(g, o, label="regrown frame")
    finally:
        r.close()


def test_a_drop_inside_the_pipeline_is_reported_at_finish(skm_mod):
    """With two frames in flight, a dropped frame followed by one that fits: skm_finish still reports
    the drop (not only the newest frame's state), the newest frame's image is readable, and the
    totals hold one drawn and one dropped frame."""
    say("case: overflowing frame, then a small frame, two in flight")
    cloud, near, far = overflow_scene(skm_mod)
    r = skm_mod.Renderer(pair_factor=1, frames_in_flight=2)
    try:
        r.upload_world(cloud)
        r.set_extent(1280, 720)
        r.apply_policy(sort_depth=32)
        r.frame_totals(reset=True)
        r.draw(r.make_frame(*near))
        r.draw(r.make_frame(*far))
        with pytest.raises(skm_mod.SkmError) as e:
            r.finish()
        assert e.value.status == skm_mod.SKM_ERR_CAPACITY
        img = r.read_frame()  # the newest frame fit
        assert img.shape == (720, 1280, 4)
        tot = r.frame_totals(reset=True)
        assert tot["dropped"] == 1 and tot["frames"] == 1, tot
        r.finish()
    finally:
        r.close()


def test_65536_ranges_match_oracle(renderer, skm_mod, oracle):
    """A frame of the most ranges the visibility pass takes (kMaxRanges, MetalVisibilityPass.h:72):
    uneven runs of 0 to 9 splats in shuffled order, so findRange's binary search
    (SplatVisibility.metal:30-37) maps every lane; one more range is rejected before any work."""
    say("case: 65536 ranges over a 300k-splat world")
    n = 300_000
    cloud = skm_mod.make_cloud(n, 1, seed=93, layout=0, spread=1.5, log_scale_mean=-3.2)
    rng = np.random.default_rng(94)
    sizes = rng.integers(0, 10, 65536)
    sizes[rng.integers(0, 65536, 200)] = 0
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    assert starts[-1] + sizes[-1] <= n
    perm = rng.permutation(65536)
    ranges = [(int(starts[k]), int(sizes[k])) for k in perm]
    view, proj, eye = skm_mod.orbit_camera(3, 640, 360, radius=3.5, height_m=0.5)
    g = gpu_frame(renderer, skm_mod, cloud, view, proj, eye, 640, 360, ranges=ranges)
    o = oracle.frame(cloud, view, proj, eye, 640, 360, ranges=ranges)
    assert o["count"] > 10_000
    assert_parity(g, o, label="65536 ranges")
    with pytest.raises(skm_mod.SkmError) as e:
        renderer.draw(renderer.make_frame(view, proj, eye, ranges=ranges + [(0, 1)]))
    assert e.value.status == skm_mod.SKM_ERR_INVALID


def test_tile_upload_while_frames_draw_other_ranges(skm_mod, oracle):
    """createSlab / uploadTile (MetalWorld.mm:59-100) in a tiled world: frames drawing tiles A and B
    are in flight (three slots) while tile C is uploaded into its own records; they give the oracle's
    result for A + B, and the next frame, drawing A + B + C, the oracle's for all three."""
    say("case: slab, tile upload during three frames in flight")
    cloud = skm_mod.make_cloud(90_000, 1, seed=95, layout=0, spread=1.2, log_scale_mean=-3.0)
    part = lambda lo, hi: {k: (v[lo:hi] if isinstance(v, np.ndarray) else v) for k, v in cloud.items()}
    view, proj, eye = skm_mod.orbit_camera(5, 400, 300, radius=3.0, height_m=0.4)
    r = skm_mod.Renderer(frames_in_flight=3)
    try:
        r.create_slab(90_000, 1)
        r.upload_tile(0, part(0, 30_000))
        r.upload_tile(30_000, part(30_000, 60_000))
        r.set_extent(400, 300)
        r.apply_policy(sort_depth=32)
        r.frame_totals(reset=True)
        ab = [(0, 30_000), (30_000, 30_000)]
        for _ in range(3):
            r.draw(r.make_frame(view, proj, eye, ranges=ab))
        r.upload_tile(60_000, part(60_000, 90_000))  # the frames above do not read these records
        r.finish()
        tot = r.frame_totals(reset=True)
        o_ab = oracle.frame(cloud, view, proj, eye, 400, 300, ranges=ab)
        assert tot["frames"] == 3 and tot["visible"] == 3 * o_ab["count"] and tot["pairs"] == 3 * o_ab["pairs"]
        assert np.abs(r.read_frame() - o_ab["image"]).max() <= 1 / 255
        g = gpu_frame(r, skm_mod, cloud, view, proj, eye, 400, 300, upload=False, ranges=ab + [(60_000, 30_000)])
        o = oracle.frame(cloud, view, proj, eye, 400, 300, ranges=ab + [(60_000, 30_000)])
        assert_parity(g, o, label="A + B + C")
    finally:
        r.close()


def test_raster_test_counting_is_opt_in(skm_mod):
    """skm_count_raster_tests: off by default (frame totals report 0 coverage tests); on, the count is
    the same for the same frame drawn twice, positive, at most 256 (P + 3 T) (each of a 16 x 16 tile's
    four 64-pixel waves walks at most its tile's P_t records, padded to groups of 4), and the pixels are
    those of the frames drawn without counting."""
    say("case: 100k SH1 720p, raster tests off then on")
    cloud = skm_mod.make_cloud(100_000, 1, seed=23, layout=1, spread=3.0, area=6.0)
    r = skm_mod.Renderer(0)
    try:
        r.upload_world(cloud)
        r.set_extent(1280, 720)
        cam = skm_mod.orbit_camera(4, 1280, 720, radius=5.0, height_m=1.0)
        r.render(*cam)
        img_off = r.read_frame()
        off = r.frame_totals(reset=True)
        assert off["frames"] == 1 and off["raster_tests"] == 0
        r.count_raster_tests(True)
        counts = []
        for _ in range(2):
            r.render(*cam)
            assert np.array_equal(r.read_frame(), img_off)
            t = r.frame_totals(reset=True)
            counts.append(t["raster_tests"])
        tiles = ((1280 + 15) // 16) * ((720 + 15) // 16)
        say(f"tests {counts}, pairs {t['pairs']}")
        assert counts[0] == counts[1] and 0 < counts[0] <= 256 * (t["pairs"] + 3 * tiles)
        r.count_raster_tests(False)
        r.render(*cam)
        assert r.frame_totals(reset=True)["raster_tests"] == 0
    finally:
        r.close()
