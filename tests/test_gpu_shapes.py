"""Raster tile sizes and block shapes, both radix pass variants, and the batch autotuner on one
MI355X. Tile pairs and ranges at each tile size match the oracle's binning at that size; pixels are
bit-identical across every shape and variant (a pixel's composite does not depend on tiles or block
shape) and within 1/255 of the oracle's FP64 image."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 1), (16, 2), (32, 2), (8, 1)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scene(skm_mod):
    cloud = skm_mod.make_cloud(150_000, 2, seed=81, layout=1, spread=3.0, area=6.0)
    view, proj, eye = skm_mod.orbit_camera(7, 1000, 600, radius=5.0)
    return cloud, view, proj, eye, 1000, 600


def test_shapes_and_sort_variants_give_identical_pixels(skm_mod, oracle, scene):
    cloud, view, proj, eye, w, h = scene
    o = oracle.frame(cloud, view, proj, eye, w, h)
    first = None
    for tile, px in SHAPES:
        for sort_kernel in (0, 1):
            radix_bits = 8 - (tile // 8 + sort_kernel) % 3  # every digit width across the shapes
            say(f"case: tile {tile} px {px} sort {sort_kernel} radix {radix_bits}")
            r = skm_mod.Renderer(0, sort_kernel=sort_kernel, tile=tile, raster_px=px)
            r.set_tuning(tile, px, sort_kernel, radix_bits)
            try:
                r.upload_world(cloud)
                r.set_extent(w, h)
                r.apply_policy(sort_depth=32)
                r.render(view, proj, eye)
                count, keys, projected = r.read_visible()
                assert count == o["count"] and np.array_equal(projected, o["projected"])
                assert np.array_equal(r.read_order(count), o["order"])
                pairs, pt, pv, ranges = r.read_tiles()
                bp, bt, bv, br = oracle.bin_tiles(o["projected"], o["order"], w, h, tile=tile)
                assert pairs == bp and np.array_equal(pt, bt) and np.array_equal(pv, bv)
                assert np.array_equal(ranges, br)
                img = r.read_frame()
                err = np.abs(img - o["image"]).max()
                say(f"pairs {pairs} max |GPU - fp64| {err:.3e}")
                assert err <= 1 / 255
                if first is None:
                    first = img
                assert np.array_equal(img.view(np.uint32), first.view(np.uint32))
            finally:
                r.close()


def test_split_heavy_tiles_give_identical_pixels(skm_mod, oracle, scene):
    """With one frame slot the raster draws the heaviest tiles as four quadrant blocks (four lanes per
    pixel, record-parallel tests): every pixel is the unsplit raster's bit for bit, over several views
    (close ones put thousands of pairs in a tile), and within 1/255 of the oracle's FP64 image."""
    cloud, view, proj, eye, w, h = scene
    cams = [(view, proj, eye)] + [skm_mod.orbit_camera(k, w, h, radius=rad, height_m=0.4)
                                  for k, rad in ((3, 2.0), (11, 1.2), (20, 3.0))]
    images = {}
    for fif in (1, 2):
        r = skm_mod.Renderer(0, frames_in_flight=fif)
        try:
            r.upload_world(cloud)
            r.set_extent(w, h)
            r.apply_policy(sort_depth=32)
            images[fif] = []
            for cam in cams:
                r.render(*cam)
                images[fif].append(r.read_frame())
                if fif == 1:
                    pairs, _, _, ranges = r.read_tiles()
                    per = ranges[:, 1].astype(np.int64) - ranges[:, 0]
                    say(f"case: split, {pairs} pairs, heaviest tile {per.max()}, tiles over 1024 pairs {(per > 1024).sum()}")
        finally:
            r.close()
    for k, cam in enumerate(cams):
        assert np.array_equal(images[1][k].view(np.uint32), images[2][k].view(np.uint32)), k
    for k in (0, 2):
        o = oracle.frame(cloud, *cams[k], w, h)
        err = np.abs(images[1][k] - o["image"]).max()
        say(f"view {k}: max |GPU - fp64| {err:.3e}")
        assert err <= 1 / 255


def test_autotuner_times_every_candidate(renderer, skm_mod, scene):
    cloud = scene[0]
    say("case: autotune a 6-view batch")
    renderer.upload_world(cloud)
    renderer.apply_policy(sort_depth=32)
    views = skm_mod.views_c2(6, seed=82, radius=(3.0, 7.0), resolutions=((640, 360), (1280, 720), (333, 217)))
    best, table = renderer.autotune_views(views, repeats=1)
    say(f"best {best}; candidates {table}")
    assert len(table) == 24 and all(ms > 0 for *_, ms in table)
    assert (best["tile_size"], best["raster_px"], best["sort_kernel"], best["radix_bits"]) in \
        [(t, p, k, rb) for t, p, k, rb, _ in table]
    renderer.render_views(views)  # the default configuration
    ref = [renderer.read_view_frame(k).copy() for k in range(6)]
    for tile, px in SHAPES:
        for rb in (6, 7):
            renderer.render_views(views, tile_size=tile, raster_px=px, sort_kernel=rb % 2, radix_bits=rb)
            for k in range(6):
                assert np.array_equal(renderer.read_view_frame(k).view(np.uint32), ref[k].view(np.uint32)), \
                    (tile, px, rb, k)
