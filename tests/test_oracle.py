"""Pins of the CPU oracle (no GPU): the reference's own known-answer tests restated, the reference's
own CPU code compiled from its sources (oracle/_ref), and closed forms/brute force for the rest.
Each test names what it pins so that a plausible slip (a dropped term, a swapped operand, a wrong
sign or index) fails at least one of them."""
import math

import numpy as np
import pytest

import cases

rng = np.random.default_rng(1234)


# ---------------------------------------------------------------- halves (Half.h, HalfTest.cpp)

def test_half_known_values(oracle):
    h = oracle.lib().orc_to_half
    # HalfTest.cpp:11-20
    assert h(0.0) == 0x0000 and h(-0.0) == 0x8000 and h(1.0) == 0x3C00 and h(-2.0) == 0xC000
    assert h(0.5) == 0x3800 and h(65504.0) == 0x7BFF
    assert h(6.103515625e-5) == 0x0400 and h(5.960464477539063e-8) == 0x0001
    # HalfTest.cpp:22-28 ties to even
    assert h(1.0 + 0.00048828125) == 0x3C00
    assert h(1.0 + 3.0 * 0.00048828125) == 0x3C02
    assert h(float(np.float32(1.0005))) == 0x3C01
    # HalfTest.cpp:30-36 saturation, NaN stays NaN
    assert h(70000.0) == 0x7BFF and h(-1e30) == 0xFBFF and h(math.inf) == 0x7BFF and h(1e-9) == 0
    assert math.isnan(oracle.lib().orc_from_half(h(math.nan)))


def test_half_round_trip_and_ieee_rule_against_numpy(oracle):
    L = oracle.lib()
    # HalfTest.cpp:38-45
    for v in (3.14159, 0.001, 123.456, 1e-6, 2.5e-7, 42.0, -0.333):
        back = L.orc_from_half(L.orc_to_half(v))
        ulp = 6e-8 if abs(v) < 6.1e-5 else abs(v) / 1024
        assert abs(back - v) <= ulp
    # The shader's half(float) rule (RNE, overflow to infinity) against numpy's IEEE conversion,
    # over random bit patterns of every magnitude plus the exact ties.
    bits = rng.integers(0, 2**32, 200_000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    ties = (np.arange(1, 2000, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0**-10)
    for v in np.concatenate([vals, ties, -ties]):
        assert L.orc_half_rne(float(v)) == int(np.float16(v).view(np.uint16)), v
    # from_half is exact for every half
    for hbits in range(0, 0x10000, 7):
        f = np.uint16(hbits).view(np.float16)
        if np.isnan(f):
            continue
        assert L.orc_from_half(hbits) == float(f)


def test_half_matches_reference_code(oracle, ref):
    bits = rng.integers(0, 2**32, 300_000, dtype=np.uint64).astype(np.uint32)
    L = oracle.lib()
    for v in bits.view(np.float32)[:300_000:3]:
        assert L.orc_to_half(float(v)) == ref.ref_to_half(float(v))


# ------------------------------------------------------------- packing (GpuLayout.cpp, GpuLayoutTest.cpp)

def two_splats(sh_degree):
    """GpuLayoutTest.cpp:16-26."""
    k = 3 * ((sh_degree + 1) ** 2 - 1)
    return {"positions": np.array([[0, 1, 2], [3, 4, 5]], np.float32),
            "colors": np.array([[1.0, 0.5, 0.0], [0.0, 0.0, 1.0]], np.float32),
            "alphas": np.array([0.5, 2.0], np.float32),
            "covariances": np.array([[1, 2, 3, 4, 5, 6], [6, 5, 4, 3, 2, 1]], np.float32),
            "sh_degree": sh_degree, "sh": np.full((2, k), 0.25, np.float32) if k else None}


def test_pack_known_answers(oracle):
    L = oracle.lib()
    p = oracle.pack_splats(two_splats(0))
    # GpuLayoutTest.cpp:27-37
    assert p[0, 2].view(np.float32) == 2.0
    assert p[0, 3] & 0xFF == 255 and (p[0, 3] >> 8) & 0xFF == 128 and (p[0, 3] >> 24) & 0xFF == 128
    assert p[0, 4] & 0xFFFF == L.orc_to_half(1.0) and p[0, 6] >> 16 == L.orc_to_half(6.0)
    assert p[0, 7] == 0
    # :39-45 opacity above one kept as float bits, clamped in the byte
    assert p[1, 7].view(np.float32) == 2.0 and (p[1, 3] >> 24) & 0xFF == 255
    # :47-58 strides and SH words
    assert L.orc_sh_stride(1) == 5 and L.orc_sh_stride(3) == 23 and L.orc_sh_stride(0) == 0
    sh = oracle.pack_sh(two_splats(2), 1)
    assert sh.shape == (2, 5)
    assert sh[0, 0] & 0xFFFF == L.orc_to_half(0.25)
    assert sh[0, 4] >> 16 == 0  # the odd half of the last word is padding


def random_cloud(n, sh_degree, seed):
    r = np.random.default_rng(seed)
    k = 3 * ((sh_degree + 1) ** 2 - 1)
    return {"positions": r.normal(0, 3, (n, 3)).astype(np.float32),
            "covariances": r.normal(0, 1, (n, 6)).astype(np.float32) * np.float32(10.0) ** r.integers(-9, 6, (n, 6)).astype(np.float32),
            "colors": r.uniform(-0.2, 1.2, (n, 3)).astype(np.float32),
            "alphas": np.where(r.uniform(size=n) < 0.2, r.uniform(1, 1000, n), r.uniform(-0.1, 1.0, n)).astype(np.float32),
            "sh_degree": sh_degree, "sh": r.normal(0, 2, (n, k)).astype(np.float32) if k else None}


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_pack_matches_reference_code(oracle, ref, degree):
    cloud = random_cloud(20_000, degree, 7 + degree)
    n = len(cloud["alphas"])
    mine = oracle.pack_splats(cloud)
    theirs = np.zeros((n, 8), np.uint32)
    ref.ref_pack_splats(n, oracle.fp(cloud["positions"]), oracle.fp(cloud["covariances"]),
                        oracle.fp(cloud["colors"]), oracle.fp(cloud["alphas"]), oracle.up(theirs))
    assert np.array_equal(mine, theirs)
    for target in range(0, degree + 1):
        if target == 0:
            continue
        mine_sh = oracle.pack_sh(cloud, target)
        theirs_sh = np.zeros((n, ref.ref_sh_stride(target)), np.uint32)
        ref.ref_pack_sh(n, degree, oracle.fp(cloud["positions"]), oracle.fp(cloud["sh"]), target, oracle.up(theirs_sh))
        assert np.array_equal(mine_sh, theirs_sh)


# ---------------------------------------------------------------- transcendentals

def test_log_exp_against_libm(oracle):
    L = oracle.lib()
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0.5, 2.0, 20000),
                         [1.0, 2.0, 0.5, 127.5, 1e-300, 1e300]])
    for x in xs:
        assert abs(L.orc_log_d(float(x)) - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x)))
    for y in rng.uniform(-700, 700, 20000):
        assert abs(L.orc_exp_d(float(y)) / math.exp(y) - 1.0) <= 2e-15


def test_fp32_log_against_libm(oracle):
    """log_f, the fp32 logarithm of the radius and of the raster's log2 peak opacity: within 3 ulp of
    the correctly rounded log over its domain (ratios 1..1e4, half opacities 2^-14..1e3), exact at 1."""
    L = oracle.lib()
    xs = np.concatenate([np.float32(2.0) ** np.arange(-14, 17, dtype=np.float32),
                         rng.uniform(1.0, 1.0e4, 20000).astype(np.float32),
                         rng.uniform(6.1035156e-05, 1.0, 20000).astype(np.float32),
                         np.nextafter(np.float32(1.0), np.float32([0.0, 2.0]))])
    assert L.orc_log_f(1.0) == 0.0
    for x in xs:
        want = np.float32(math.log(float(x)))
        got = np.float32(L.orc_log_f(float(x)))
        ulp = np.spacing(np.abs(want)) if want != 0 else np.float32(1e-45)
        assert abs(float(got) - math.log(float(x))) <= 3 * float(ulp) + 1e-12, x


# ---------------------------------------------------------------- camera (MetalFrameCompute.mm:141-164)

def test_camera_uniform_closed_form(oracle, ref):
    proj = cases.perspective(1.0, 2.0, 0.1, 100.0)
    theirs = np.zeros(16, np.float32)
    ref.ref_perspective(1.0, 2.0, 0.1, 100.0, oracle.fp(theirs))
    assert np.array_equal(proj, theirs)
    cam = oracle.make_camera(cases.identity(), proj, 200, 100, (1, 2, 3), output_linear=1)
    # MetalFrameComputeTest.mm:141-156
    assert cam.screen_size[0] == 200.0 and cam.screen_size[1] == 100.0
    assert cam.focal[0] == np.float32(200.0) * proj[0] / np.float32(2)
    assert cam.focal[1] == np.float32(100.0) * proj[5] / np.float32(2)
    assert cam.tan_half_fov[0] == np.float32(1) / proj[0]
    assert cam.output_linear == 1 and cam.camera_position[0] == 1.0 and cam.camera_position[2] == 3.0
    assert math.isinf(cam.reveal[0])
    # Mat4Test.cpp:37-43: near maps to depth 0, far to 1
    p = proj.reshape(4, 4).T.astype(np.float64)
    n = p @ [0, 0, -0.1, 1]
    f = p @ [0, 0, -100.0, 1]
    assert abs(n[2] / n[3]) < 1e-6 and abs(f[2] / f[3] - 1) < 1e-5


# ---------------------------------------------------------------- visibility cases (MetalVisibilityPassTest.mm)

def run_cull(oracle, cloud, view=None, proj=None, width=cases.VIS_W, height=cases.VIS_H, min_px=0.0,
             ranges=None, key16=False, cam_pos=(0, 0, 0)):
    view = cases.identity() if view is None else view
    proj = cases.VIS_PROJ if proj is None else proj
    deg = cloud["sh_degree"] if cloud.get("sh") is not None else 0
    cam = oracle.make_camera(view, proj, width, height, cam_pos, 0, min_px)
    return oracle.cull(cam, oracle.pack_splats(cloud), oracle.pack_sh(cloud, deg), deg, ranges=ranges, key16=key16)


def test_ranges_cull_and_order_front_to_back(oracle):
    cloud, ranges, expected = cases.case_ranges_front_to_back()
    count, keys, proj = run_cull(oracle, cloud, ranges=ranges)
    assert count == 3
    order = oracle.argsort(keys)
    assert list(proj[order, 7]) == expected


def test_subpixel_and_cutoff_cases(oracle):
    cloud, cutoff, kept = cases.case_subpixel_dropped()
    assert run_cull(oracle, cloud, min_px=cutoff)[0] == kept
    cloud, rows = cases.case_camera_cutoff()
    for cutoff, kept in rows:
        assert run_cull(oracle, cloud, min_px=cutoff)[0] == kept


def test_edge_and_transparent(oracle):
    cloud, cutoff, kept = cases.case_edge_and_transparent()
    count, _, proj = run_cull(oracle, cloud, min_px=cutoff)
    assert sorted(proj[:count, 7]) == kept


def test_footprint_planes_and_radius_cap(oracle):
    cloud, cutoff, kept, radius = cases.case_footprint_planes_radius()
    count, _, proj = run_cull(oracle, cloud, min_px=cutoff)
    assert sorted(proj[:count, 7]) == kept
    r5 = proj[list(proj[:, 7]).index(5), 6].view(np.float32)
    assert abs(r5 - radius) <= 0.02


def test_sixteen_bit_keys(oracle):
    cloud, view, proj, eye, depths, expected = cases.case_depth_keys16()
    count, keys, pr = run_cull(oracle, cloud, view=view, proj=proj, key16=True, cam_pos=eye)
    assert count == len(expected)
    order = oracle.argsort(keys)
    assert list(pr[order, 7]) == expected
    sk = np.sort(keys)
    for i, idx in enumerate(expected):
        assert abs(int(sk[i]) - cases.relative_depth_key(depths[idx], 0.05)) <= 1
    assert sk[1] < sk[2] and sk[3] == sk[4] and sk[-1] < 65535
    cam = oracle.make_camera(view, proj, 1000, 1000, eye)
    for d in (0.06, 1.0, 2.0001, 150.0, 199.0, 1e9):
        assert oracle.lib().orc_depth_key16(cam, d) == cases.relative_depth_key(d, np.float32(0.05))


def test_compaction_of_tails(oracle):
    cloud, sizes = cases.case_simd_tails()
    for n in sizes:
        count, _, proj = run_cull(oracle, cloud, ranges=[(0, n)] if n else [])
        assert count == n - (n + 2) // 3
        ids = proj[:count, 7]
        assert len(set(ids)) == count and all(i % 3 != 0 and i < n for i in ids)
        assert list(ids) == sorted(ids)  # compaction keeps lane order


def test_isotropic_projection_closed_form(oracle):
    """EWA for a splat on the axis: sigma_px^2 = (f sigma / z)^2 + 0.3 (projectCovariance,
    SplatProjection.metalh:6-24), with the eigen-gap floor of ellipseAxes."""
    for z, s in ((-2.0, 0.01), (-5.0, 0.2), (-20.0, 1.0)):
        cloud = cases.iso_cloud([(0, 0, z)], sigmas=s)
        count, _, proj = run_cull(oracle, cloud)
        assert count == 1
        a1 = proj[0, 2:4].view(np.uint32)
        var = (500.0 * np.float32(np.float32(s) ** 2).astype(np.float64) ** 0.5 / -z) ** 2 + 0.3
        h = lambda v: float(np.uint16(v).view(np.float16))
        # A round splat has b == 0 and a == d: ellipseAxes floors the eigen gap at 0.1
        # (SplatProjection.metalh:30), so the axes are sqrt(var + 0.1) along y, sqrt(var - 0.1) along x.
        assert h(a1[0] & 0xFFFF) == 0.0 and abs(h(a1[0] >> 16) - math.sqrt(var + 0.1)) <= 2e-3 * math.sqrt(var)
        assert abs(abs(h(a1[1] & 0xFFFF)) - math.sqrt(var - 0.1)) <= 2e-3 * math.sqrt(var) and h(a1[1] >> 16) == 0.0
        alpha = 1.0
        r = min(2.5, math.sqrt(2 * math.log(alpha * 255 / 2)))
        assert abs(proj[0, 6].view(np.float32) - r) < 1e-6


def test_radius_closed_form(oracle):
    for alpha in (0.01, 0.05, 0.1, 0.5):
        cloud = cases.iso_cloud([(0, 0, -3)], sigmas=0.05, alphas=[alpha])
        count, _, proj = run_cull(oracle, cloud)
        a_packed = round(alpha * 255) / 255
        expect = min(2.5, math.sqrt(2 * math.log(a_packed / (2 / 255))))
        assert abs(proj[0, 6].view(np.float32) - expect) < 2e-6


def sh_unit_response(oracle, k, dirs):
    """Colour response of SH coefficient k (red channel, amplitude 0.25) over directions, read
    through the oracle's projection with a camera looking along each direction."""
    cloud = cases.iso_cloud([(0, 0, 0)], sigmas=0.1, colors=np.array([[1.0, 1.0, 1.0]], np.float32), sh_degree=3)
    cloud["sh"][0, 3 * k] = 0.25
    splat = oracle.pack_splats(cloud)[0]
    shw = oracle.pack_sh(cloud, 3)[0]
    out = np.zeros(8, np.uint32)
    vals = []
    for d in dirs:
        eye = -2.0 * d
        up = (0, 1, 0) if abs(d[1]) < 0.9 else (1, 0, 0)
        view = look_at(eye, (0, 0, 0), up)
        cam = oracle.make_camera(view, cases.VIS_PROJ, 100, 100, eye)
        assert oracle.lib().orc_project(cam, oracle.up(splat), oracle.up(shw), 3, 0, oracle.up(out)) == 1
        vals.append((float(np.uint16(out[4] & 0xFFFF).view(np.float16)) - 1.0) / 0.25)
    return np.array(vals)


def look_at(eye, target, up):
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m.T.reshape(16).astype(np.float32)


def test_sh_basis_is_orthonormal(oracle):
    """The 15 real SH functions of bands 1-3 (shColor, SplatProjection.metalh:50-76) are orthonormal
    over the sphere: a wrong constant, sign pattern or polynomial breaks the Gram matrix."""
    n = 1500
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    theta = math.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], 1)
    resp = np.stack([sh_unit_response(oracle, k, dirs) for k in range(15)])
    gram = resp @ resp.T * (4 * math.pi / n)
    assert np.abs(gram - np.eye(15)).max() < 0.03


# ---------------------------------------------------------------- sort (MetalRadixSortTest.mm:58-139)

@pytest.mark.parametrize("n,mask,seed", [(1000003, 0xFFFFFFFF, 1), (70000, 0x7, 2), (17, 0xFFFFFFFF, 3),
                                         (4096, 0xFFFFFFFF, 4), (0, 0xFFFF, 16), (4097, 0xFFFF, 16)])
def test_argsort_is_stable_sort(oracle, n, mask, seed):
    keys = (np.random.default_rng(seed).integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) & np.uint32(mask))
    assert np.array_equal(oracle.argsort(keys), np.argsort(keys, kind="stable"))


# ---------------------------------------------------------------- tiles and raster

def small_frame(oracle, skm_mod, n=400, w=96, h=80, seed=3, deg=1, key16=False):
    cloud = skm_mod.make_cloud(n, deg, seed=seed, layout=0, spread=0.6, log_scale_mean=-3.0)
    eye = np.array([0.3, 0.2, 3.0], np.float32)
    view = skm_mod.look_at(eye)
    proj = skm_mod.perspective(math.pi / 3, w / h, 0.2, 100.0)
    return cloud, oracle.frame(cloud, view, proj, eye, w, h, key16=key16), w, h


def test_tile_lists_cover_every_accepted_pixel(oracle, skm_mod):
    """Brute force: every (pixel, survivor) the coverage rule accepts has the survivor listed in
    the pixel's tile, and each tile list is in (depth key, compaction index) order."""
    cloud, o, w, h = small_frame(oracle, skm_mod)
    tiles_x = (w + 15) // 16
    rank = np.empty(o["count"], np.int64)
    rank[o["order"]] = np.arange(o["count"])
    for t, (s, e) in enumerate(o["ranges"]):
        vals = o["pair_values"][s:e]
        assert np.all(o["pair_tiles"][s:e] == t)
        assert np.all(np.diff(rank[vals]) > 0)
        listed = set(vals.tolist())
        tx, ty = t % tiles_x, t // tiles_x
        for j in range(o["count"]):
            if j in listed:
                continue
            for py in range(ty * 16, min(ty * 16 + 16, h), 3):
                for px in range(tx * 16, min(tx * 16 + 16, w), 3):
                    assert not oracle.decide(o["projected"][j], w, h, px, py)[0]


@pytest.mark.parametrize("w,h,n,log_scale,lod_alpha,seed", [
    (96, 80, 400, -3.0, False, 3), (320, 180, 3000, -2.5, False, 11), (256, 200, 2000, -2.0, True, 12),
    (160, 90, 5000, -4.0, True, 13)])
def test_binning_misses_no_accepted_pixel(oracle, skm_mod, w, h, n, log_scale, lod_alpha, seed):
    """Every pixel of every record's quad is tested with the coverage rule (in C): each accepted
    (pixel, record) has the record in the pixel's tile list, with the tile rectangles cut to the
    footprint box (quad box and cutoff disc box) rather than the quad's box."""
    cloud = skm_mod.make_cloud(n, 1, seed=seed, layout=0, spread=0.8, log_scale_mean=log_scale)
    if lod_alpha:
        cloud["alphas"][::3] = np.float32(1.5) + np.arange(len(cloud["alphas"][::3]), dtype=np.float32) % 40
    eye = np.array([0.2, 0.3, 3.5], np.float32)
    view = skm_mod.look_at(eye)
    proj = skm_mod.perspective(math.pi / 3, w / h, 0.2, 100.0)
    o = oracle.frame(cloud, view, proj, eye, w, h, image=False)
    missed, accepted = oracle.missed_coverage(o["projected"], o["count"], o["pair_values"], o["ranges"], w, h)
    assert accepted > 0 and missed == 0
    loose = 0
    for j in range(o["count"]):
        p = o["projected"][j]
        f16 = lambda v: float(np.uint16(v).view(np.float16))
        r = float(p[6].view(np.float32))
        ex = r * (abs(f16(p[2] & 0xFFFF)) + abs(f16(p[3] & 0xFFFF)))
        ey = r * (abs(f16(p[2] >> 16)) + abs(f16(p[3] >> 16)))
        cx = (float(p[0].view(np.float32)) * 0.5 + 0.5) * w
        cy = (0.5 - float(p[1].view(np.float32)) * 0.5) * h
        x0, x1 = max(0, math.ceil(cx - ex - 0.5 - 1 / 16)), min(w - 1, math.floor(cx + ex - 0.5 + 1 / 16))
        y0, y1 = max(0, math.ceil(cy - ey - 0.5 - 1 / 16)), min(h - 1, math.floor(cy + ey - 0.5 + 1 / 16))
        if x0 <= x1 and y0 <= y1:
            loose += (x1 // 16 - x0 // 16 + 1) * (y1 // 16 - y0 // 16 + 1)
    assert o["pairs"] <= loose


@pytest.mark.parametrize("region", [8, 16])
@pytest.mark.parametrize("w,h,n,log_scale,lod_alpha,seed", [
    (320, 180, 3000, -2.5, False, 21), (256, 200, 2000, -2.0, True, 22), (160, 90, 5000, -4.0, True, 23)])
def test_raster_region_lists_drop_no_accepted_pixel(oracle, skm_mod, region, w, h, n, log_scale, lod_alpha, seed):
    """The raster's per-region record test (quad slabs widened over the region, footprint box) keeps
    every record that the coverage rule accepts at some pixel of the region."""
    cloud = skm_mod.make_cloud(n, 1, seed=seed, layout=0, spread=0.8, log_scale_mean=log_scale)
    if lod_alpha:
        cloud["alphas"][::3] = np.float32(1.5) + np.arange(len(cloud["alphas"][::3]), dtype=np.float32) % 40
    eye = np.array([0.2, 0.3, 3.5], np.float32)
    o = oracle.frame(cloud, skm_mod.look_at(eye), skm_mod.perspective(math.pi / 3, w / h, 0.2, 100.0), eye, w, h,
                     image=False)
    missed, accepted = oracle.region_misses(o["projected"], o["count"], w, h, region)
    assert accepted > 0 and missed == 0


@pytest.mark.parametrize("w,h,seed", [(3840, 2160, 24), (7680, 4320, 25)])
def test_region_and_tile_lists_at_large_pixel_coordinates(oracle, skm_mod, w, h, seed):
    """The coverage rule is affine in the pixel centre (gu = g00 fx + h01 fy + u0), so its fp32
    rounding grows with the pixel coordinates (about 3 W 2^-24 (|g00| + |g01|) at width W). At 4K and
    8K the region test's margins (the 1/16-pixel pad on its half span, then the slack) and the tile
    rectangles' pad still keep every accepted (pixel, record)."""
    cloud = skm_mod.make_cloud(400, 1, seed=seed, layout=0, spread=0.8, log_scale_mean=-4.5)
    cloud["alphas"][::4] = np.float32(2.0)
    eye = np.array([0.2, 0.3, 3.5], np.float32)
    o = oracle.frame(cloud, skm_mod.look_at(eye), skm_mod.perspective(math.pi / 3, w / h, 0.2, 100.0), eye, w, h,
                     image=False)
    for region in (8, 16):
        missed, accepted = oracle.region_misses(o["projected"], o["count"], w, h, region)
        assert accepted > 0 and missed == 0
    missed, accepted = oracle.missed_coverage(o["projected"], o["count"], o["pair_values"], o["ranges"], w, h)
    assert accepted > 0 and missed == 0


def test_raster_against_numpy_brute_force(oracle, skm_mod):
    """An independent FP64 composite in numpy over all survivors in depth order (exact 2x2 solve of
    the half axes, alpha = min(1, a0 exp(-q/2)), cutoff 1/255) matches the oracle within 1/255."""
    cloud, o, w, h = small_frame(oracle, skm_mod, n=150, w=48, h=40, seed=5)
    f16 = lambda v: float(np.uint16(v).view(np.float16))
    img = np.zeros((h, w, 3))
    T = np.ones((h, w))
    ys, xs = np.mgrid[0:h, 0:w] + 0.5
    for j in o["order"]:
        p = o["projected"][j]
        cx = (float(p[0].view(np.float32)) * 0.5 + 0.5) * w
        cy = (0.5 - float(p[1].view(np.float32)) * 0.5) * h
        A = np.array([[f16(p[2] & 0xFFFF), f16(p[3] & 0xFFFF)], [f16(p[2] >> 16), f16(p[3] >> 16)]])
        r = float(p[6].view(np.float32))
        a0 = f16(p[5] >> 16)
        col = np.clip([f16(p[4] & 0xFFFF), f16(p[4] >> 16), f16(p[5] & 0xFFFF)], 0, 4)
        uv = np.linalg.solve(A, np.stack([xs - cx, cy - ys]).reshape(2, -1)).reshape(2, h, w)
        a = np.minimum(1.0, a0 * np.exp(-0.5 * (uv ** 2).sum(0)))
        take = (np.abs(uv) <= r).all(0) & (a0 * np.exp(-0.5 * (uv ** 2).sum(0)) >= 1 / 255)
        a = np.where(take, a, 0.0)
        img += (T * a)[..., None] * col
        T *= 1 - a
    img += T[..., None] * np.array(cases.BACKGROUND)
    assert np.abs(img - o["image"][..., :3]).max() <= 1 / 255


def test_raster32_termination_error_is_bounded(oracle, skm_mod):
    cloud, o, w, h = small_frame(oracle, skm_mod, n=3000, seed=9, deg=2)
    r32 = oracle.raster32(o["projected"], o["pair_values"], o["ranges"], w, h)
    assert np.abs(r32 - o["image"]).max() <= 1 / 255


def pixel_frame(oracle, cloud, deg=0):
    return oracle.frame(cloud, cases.identity(), cases.PIX_PROJ, (0, 0, 0), cases.PIX, cases.PIX)


def test_pixel_closed_forms(oracle):
    """MetalSplatPipelineTest.mm:96-129,173-232 restated on the oracle's composite."""
    o = pixel_frame(oracle, cases.iso_cloud(np.zeros((0, 3))))
    assert np.allclose(o["image"][32, 32, :3], cases.BACKGROUND, atol=1e-7)
    o = pixel_frame(oracle, cases.case_overlap_red_green())
    px = o["image"][32, 32]
    assert px[0] * 255 > 150 and px[1] * 255 < 80
    for alpha in (0.5, 0.1, 4.0):
        cloud, cov = cases.case_single_colour(alpha)
        px = pixel_frame(oracle, cloud)["image"][32, 32]
        for c in range(3):
            expect = cov * cloud["colors"][0, c] + (1 - cov) * cases.BACKGROUND[c]
            assert abs(px[c] - expect) * 255 <= 1.0
