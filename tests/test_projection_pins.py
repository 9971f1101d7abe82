"""Pins of the oracle's projection (projectCovariance, ellipseAxes, the FOV clamp, the radius and
the footprint and sub-pixel culls; SplatProjection.metalh:6-37,78-124) on anisotropic, off-axis
splats under rotated, non-square cameras, against an FP64 restatement built from matrix products
and numpy.linalg.eigh (tests/projection_cases.py), plus rigid-motion invariance.

What each test fails on (each was checked against a deliberately broken copy of the oracle):
  * W taken transposed in T = J W: every rotated camera (fp64 axes; rigid motion);
  * the sign of J's third column (j02, j12) flipped: every off-axis splat (fp64 axes);
  * the eigenvector taken from the larger eigenvalue's complement, (b, d - lambda1): every
    anisotropic splat (fp64 axes);
  * the y clamp built from the x half-FOV tangent: splats past 1.3 tan(half FOV) in y under a
    non-square camera (fp64 axes of the clamped splats).
"""
import math

import numpy as np
import pytest

import cases
import projection_cases as pc


def oracle_project(oracle, cam, splats):
    """orc_project on each packed splat: (visible flags, records)."""
    L = oracle.lib()
    out = np.zeros((len(splats), 8), np.uint32)
    vis = np.zeros(len(splats), bool)
    rec = np.zeros(8, np.uint32)
    for i, s in enumerate(splats):
        s = np.ascontiguousarray(s, np.uint32)
        vis[i] = L.orc_project(cam, oracle.up(s), None, 0, i, oracle.up(rec)) == 1
        out[i] = rec
    return vis, out


def decided(ref, min_px):
    """True / False when the FP64 projection decides visibility with room to spare (1% of the
    margins), None when the splat sits too close to a cull boundary for fp32 to be held to it."""
    if ref is None:
        return False
    w = ref["clip"][3]
    fp = ref["footprint"]
    if np.any(fp < -0.01 * w):
        return False
    sub = ref["subpixel"] - min_px
    if np.all(fp > 0.01 * w) and sub > 0.01 * max(min_px, 1e-3):
        return True
    if np.all(fp > 0.01 * w) and sub < -0.01 * max(min_px, 1e-3):
        return False
    return None


def check_scene(oracle, name, w, h, view, proj, eye, cloud, min_px):
    splats = oracle.pack_splats(cloud)
    cam = oracle.make_camera(view, proj, w, h, eye, 0, min_px)
    vis, recs = oracle_project(oracle, cam, splats)
    stats = {"compared": 0, "clamped_x": 0, "clamped_y": 0, "off_axis": 0, "aniso": 0}
    for i in range(len(splats)):
        ref = pc.fp64_project(splats[i], view, proj, w, h)
        want = decided(ref, min_px)
        if want is None:
            continue
        assert vis[i] == want, (name, i, ref and ref["footprint"], ref and ref["subpixel"])
        if not want:
            continue
        rec = recs[i]
        clip = ref["clip"]
        cx, cy = recs[i, 0].view(np.float32), recs[i, 1].view(np.float32)
        assert abs(cx - clip[0] / clip[3]) <= 1e-5 * (1 + abs(clip[0] / clip[3])), (name, i)
        assert abs(cy - clip[1] / clip[3]) <= 1e-5 * (1 + abs(clip[1] / clip[3])), (name, i)
        r = float(rec[6].view(np.float32))
        assert abs(r - ref["radius"]) <= 2e-5 * ref["radius"], (name, i, r, ref["radius"])
        a, b, d = ref["cov"]
        # Axes only where the eigenvector is well conditioned (gap above the 0.1 floor and 2% of
        # the major eigenvalue); the record holds halves (2^-11 relative) of fp32 arithmetic.
        if ref["gap"] > max(0.2, 0.02 * (a + d)):
            err = pc.axes_error(rec, ref)
            assert err <= 2.5e-3, (name, i, err, pc.record_axes(rec), ref["a1"], ref["a2"])
            stats["compared"] += 1
            stats["clamped_x"] += int(ref["clamped"][0])
            stats["clamped_y"] += int(ref["clamped"][1])
            v = ref["view"]
            stats["off_axis"] += int(abs(v[0] / v[2]) > 0.2 and abs(v[1] / v[2]) > 0.2)
            stats["aniso"] += int(ref["gap"] > 0.5 * (a + d))
    return stats


@pytest.mark.parametrize("seed", [11, 12])
def test_projection_matches_fp64_linear_algebra(oracle, seed):
    """Survivors, centres, radii and axes of the oracle against the FP64 matrix-product projection,
    per camera shape (16:9, the phone's portrait aspect, 2:1), each under a random rotation."""
    for name, w, h, view, proj, eye, cloud in pc.scene(seed):
        for min_px in (0.0, 1.0):
            st = check_scene(oracle, name, w, h, view, proj, eye, cloud, min_px)
            # The cases the four slips need: enough well-conditioned anisotropic, off-axis splats, and
            # survivors whose Jacobian uses the clamped coordinate on each axis.
            assert st["compared"] >= 150 and st["off_axis"] >= 40 and st["aniso"] >= 40, (name, st)
            assert st["clamped_x"] >= 5 and st["clamped_y"] >= 5, (name, st)


def test_fov_clamp_binds_per_axis(oracle):
    """A splat past 1.3 tan(half FOV) on y only, under a 2:1 camera whose x tangent is twice its y
    tangent: the projected covariance equals the FP64 one with y clamped at 1.3 tan_y, and differs
    from the unclamped one (so the clamp is what the comparison sees)."""
    w, h = 1000, 500
    proj = cases.perspective(math.radians(50.0), w / h, 0.1, 100.0)
    tan_y = 1.0 / float(proj[5])
    view = cases.identity()
    z = -4.0
    y = 1.6 * tan_y * -z
    cloud = cases.iso_cloud([(0.3, y, z)])
    r = pc.quat_to_matrix(np.array([0.9, 0.2, -0.3, 0.1]))
    c = r @ np.diag([1.2, 0.5, 0.3]) ** 2 @ r.T
    cloud["covariances"][0] = [c[0, 0], c[0, 1], c[0, 2], c[1, 1], c[1, 2], c[2, 2]]
    splats = oracle.pack_splats(cloud)
    cam = oracle.make_camera(view, proj, w, h, (0, 0, 0), 0, 0.0)
    vis, recs = oracle_project(oracle, cam, splats)
    assert vis[0]
    ref = pc.fp64_project(splats[0], view, proj, w, h)
    assert ref["clamped"] == (False, True)
    assert pc.axes_error(recs[0], ref) <= 2.5e-3
    # The same splat with the clamp taken from the x tangent: a different covariance.
    wide = dict(ref)
    P = pc.row_major(proj)
    v = ref["view"]
    focal = np.array([w * P[0, 0] / 2.0, h * P[1, 1] / 2.0])
    lim = 1.3 / P[0, 0]
    ty = np.clip(v[1] / v[2], -lim, lim) * v[2]
    J = np.array([[focal[0] / v[2], 0, -focal[0] * v[0] / v[2] ** 2], [0, focal[1] / v[2], -focal[1] * ty / v[2] ** 2]])
    xx, xy = pc.unpack_halves(splats[0][4])
    xz, yy = pc.unpack_halves(splats[0][5])
    yz, zz = pc.unpack_halves(splats[0][6])
    S = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])
    cov = J @ S @ J.T + 0.3 * np.eye(2)
    lam, vec = np.linalg.eigh(cov)
    wide["a1"] = vec[:, 1] * math.sqrt(lam[1])
    wide["a2"] = np.array([vec[1, 1], -vec[0, 1]]) * math.sqrt(lam[0])
    assert pc.axes_error(recs[0], wide) > 2e-2


def test_rigid_motion_invariance(oracle):
    """Rotating and moving the world and the camera together leaves view space unchanged: the same
    survivors, depth keys within fp32 rounding, and records within the half rounding of the moved
    covariances (centres, radii, axes up to their common sign)."""
    rng = np.random.default_rng(77)
    for name, w, h, view, proj, eye, cloud in pc.scene(31, n=500):
        q = pc.quat_to_matrix(rng.normal(size=4))
        t = rng.uniform(-5, 5, 3)
        M = pc.rigid(q, t)
        moved = dict(cloud)
        p = cloud["positions"].astype(np.float64)
        moved["positions"] = (p @ q.T + t).astype(np.float32)
        cov = np.zeros_like(cloud["covariances"])
        for i, c6 in enumerate(cloud["covariances"].astype(np.float64)):
            S = np.array([[c6[0], c6[1], c6[2]], [c6[1], c6[3], c6[4]], [c6[2], c6[4], c6[5]]])
            Sm = q @ S @ q.T
            cov[i] = [Sm[0, 0], Sm[0, 1], Sm[0, 2], Sm[1, 1], Sm[1, 2], Sm[2, 2]]
        moved["covariances"] = cov
        view2 = pc.col_major(pc.row_major(view) @ np.linalg.inv(M))
        eye2 = (q @ eye + t).astype(np.float32)
        results = []
        for cl, vw, ey in ((cloud, view, eye), (moved, view2, eye2)):
            cam = oracle.make_camera(vw, proj, w, h, ey, 0, 1.0)
            splats = oracle.pack_splats(cl)
            vis, recs = oracle_project(oracle, cam, splats)
            depth = np.array([-(pc.row_major(vw) @ np.append(s[:3].view(np.float32).astype(np.float64), 1.0))[2]
                              for s in splats])
            refs = [pc.fp64_project(s, vw, proj, w, h) for s in splats]
            results.append((vis, recs, depth, refs))
        (v0, r0, d0, f0), (v1, r1, d1, f1) = results
        compared = 0
        for i in range(len(v0)):
            a, b = decided(f0[i], 1.0), decided(f1[i], 1.0)
            if a is None or b is None:
                continue
            assert a == b == v0[i] == v1[i], (name, i)
            assert abs(d0[i] - d1[i]) <= 1e-5 * abs(d0[i])
            if not v0[i]:
                continue
            for k in (0, 1):
                assert abs(r0[i, k].view(np.float32) - r1[i, k].view(np.float32)) <= 2e-5
            ra, rb = float(r0[i, 6].view(np.float32)), float(r1[i, 6].view(np.float32))
            assert abs(ra - rb) <= 1e-3 * ra
            gap_ok = f0[i]["gap"] > max(0.2, 0.05 * (f0[i]["cov"][0] + f0[i]["cov"][2]))
            if gap_ok:
                assert pc.axes_error(r1[i], f0[i]) <= 6e-3, (name, i)
                compared += 1
        assert compared >= 100, (name, compared)


def test_depth_keys_match_fp64_view_depth(oracle):
    """The cull's 32-bit key is the float bits of -(view p).z (SplatVisibility.metal:90-92): within
    fp32 rounding of the FP64 depth under rotated cameras, and ordered like it."""
    for name, w, h, view, proj, eye, cloud in pc.scene(41, n=300):
        cam = oracle.make_camera(view, proj, w, h, eye, 0, 0.0)
        count, keys, recs = oracle.cull(cam, oracle.pack_splats(cloud), None, 0)
        idx = recs[:count, 7]
        V = pc.row_major(view)
        depth = np.array([-(V @ np.append(cloud["positions"][i].astype(np.float64), 1.0))[2] for i in idx])
        got = keys.view(np.float32).astype(np.float64)
        assert np.all(np.abs(got - depth) <= 4e-7 * np.abs(depth) + 1e-6), name
        assert count > 100
