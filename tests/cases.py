"""Hand-built scenes restated from the reference's own GPU tests, as SplatCloud arrays, with the
results those tests expect. The same cases run through the CPU oracle (test_oracle_*.py) and
through the HIP renderer (test_gpu_*.py).

Cameras: the reference's visibility tests pair a 1000x1000 screen with focal 500 px and tan(half
FOV) 1 (MetalVisibilityPassTest.mm:104-113); here the projection is built with a 90 degree FOV so
that the derived focal and tan(half FOV) are exactly those values.
"""
import math

import numpy as np


def iso_cloud(positions, sigmas=1.0, colors=1.0, alphas=1.0, sh_degree=0):
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    n = len(p)
    s = np.broadcast_to(np.asarray(sigmas, np.float32), (n,))
    cov = np.zeros((n, 6), np.float32)
    var = (s.astype(np.float64) ** 2).astype(np.float32)
    cov[:, 0] = var
    cov[:, 3] = var
    cov[:, 5] = var
    col = np.broadcast_to(np.asarray(colors, np.float32), (n, 3)).copy() if np.ndim(colors) else np.full((n, 3), colors, np.float32)
    a = np.broadcast_to(np.asarray(alphas, np.float32), (n,)).copy()
    cloud = {"positions": p.copy(), "covariances": cov, "colors": col, "alphas": a, "sh_degree": sh_degree,
             "sh": np.zeros((n, 3 * ((sh_degree + 1) ** 2 - 1)), np.float32) if sh_degree else None}
    return cloud


def identity():
    return np.eye(4, dtype=np.float32).reshape(16)


def perspective(fovy, aspect, near, far):
    """Mat4::perspective (Mat4.h:75-86), column-major, computed in float32 like the reference."""
    f32 = np.float32
    ys = f32(1) / f32(math.tan(f32(fovy) * f32(0.5)))
    xs = ys / f32(aspect)
    zs = f32(far) / (f32(near) - f32(far))
    m = np.zeros(16, np.float32)
    m[0], m[5], m[10], m[14], m[11] = xs, ys, zs, f32(near) * zs, -1
    return m


VIS_W = VIS_H = 1000
VIS_PROJ = perspective(math.pi / 2, 1.0, 0.1, 100.0)  # focal 500 px, tan(half FOV) 1


def case_ranges_front_to_back():
    """MetalVisibilityPassTest.mm:205-263: two ranges of six splats; drawn front to back {2, 4, 0}."""
    pos = [(0, 0, -10), (0, 0, 5), (0, 0, -2), (50, 0, -1), (0, 0, -5), (0, 0, -3)]
    return iso_cloud(pos), [(0, 3), (3, 2)], [2, 4, 0]


def case_subpixel_dropped():
    """MetalVisibilityPassTest.mm:265-304: a 1e-8 variance splat never takes a slot (cutoff 0.5)."""
    c = iso_cloud([(0, 0, -2), (0, 0, -2)])
    c["covariances"][1, [0, 3, 5]] = 1.0e-8
    return c, 0.5, 1


def case_camera_cutoff():
    """MetalVisibilityPassTest.mm:307-333: sigma 0.3 px x 2.5 = 0.75 px: kept at 0.5, dropped at 1."""
    c = iso_cloud([(0, 0, -20)])
    c["covariances"][0, [0, 3, 5]] = 1.44e-4
    return c, [(0.5, 1), (1.0, 0)]


def case_edge_and_transparent():
    """MetalVisibilityPassTest.mm:335-372: depth 2; edge-crossing quad kept, opacity 0 culled: {0, 3}."""
    edge = 2.0 / float(VIS_PROJ[0])
    cases = [(1.25 * edge, 0.3, 1.0), (1.25 * edge, 0.05, 1.0), (0.0, 0.1, 0.0), (0.0, 0.1, 1.0)]
    c = iso_cloud([(x, 0, -2) for x, _, _ in cases], sigmas=[s for _, s, _ in cases],
                  alphas=[a for _, _, a in cases])
    return c, 0.0, [0, 3]


def case_footprint_planes_radius():
    """MetalVisibilityPassTest.mm:374-423: kept {0, 4, 5}; splat 5's radius is 512/400 sigma."""
    rows = [(-2.0, 0.0024, 1.0, 1.0), (-2.0, 0.0008, 1.0, 1.0), (-0.05, 0.1, 1.0, 1.0),
            (-150.0, 1.0, 1.0, 1.0), (-2.0, 0.1, 0.0, 2.0), (-2.0, 1.6, 1.0, 1.0)]
    c = iso_cloud([(0, 0, z) for z, _, _, _ in rows], sigmas=[s for _, s, _, _ in rows],
                  colors=np.array([[v, v, v] for _, _, v, _ in rows], np.float32),
                  alphas=[a for _, _, _, a in rows])
    return c, 1.0, [0, 4, 5], 512.0 / 400.0


def case_depth_keys16():
    """MetalVisibilityPassTest.mm:427-493: 16-bit keys, near 0.05, far 200, camera at z = 10."""
    depths = [150, 2.0002, 2.0001, 0.4003, 0.4002, 0.06, 199, 250, -1]
    c = iso_cloud([(0, 0, np.float32(10) - np.float32(d)) for d in depths])
    view = identity()
    view[14] = -10.0
    proj = perspective(1.0, 1.0, 0.05, 200.0)
    return c, view, proj, np.array([0, 0, 10], np.float32), depths, [5, 4, 3, 1, 2, 0, 6]


def relative_depth_key(depth, z_near):
    """The reference test's closed form (MetalVisibilityPassTest.mm:431-441)."""
    bits = lambda v: int(np.float32(v).view(np.uint32))
    return min((bits(depth) - bits(z_near)) >> 11, 65535)


def case_simd_tails():
    """MetalVisibilityPassTest.mm:495-537: every third splat behind the camera."""
    n = 513
    pos = np.zeros((n, 3), np.float32)
    pos[:, 2] = -2
    pos[::3, 2] = 2
    return iso_cloud(pos), [1, 31, 32, 33, 255, 256, 257, 513, 0]


# Pixel cases (MetalSplatPipelineTest.mm): a 64x64 target, perspective(1, 1, 0.1, 100).
PIX = 64
PIX_PROJ = perspective(1.0, 1.0, 0.1, 100.0)
BACKGROUND = (0.05, 0.05, 0.08)


def case_overlap_red_green():
    """MetalSplatPipelineTest.mm:40-49,173-192: green at 2.5 m, red at 1.8 m; red wins at the centre."""
    c = iso_cloud([(0, 0, -2.5), (0, 0, -1.8)], sigmas=math.sqrt(0.05),
                  colors=np.array([[0, 1, 0], [1, 0, 0]], np.float32))
    c["covariances"][:, [0, 3, 5]] = 0.05
    return c


def case_single_colour(alpha):
    """MetalSplatPipelineTest.mm:197-232: colour (0.6, 0.2, 0.4) at 2.5 m; the centre pixel is
    coverage * colour + (1 - coverage) * background, coverage = alpha * 0.991 (1 when alpha > 1)."""
    c = iso_cloud([(0, 0, -2.5)], colors=np.array([[0.6, 0.2, 0.4]], np.float32), alphas=[alpha])
    c["covariances"][:, [0, 3, 5]] = 0.05
    return c, min(1.0, alpha * 0.991)
