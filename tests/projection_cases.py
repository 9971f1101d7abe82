"""Anisotropic, off-axis splats under rotated, non-square cameras, and an FP64 restatement of the
projection written from the shader (SplatProjection.metalh:6-37,78-124) with numpy's linear algebra,
independent of the oracle's C++: Sigma' = J W Sigma W^T J^T + 0.3 I built as matrix products, the
axes from numpy.linalg.eigh. The oracle pins (tests/test_projection_pins.py) compare the oracle with
it; the GPU parity test (tests/test_gpu_parity.py) draws the same scenes against the oracle.

Matrices are column-major float[16] (Mat4.h:10-14): M[r][c] = m[c * 4 + r].
"""
import math

import numpy as np

import cases

K_LOWPASS = 0.3
K_FOV_CLAMP = 1.3
K_MIN_ALPHA = 2.0 / 255.0
K_MAX_PIXEL_RADIUS = 512.0


def quat_to_matrix(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rigid(rot, trans):
    m = np.eye(4)
    m[:3, :3] = rot
    m[:3, 3] = trans
    return m


def col_major(m):
    return np.ascontiguousarray(np.asarray(m, np.float64).T.reshape(16), dtype=np.float32)


def row_major(m16):
    return np.asarray(m16, np.float64).reshape(4, 4).T


def random_view(rng):
    """A camera with a random orientation at a random place: view = [R | -R c]."""
    rot = quat_to_matrix(rng.normal(size=4))
    eye = rng.uniform(-4, 4, 3)
    return rigid(rot, -rot @ eye), eye


def anisotropic_cloud(rng, view_m, tan_x, tan_y, n, spread=1.6, depth=(1.5, 12.0), sigma=(0.01, 0.25),
                      ratio=8.0):
    """n splats placed in view space at |x/z| and |y/z| up to `spread` x the half-FOV tangents (past
    the 1.3 clamp on both axes), depths in `depth`, covariances R diag(s^2) R^T with axis ratios up to
    `ratio`; returned in world space through the inverse of the view."""
    z = -rng.uniform(*depth, n)
    x = rng.uniform(-spread, spread, n) * tan_x * -z
    y = rng.uniform(-spread, spread, n) * tan_y * -z
    inv = np.linalg.inv(view_m)
    world = (inv @ np.stack([x, y, z, np.ones(n)]))[:3].T
    cov = np.zeros((n, 6))
    for i in range(n):
        base = rng.uniform(*sigma)
        s = base * np.array([1.0, rng.uniform(1.0 / ratio, 1.0), rng.uniform(1.0 / ratio, 1.0)])
        r = quat_to_matrix(rng.normal(size=4))
        c = r @ np.diag(s * s) @ r.T
        cov[i] = c[0, 0], c[0, 1], c[0, 2], c[1, 1], c[1, 2], c[2, 2]
    cloud = cases.iso_cloud(world.astype(np.float32))
    cloud["covariances"] = cov.astype(np.float32)
    cloud["colors"] = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    cloud["alphas"] = rng.uniform(0.05, 1.0, n).astype(np.float32)
    return cloud


def edge_cloud(rng, view_m, tan_x, tan_y, n, axis):
    """Large splats centred between 1.3 and 2 times the half-FOV tangent on one axis: their quads
    still reach the screen, so they survive, and their Jacobian uses the clamped coordinate."""
    z = -rng.uniform(2.0, 6.0, n)
    side = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    t_main, t_other = (tan_x, tan_y) if axis == 0 else (tan_y, tan_x)
    main = side * rng.uniform(1.35, 1.9, n) * t_main * -z
    other = rng.uniform(-0.6, 0.6, n) * t_other * -z
    x, y = (main, other) if axis == 0 else (other, main)
    inv = np.linalg.inv(view_m)
    world = (inv @ np.stack([x, y, z, np.ones(n)]))[:3].T
    cov = np.zeros((n, 6))
    for i in range(n):
        s = rng.uniform(0.2, 0.6) * -z[i] * np.array([1.0, rng.uniform(0.25, 1.0), rng.uniform(0.25, 1.0)])
        r = quat_to_matrix(rng.normal(size=4))
        c = r @ np.diag(s * s) @ r.T
        cov[i] = c[0, 0], c[0, 1], c[0, 2], c[1, 1], c[1, 2], c[2, 2]
    cloud = cases.iso_cloud(world.astype(np.float32))
    cloud["covariances"] = cov.astype(np.float32)
    cloud["colors"] = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    cloud["alphas"] = rng.uniform(0.3, 1.0, n).astype(np.float32)
    return cloud


def cameras():
    """(name, width, height, fovy): landscape 16:9, portrait (the phone's 1206 x 2622 aspect), 2:1."""
    return [("landscape", 1280, 720, math.radians(60.0)), ("portrait", 603, 1311, math.radians(65.0)),
            ("wide", 1000, 500, math.radians(50.0))]


def half_to_float(bits):
    return float(np.uint16(bits).view(np.float16))


def unpack_halves(word):
    word = int(word)
    return half_to_float(word & 0xFFFF), half_to_float(word >> 16)


def fp64_project(splat8, view16, proj16, width, height):
    """The shader's projection in FP64 from one packed GpuSplat (positions fp32, covariance as its
    six halves, opacity byte or LOD float): a dict with the view-space point, the clip-space centre,
    the projected covariance (a, b, d), the FP64 axes, the radius and the two cull margins; None
    when the centre tests reject it (behind the camera, outside near/far, opacity below 2/255)."""
    s = np.asarray(splat8, np.uint32)
    p = s[:3].view(np.float32).astype(np.float64)
    rgba = int(s[3])
    lod = int(s[7]) != 0
    alpha = float(np.uint32(s[7]).view(np.float32)) if lod else (rgba >> 24) / 255.0
    if alpha < K_MIN_ALPHA:
        return None
    V = row_major(view16)
    P = row_major(proj16)
    v = V @ np.append(p, 1.0)
    if v[2] >= 0:
        return None
    clip = P @ v
    near = P[2, 3] / P[2, 2]
    if clip[2] < 0 or clip[2] > clip[3] or clip[3] <= near:
        return None
    xx, xy = unpack_halves(s[4])
    xz, yy = unpack_halves(s[5])
    yz, zz = unpack_halves(s[6])
    sigma = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])
    focal = np.array([width * P[0, 0] / 2.0, height * P[1, 1] / 2.0])
    tan_half = np.array([1.0 / P[0, 0], 1.0 / P[1, 1]])
    lim = K_FOV_CLAMP * tan_half
    tx = np.clip(v[0] / v[2], -lim[0], lim[0]) * v[2]
    ty = np.clip(v[1] / v[2], -lim[1], lim[1]) * v[2]
    J = np.array([[focal[0] / v[2], 0.0, -focal[0] * tx / v[2] ** 2],
                  [0.0, focal[1] / v[2], -focal[1] * ty / v[2] ** 2]])
    W = V[:3, :3]
    T = J @ W
    cov = T @ sigma @ T.T + K_LOWPASS * np.eye(2)
    lam, vec = np.linalg.eigh(cov)  # ascending
    mean = 0.5 * (lam[0] + lam[1])
    dist = max(0.1, 0.5 * (lam[1] - lam[0]))
    l1, l2 = mean + dist, mean - dist
    e1 = vec[:, 1]
    e2 = np.array([e1[1], -e1[0]])
    a1 = e1 * math.sqrt(l1)
    a2 = e2 * math.sqrt(max(l2, 0.0))
    radius = min(5.0 if lod else 2.5, math.sqrt(2.0 * math.log(max(alpha / K_MIN_ALPHA, 1.0))))
    radius = min(radius, K_MAX_PIXEL_RADIUS / max(math.sqrt(l1), 1e-6))
    ext = radius * (np.abs(a1) + np.abs(a2))
    margin = (ext + 0.5) * 2.0 * clip[3] / np.array([width, height])
    footprint = clip[3] + margin - np.abs(clip[:2])  # > 0 on both axes: the quad reaches the screen
    source_sigma = math.sqrt(max(lam[1] - K_LOWPASS, 0.0))
    return {"view": v, "clip": clip, "cov": (cov[0, 0], cov[0, 1], cov[1, 1]), "a1": a1, "a2": a2,
            "radius": radius, "footprint": footprint, "subpixel": source_sigma * radius, "alpha": alpha,
            "gap": lam[1] - lam[0], "clamped": (abs(v[0] / v[2]) > lim[0], abs(v[1] / v[2]) > lim[1])}


def record_axes(rec8):
    a1 = np.array(unpack_halves(rec8[2]))
    a2 = np.array(unpack_halves(rec8[3]))
    return a1, a2


def axes_error(rec8, ref):
    """Distance between a record's half axes and the FP64 axes, up to the common sign of the
    eigenvector (e2 is e1 turned, so both axes flip together), relative to the major axis."""
    a1, a2 = record_axes(rec8)
    plus = np.abs(a1 - ref["a1"]).max() + np.abs(a2 - ref["a2"]).max()
    minus = np.abs(a1 + ref["a1"]).max() + np.abs(a2 + ref["a2"]).max()
    return min(plus, minus) / np.linalg.norm(ref["a1"])


def scene(seed, n=600):
    """One test scene per camera shape: (name, width, height, view16, proj16, eye, cloud)."""
    rng = np.random.default_rng(seed)
    out = []
    for name, w, h, fovy in cameras():
        view_m, eye = random_view(rng)
        proj = cases.perspective(fovy, w / h, 0.1, 100.0)
        tan_y = 1.0 / float(proj[5])
        tan_x = 1.0 / float(proj[0])
        parts = [anisotropic_cloud(rng, view_m, tan_x, tan_y, n), edge_cloud(rng, view_m, tan_x, tan_y, n // 6, 0),
                 edge_cloud(rng, view_m, tan_x, tan_y, n // 6, 1)]
        cloud = {k: np.concatenate([c[k] for c in parts]) for k in ("positions", "covariances", "colors", "alphas")}
        cloud["sh_degree"] = 0
        cloud["sh"] = None
        out.append((name, w, h, col_major(view_m), proj, eye.astype(np.float32), cloud))
    return out
