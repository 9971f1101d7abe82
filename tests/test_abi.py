"""The C-ABI libraries load on a machine without a GPU and export every function their headers
declare; the Python binding covers the whole ABI; the scene generator is deterministic. No compute
calls are made here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(skm_\w+)\s*\(", text))


def exported(path):
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_skm_exports_every_declared_function(skm_mod):
    names = declared("skm.h")
    assert len(names) > 40
    syms = exported(skm_mod.LIB_PATH)
    assert names <= syms, sorted(names - syms)
    ctypes.CDLL(skm_mod.LIB_PATH)  # loads without a device
    assert set(skm_mod.API) == names


def test_server_exports_every_declared_function(skm_mod):
    names = declared("skm_server.h")
    assert len(names) >= 9
    assert names <= exported(skm_mod.LIB_PATH), sorted(names - exported(skm_mod.LIB_PATH))
    assert set(skm_mod.SERVER_API) == names
    assert ctypes.sizeof(skm_mod.Request) == 168 and ctypes.sizeof(skm_mod.Completion) == 40
    assert ctypes.sizeof(skm_mod.ServerConfig) == 24 and ctypes.sizeof(skm_mod.ServerStats) == 96


def test_every_header_is_covered():
    """Each public header's functions are exported by one of the libraries."""
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["skm.h", "skm_lod.h", "skm_scene.h", "skm_server.h"]


def test_scene_exports_every_declared_function(skm_mod):
    names = declared("skm_scene.h")
    assert names <= exported(skm_mod.SCENE_PATH)


def test_struct_layouts_match_the_reference(skm_mod):
    C = ctypes
    # CameraUniform 208 bytes (MetalShaderTypes.h:26-30) is the oracle's Camera; RevealUniform 32 B.
    assert C.sizeof(skm_mod.Reveal) == 32
    assert C.sizeof(skm_mod.Range) == 8
    assert skm_mod.RenderPolicy.background_color.offset == 24 and C.sizeof(skm_mod.RenderPolicy) == 48
    # The batch extension's structs (static_assert-ed to the same sizes in renderer.cpp).
    assert C.sizeof(skm_mod.View) == 152 and C.sizeof(skm_mod.BatchTuning) == 24
    assert C.sizeof(skm_mod.Config) == 24
    assert C.sizeof(skm_mod.BatchStats) == 40


def test_renderer_creation_fails_cleanly_without_a_device(skm_mod):
    import subprocess
    import sys

    # In a child process: HIP initialisation must not touch this test process.
    code = ("import sys; sys.path.insert(0, %r); import skm\n"
            "try:\n    skm.Renderer()\nexcept skm.SkmError as e:\n    print('status', e.status)\n"
            "else:\n    print('created')\n") % os.path.join(ROOT, "splatkit-ios_amd")
    out = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=120).stdout
    has_gpu = os.path.exists("/dev/kfd") and "created" in out
    assert has_gpu or "status -2" in out, out


def test_scene_is_deterministic_and_thread_independent(skm_mod):
    a = skm_mod.make_cloud(20000, 3, seed=5, layout=1, spread=5.0, threads=1)
    b = skm_mod.make_cloud(20000, 3, seed=5, layout=1, spread=5.0, threads=7)
    for k in ("positions", "covariances", "colors", "alphas", "sh"):
        assert np.array_equal(a[k], b[k])
    c = skm_mod.make_cloud(20000, 3, seed=6, layout=1, spread=5.0)
    assert not np.array_equal(a["positions"], c["positions"])


def test_scene_distributions(skm_mod):
    c = skm_mod.make_cloud(200_000, 1, seed=11, layout=0, spread=2.0)
    assert abs(c["positions"].std() - 2.0) < 0.02 and abs(c["positions"].mean()) < 0.02
    # trace of R diag(s)^2 R^T = sum of squared scales; log scale ~ N(-4.5, 0.5)
    tr = c["covariances"][:, 0] + c["covariances"][:, 3] + c["covariances"][:, 5]
    expect = 3 * np.exp(2 * -4.5 + 2 * 0.5 ** 2)
    assert abs(tr.mean() / expect - 1) < 0.02
    assert c["alphas"].min() > 0 and c["alphas"].max() < 1 and abs(np.median(c["alphas"]) - 0.5) < 0.01
    assert c["colors"].min() >= 0 and c["colors"].max() <= 1
    assert abs(c["sh"].std() - 0.05) < 0.001
    s = skm_mod.make_cloud(100_000, 0, seed=12, layout=1, spread=5.0, area=10.0)
    near_ground = np.abs(s["positions"][:, 1]) < 0.2
    assert 0.3 < near_ground.mean() < 0.5


def test_scene_covariance_matches_reference_code(skm_mod, oracle, ref):
    """The generator's R diag(s)^2 R^T against covarianceFromQuaternionScale (Covariance.cpp:7-17)."""
    r = np.random.default_rng(3)
    for _ in range(200):
        q = r.normal(size=4).astype(np.float32)
        q /= np.float32(np.linalg.norm(q))
        s = np.exp(r.normal(-4.5, 0.5, 3)).astype(np.float32)
        out = np.zeros(6, np.float32)
        ref.ref_covariance(oracle.fp(q), oracle.fp(s), oracle.fp(out))
        R = np.array([[1 - 2 * (q[1] ** 2 + q[2] ** 2), 2 * (q[0] * q[1] - q[3] * q[2]), 2 * (q[0] * q[2] + q[3] * q[1])],
                      [2 * (q[0] * q[1] + q[3] * q[2]), 1 - 2 * (q[0] ** 2 + q[2] ** 2), 2 * (q[1] * q[2] - q[3] * q[0])],
                      [2 * (q[0] * q[2] - q[3] * q[1]), 2 * (q[1] * q[2] + q[3] * q[0]), 1 - 2 * (q[0] ** 2 + q[1] ** 2)]],
                     np.float64)
        S = R @ np.diag(s.astype(np.float64) ** 2) @ R.T
        np.testing.assert_allclose(out, S[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]], rtol=1e-4, atol=1e-10)


def test_look_at_and_perspective(skm_mod):
    v = skm_mod.look_at((0, 0, 6)).reshape(4, 4).T
    assert np.allclose(v @ [0, 0, 0, 1], [0, 0, -6, 1])
    v = skm_mod.look_at((3, 1, 2), (0, 0, 0)).reshape(4, 4).T
    assert np.allclose(v @ [3, 1, 2, 1], [0, 0, 0, 1], atol=1e-6)
    assert np.allclose((v @ [0, 0, 0, 1])[:2], 0, atol=1e-6)
    import cases

    assert np.array_equal(skm_mod.perspective(1.0, 1.5, 0.1, 100.0), cases.perspective(1.0, 1.5, 0.1, 100.0))


def test_lod_header_is_exported(skm_mod):
    names = declared("skm_lod.h")
    assert names <= exported(skm_mod.LIB_PATH)
    assert set(skm_mod.LOD_API) == names
