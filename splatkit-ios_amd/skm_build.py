"""Builds the in-tree shared libraries of the render path.

  libskm.so        HIP kernels for gfx950 + the C-ABI renderer (include/skm.h)
  libskm_scene.so  host-only synthetic scenes and camera matrices (include/skm_scene.h)

Every HIP source is compiled with -ffp-contract=off and without fast-math (each fp32 operation is
one IEEE rounding, which the bit-exact parity of the cull relies on) and with packed-fp32
instructions disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SKM_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in ("kernels.hip", "renderer.cpp", "batch.cpp", "lod_build.cpp", "lod.hip", "server.cpp", "noalloc.hip", "lod_file.cpp"):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append(([HIPCC] + HIP_FLAGS + lang + ["-c", path, "-o", obj], obj))
    # One compiler per source, side by side.
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 1))) as pool:
        outs = list(pool.map(lambda job: _run(job[0]), jobs))
    if verbose:
        for out in outs:
            if out.strip():
                print(out)
    objs = [obj for _, obj in jobs]
    lib = os.path.join(HERE, "libskm.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib + ".tmp"] + objs)
    os.replace(lib + ".tmp", lib)
    scene = os.path.join(HERE, "libskm_scene.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-pthread", f"-I{INCLUDE}",
          os.path.join(CSRC, "scene.cpp"), "-o", scene + ".tmp"])
    os.replace(scene + ".tmp", scene)
    return lib, scene


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
