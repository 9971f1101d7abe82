"""Throughput of the per-frame splat render path on one MI355X.

Workload (default): the 3M-Gaussian SH3 scene at 1920x1080 on a 60-frame orbit (R = 8 m,
h = 1.5 m, 6 degrees per frame, 60 degree vertical FOV), full cull + 32-bit radix sort + tile
binning + raster each frame, scene resident in HBM before timing. A step is one frame.

Prints one JSON line: frames/s (value), the per-stage GPU times measured with HIP events on the
renderer's stream over the timed frames, the roofline of the dominant kernel, and the CPU
baseline (the C++ oracle's FP32 frame on the host cores, bounded to a few frames).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--cpu-frames F]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "splatkit-ios_amd"))

HBM_PEAK = 8.0e12          # bytes/s, MI355X HBM3E
# VALU issue: 256 CUs x 4 SIMDs x 16 lanes per clock at 2.4 GHz, one lane operation per lane per
# instruction (a wave64 instruction takes 4 clocks on a 16-lane SIMD). The packed-fp32 figure
# (157.3 TFLOP/s) counts 2 per FMA and 2 per packed pair; this build disables packed fp32.
VALU_LANE_OPS_PEAK = 256 * 4 * 16 * 2.4e9
WIDTH, HEIGHT = 1920, 1080
N_SPLATS = 3_000_000
SH_DEGREE = 3
# Frames a draw may queue before it waits (each with its own HIP stream and scratch, so consecutive
# frames overlap on the GPU); the reference keeps two.
FRAMES_IN_FLIGHT = 3
SH_BYTES = {0: 0, 1: 20, 2: 48, 3: 92}
ORBIT_FRAMES = 60


def measured_valu(kernel, steps):
    """Wave64 VALU instructions per launch of `kernel` over the orbit frames a timed region of `steps`
    draws (frame i mod 60), from the committed SQ_INSTS_VALU counter run of this workload
    (profiles/sq_counters_3M_sh3_1080p.json, one value per orbit frame), or None when absent."""
    path = os.path.join(ROOT, "profiles", "sq_counters_3M_sh3_1080p.json")
    if not os.path.exists(path):
        return None
    per_frame = json.load(open(path)).get("kernels", {}).get(kernel, {}).get("valu_by_orbit_frame")
    if not per_frame or len(per_frame) != ORBIT_FRAMES:
        return None
    return sum(per_frame[i % ORBIT_FRAMES] for i in range(steps)) / steps


def measured_one_slot_valu():
    """Wave64 VALU instructions per raster launch of the 60 orbit frames drawn with one frame slot (the
    split raster), from the committed counter run (profiles/sq_counters_3M_sh3_1080p.json, "one_slot"),
    or None when absent."""
    path = os.path.join(ROOT, "profiles", "sq_counters_3M_sh3_1080p.json")
    if not os.path.exists(path):
        return None
    per_frame = json.load(open(path)).get("one_slot", {}).get("valu_by_orbit_frame")
    return sum(per_frame) / len(per_frame) if per_frame else None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload, from the committed counter runs (fetched
    and written bytes, each scaled by a 1 GiB copy's reading), or None when they are absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_3M_sh3_1080p.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path))
    for name, v in data["kernels"].items():
        if name == kernel:
            return v["fetch_bytes"] + v["write_bytes"]
    return None


def measured_frame_valu():
    """Wave64 VALU instructions of one frame, summed over its kernels (launches per frame x instructions
    per launch, from the committed counter run), or None when absent."""
    path = os.path.join(ROOT, "profiles", "sq_counters_3M_sh3_1080p.json")
    if not os.path.exists(path):
        return None
    kernels = json.load(open(path)).get("kernels", {})
    return sum(e["launches_per_frame"] * e["per_launch_mean"].get("SQ_INSTS_VALU", 0.0) for e in kernels.values()) or None


def stage_bytes(N, V, P, T, depth_passes, tile_passes):
    """Algorithmic HBM bytes per frame of each stage, in the reference's packed-record units."""
    k = 4
    return {
        "cull": 32 * N + SH_BYTES[SH_DEGREE] * V + 36 * V,
        "depth_sort": V * 2 * (k + 4) * depth_passes + k * V,
        "bin_tile_sort_ranges": (4 + 32) * V + 8 * P + P * 2 * (k + 4) * tile_passes + k * P + k * P + 8 * T,
        "raster": (4 + 32) * P + 16 * WIDTH * HEIGHT,
    }


def spatially_ordered(r, cloud):
    """The world as the reference's loader hands it to a renderer: reordered along the Morton curve over
    its bounds (reorderSpatially, SplatWorldLoader.cpp:197), here with skm_reorder_spatially on the GPU.
    Loading, outside every timed region; returns the seconds it took."""
    t0 = time.perf_counter()
    r.reorder_spatially(cloud)
    return time.perf_counter() - t0


def tune_single_draws(r, frames, draws=12):
    """Times single draws of the orbit under every candidate the autotuner knows (its table of a
    one-view call gives the candidate set); returns the fastest tuning and (tile, px, sort, radix, ms)."""
    _, candidates = r.autotune_views([frames[0] + (WIDTH, HEIGHT)], repeats=1)
    table = []
    for t, p, k, rb, _ in candidates:
        r.set_tuning(tile_size=t, raster_px=p, sort_kernel=k, radix_bits=rb)
        for i in range(2):
            r.draw(r.make_frame(*frames[(7 * i) % 60]))
        r.finish()
        t0 = time.perf_counter()
        for i in range(draws):
            r.draw(r.make_frame(*frames[(5 * i) % 60]))
        r.finish()
        table.append((t, p, k, rb, (time.perf_counter() - t0) * 1e3 / draws))
    t, p, k, rb, _ = min(table, key=lambda row: row[4])
    return {"tile_size": t, "raster_px": p, "sort_kernel": k, "radix_bits": rb}, table


def radix_sort(r, skm, runs):
    """The stable LSD argsort alone (skm_argsort_u32_device) on uniformly random keys resident in HBM,
    for (n, key bits, label) runs: the depth sort's algorithm at and above the frame's sizes, and the
    tile sort at the frame's pair count P with the frame's tile-id width. Algorithmic bytes in the
    reference's units, passes actually executed: n (2 (k + 4) passes + k) with k = 4."""
    out = []
    rng = np.random.default_rng(0x5EED0005)
    for n, bits, label in runs:
        keys = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
        src, work, order = skm.DeviceBuffer(4 * n), skm.DeviceBuffer(4 * n), skm.DeviceBuffer(4 * n)
        src.upload(keys)
        ms = []
        for _ in range(4):
            work.copy_from(src, 4 * n)  # the sort overwrites its keys
            ms.append(r.argsort_device(work.ptr.value, n, bits, order.ptr.value))
        if n <= 16_000_000:
            assert np.array_equal(order.download(np.uint32, n), np.argsort(keys, kind="stable"))
        t = min(ms[1:])
        passes = -(-bits // 8)
        byt = n * (2 * (4 + 4) * passes + 4)
        out.append({"label": label, "keys": n, "key_bits": bits, "passes": passes, "ms": round(t, 4),
                    "GB_s": round(byt / t / 1e6, 1), "hbm_frac": round(byt / t / 1e6 / (HBM_PEAK / 1e9), 4)})
        for b in (src, work, order):
            b.close()
    return {"workload": "stable argsort of uniformly random keys resident in HBM (ceil(bits / 8) passes, bits split "
                        "evenly)", "runs": out}


def batched_vs_loop(r, skm, batches, loops, single_tuning):
    """The 64-view batch of the same scene (720p/1080p/1440p views on spheres of radius 4-12 m): views/s
    of one graph launch per batch against the plain path drawing the same views one per launch (under
    the single-draw tuning)."""
    views = skm.views_c2(64)
    arr = r.make_views(views)
    best, table = r.autotune_views(arr, repeats=2)
    for _ in range(2):
        r.render_views(arr, **best)
    t0 = time.perf_counter()
    for _ in range(batches):
        r.draw_views(arr, **best)
    r.finish()
    batch_s = (time.perf_counter() - t0) / batches
    st = r.batch_stats()
    frames = [(r.make_frame(v, p, e), w, h) for v, p, e, w, h in views]
    r.set_tuning(**single_tuning)
    # One untimed pass over every view first: a view whose tile pairs outgrow the capacity is dropped
    # (skm_finish reports it) and the capacity grows, so the timed passes draw every view whole.
    for attempt in range(3):
        for f, w, h in frames:
            r.set_extent(w, h)
            r.draw(f)
        try:
            r.finish()
            break
        except skm.SkmError as e:
            if e.status != skm.SKM_ERR_CAPACITY or attempt == 2:
                raise
    r.frame_totals(reset=True)
    t0 = time.perf_counter()
    for _ in range(loops):
        for f, w, h in frames:
            r.set_extent(w, h)
            r.draw(f)
    r.finish()
    loop_s = (time.perf_counter() - t0) / loops
    drawn = r.frame_totals(reset=True)
    if drawn["dropped"] != 0 or drawn["frames"] != loops * len(frames):
        raise SystemExit(f"per-view loop: {drawn['dropped']} views dropped, {drawn['frames']} drawn")
    return {"workload": "64 views of the 3M SH3 scene, resolutions uniform over 720p/1080p/1440p, eyes on "
                        "spheres of radius U(4,12) m", "views": 64,
            "batched_frames_per_s": round(64 / batch_s, 1), "batched_ms_per_batch": round(batch_s * 1000, 3),
            "batched_gpu_ms_per_batch": round(st["gpu_ms"], 3), "graph": st["graph"],
            "per_view_loop_frames_per_s": round(64 / loop_s, 1), "per_view_loop_ms": round(loop_s * 1000, 3),
            "speedup": round(loop_s / batch_s, 3), "visible_per_batch": st["visible"], "pairs_per_batch": st["pairs"],
            "batches": batches, "loops": loops, "autotuned": best,
            "autotune_ms": [{"tile": t, "raster_px": p, "sort_kernel": k, "radix_bits": rb, "gpu_ms": round(ms, 3)}
                            for t, p, k, rb, ms in table]}


SERVING_SIZES = (500_000, 1_000_000, 2_000_000, 3_000_000, 5_000_000, 6_000_000, 8_000_000, 10_000_000)
SERVING_RESOLUTIONS = ((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))


# Continuous batching: up to 3 requests of a scene per launch and 3 launches in flight (measured on
# MI355X against 2, 4, 6, 8 and 64 requests per launch: 4 saturates within 1% of 3 but doubles the p99
# near saturation, larger launches are slower; 3 launches plus the process's own stream fill its 4
# hardware queues).
SERVING_MODES = (("continuous_batching", 3, 3), ("one_request_per_launch", 1, 1))


def pixels_to_host(block_bytes, copies=96, streams=3):
    """Device-to-host copy rate into pinned host memory for blocks of one served request's BGRA8 pixels
    (the last step of every request), with copies on `streams` streams, after a warm-up round."""
    import torch

    n = int(block_bytes)
    src = torch.empty(8 * n, dtype=torch.uint8, device="cuda")
    dst = [torch.empty(n, dtype=torch.uint8, pin_memory=True) for _ in range(8)]
    ss = [torch.cuda.Stream() for _ in range(streams)]
    rate = 0.0
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(copies):
            with torch.cuda.stream(ss[i % streams]):
                dst[i % 8].copy_(src[(i % 8) * n:(i % 8 + 1) * n], non_blocking=True)
        torch.cuda.synchronize()
        rate = copies * n / (time.perf_counter() - t0)
    del src, dst
    return rate


SERVING_MEAN_BYTES = 4.0 * sum(w * h for w, h in SERVING_RESOLUTIONS) / len(SERVING_RESOLUTIONS)


def serving(skm, rates, per_rate, modes=SERVING_MODES, d2h=None):
    """8 resident scenes (0.5M-10M splats, SH0-3 cyclic, drawn like the 3M scene), requests with a
    random scene, a random camera on a sphere of radius 4-12 m and a resolution from 720p to 2160p,
    arriving as a Poisson stream at each rate. Continuous batching against one request per launch."""
    import numpy as np

    clouds = [skm.make_cloud(n, k % 4, seed=0x5EED0004 + k, layout=1, spread=5.0, area=10.0)
              for k, n in enumerate(SERVING_SIZES)]
    loader = skm.Renderer(0)
    for c in clouds:
        spatially_ordered(loader, c)
    loader.close()
    out = {"workload": "8 resident scenes of 0.5/1/2/3/5/6/8/10M splats, SH0-3 cyclic; requests uniform over "
                       "scenes, cameras on spheres of radius U(4,12) m, resolutions uniform over "
                       "720p/1080p/1440p/2160p; Poisson arrivals; BGRA8 pixels delivered to host memory",
           "requests_per_rate": per_rate, "modes": {}}
    for mode, batch, flight in modes:
        srv = skm.Server(max_batch_views=batch, max_in_flight=flight)
        for c in clouds:
            srv.add_scene(c)
        rows = []
        for k, rate in enumerate(rates):
            rng = np.random.default_rng(1000 + k)
            views = skm.views_c2(per_rate, seed=2000 + k, resolutions=SERVING_RESOLUTIONS)
            scenes = rng.integers(0, len(clouds), per_rate)
            if k == 0:  # warm-up: scratch growth outside the measured window
                for t in srv.wait_all([srv.submit(int(s), *v) for s, v in zip(scenes[:32], views[:32])], timeout_s=300):
                    srv.release(t)
                srv.stats(reset=True)
            arrivals = np.cumsum(rng.exponential(1.0 / rate, per_rate))
            # The client takes each result as it completes and releases it (its pinned output block
            # returns to the server's pool), as a serving client would.
            done = {}

            def take(timeout_ms):
                for c in srv.poll(timeout_ms=timeout_ms):
                    done[c["ticket"]] = c
                    srv.release(c["ticket"])

            t0 = time.perf_counter()
            last_poll = 0.0
            for s, v, at in zip(scenes, views, arrivals):
                while (now := time.perf_counter() - t0) < at:
                    if now - last_poll > 2e-4:
                        take(0.0)
                        last_poll = now
                srv.submit(int(s), *v)
            deadline = time.perf_counter() + 300
            while len(done) < per_rate and time.perf_counter() < deadline:
                take(50.0)
            wall = time.perf_counter() - t0
            st = srv.stats(reset=True)
            ok = sum(1 for c in done.values() if c["status"] == 0)
            rows.append({"offered_per_s": rate, "frames_per_s": round(ok / wall, 1), "p50_ms": round(st["p50_ms"], 2),
                         "p99_ms": round(st["p99_ms"], 2), "mean_ms": round(st["mean_ms"], 2),
                         "mean_batch_views": round(st["mean_batch_views"], 2), "launches": st["launches"],
                         "rejected": st["rejected"]})
        out["modes"][mode] = rows
        srv.close()
    best = {m: max(r["frames_per_s"] for r in rows) for m, rows in out["modes"].items()}
    out["max_frames_per_s"] = best
    if "continuous_batching" in best and "one_request_per_launch" in best:
        out["speedup_at_saturation"] = round(best["continuous_batching"] / best["one_request_per_launch"], 3)
    # Every request's pixels cross PCIe to host memory: the copy rate bounds the requests per second.
    if d2h:
        out["pixels_to_host"] = {"mean_request_bytes": int(SERVING_MEAN_BYTES), "d2h_GB_s": round(d2h / 1e9, 1),
                                 "ceiling_requests_per_s": round(d2h / SERVING_MEAN_BYTES, 1),
                                 "note": "device-to-host copies of mean-request-sized BGRA8 blocks into pinned host "
                                         "memory on 3 streams, measured in a child process before this one uses "
                                         "the GPU; requests/s cannot exceed d2h / mean_request_bytes"}
    return out


def counter_leg(skm, leg):
    """One leg's GPU work once, untimed and with nothing printed but a summary line, for a profiler's
    counter run (which serialises dispatches). batch: two 64-view launches of the 3M scene (the second is
    the measured one); loop: the same 64 views drawn one per launch twice (the first pass lets the pair
    buffers grow); serve / serve_single: 288 requests of the 8-scene serving mix, the first 32 a warm-up,
    with continuous batching (3 per launch, 3 in flight) or one request per launch; one_slot: the 60 orbit
    frames with one frame slot (the split raster), after 5 untimed ones."""
    if leg == "one_slot":
        cloud = skm.scene_c1(N_SPLATS, SH_DEGREE)
        r = skm.Renderer(0, frames_in_flight=1)
        spatially_ordered(r, cloud)
        r.upload_world(cloud)
        r.set_extent(WIDTH, HEIGHT)
        r.apply_policy(sort_depth=32)
        r.set_tuning(tile_size=16, raster_px=1, sort_kernel=0, radix_bits=7)
        frames = [skm.orbit_camera(f, WIDTH, HEIGHT) for f in range(ORBIT_FRAMES)]
        for i in range(5):
            r.render(*frames[i])
        r.frame_totals(reset=True)
        for f in frames:
            r.draw(r.make_frame(*f))
        r.finish()
        return {"leg": leg, **r.frame_totals(reset=True)}
    if leg in ("batch", "loop"):
        cloud = skm.scene_c1(N_SPLATS, SH_DEGREE)
        r = skm.Renderer(0, frames_in_flight=FRAMES_IN_FLIGHT)
        spatially_ordered(r, cloud)
        r.upload_world(cloud)
        r.apply_policy(sort_depth=32)
        views = skm.views_c2(64)
        tuning = dict(tile_size=16, raster_px=1, sort_kernel=0, radix_bits=8)
        if leg == "batch":
            arr = r.make_views(views)
            for _ in range(2):
                r.draw_views(arr, **tuning)
                r.finish()
            return {"leg": leg, **r.batch_stats()}
        r.set_tuning(**tuning)
        frames = [(r.make_frame(v, p, e), w, h) for v, p, e, w, h in views]
        for attempt in range(3):
            for f, w, h in frames:
                r.set_extent(w, h)
                r.draw(f)
            try:
                r.finish()
                break
            except skm.SkmError as e:
                if e.status != skm.SKM_ERR_CAPACITY or attempt == 2:
                    raise
        r.frame_totals(reset=True)
        for f, w, h in frames:
            r.set_extent(w, h)
            r.draw(f)
        r.finish()
        return {"leg": leg, **r.frame_totals(reset=True)}
    clouds = [skm.make_cloud(n, k % 4, seed=0x5EED0004 + k, layout=1, spread=5.0, area=10.0)
              for k, n in enumerate(SERVING_SIZES)]
    loader = skm.Renderer(0)
    for c in clouds:
        spatially_ordered(loader, c)
    loader.close()
    _, batch, flight = SERVING_MODES[0 if leg == "serve" else 1]
    srv = skm.Server(max_batch_views=batch, max_in_flight=flight)
    for c in clouds:
        srv.add_scene(c)
    rng = np.random.default_rng(1003)
    views = skm.views_c2(288, seed=2003, resolutions=SERVING_RESOLUTIONS)
    scenes = rng.integers(0, len(clouds), 288)
    failed = []
    for lo, hi in ((0, 32), (32, 288)):
        tickets = [srv.submit(int(s), *v) for s, v in zip(scenes[lo:hi], views[lo:hi])]
        done = srv.wait_all(tickets, timeout_s=300)
        for k, t in enumerate(tickets):
            if done[t]["status"] != 0:
                failed.append({"request": lo + k, "scene": int(scenes[lo + k]), "size": list(views[lo + k][3:5]),
                               "status": done[t]["status"], "batch_views": done[t]["batch_views"]})
            srv.release(t)
    return {"leg": leg, "requests": 288, **srv.stats(reset=True), "failed": failed}


def fly_camera(frame, frames, width=1920, height=1080):
    """A straight pass over the 200 m square at 4 m height, weaving 30 m either side, looking ahead
    and slightly down."""
    import numpy as np
    import skm

    t = frame / max(1, frames - 1)
    eye = np.array([-90.0 + 180.0 * t, 4.0, 30.0 * math.sin(2 * math.pi * t)], np.float32)
    ahead = eye + np.array([10.0, -1.0, 30.0 * 2 * math.pi / 180.0 * math.cos(2 * math.pi * t)], np.float32)
    return skm.look_at(eye, ahead), skm.perspective(math.pi / 3, width / height, 0.2, 400.0), eye


LOD_PASSES = 5


def lod_flythrough(skm, frames, leaves, frames_in_flight=FRAMES_IN_FLIGHT):
    """The LOD world: `leaves` splats at SH1 drawn like the 3M scene over a 200 x 200 m square, an
    octree with at most 256 leaves per cluster, a cut per frame at 1.5 px error (color weight 4,
    culling on), 1920x1080, over a 120-frame fly-through, with 16-bit and 32-bit depth keys."""
    t0 = time.time()
    cloud = skm.make_cloud(leaves, 1, seed=0x5EED0003, layout=1, spread=100.0, area=200.0)
    gen_s = time.time() - t0
    r = skm.Renderer(0, frames_in_flight=frames_in_flight)
    spatially_ordered(r, cloud)  # the loader reorders before it builds the tree (SplatWorldLoader.cpp:197-205)
    t0 = time.time()
    built = skm.LodTree.build(cloud, 21, 256, threads=min(16, os.cpu_count() or 1))
    build_s = time.time() - t0
    del cloud
    # The world reaches the renderer from a LODSPLAT file, as the reference's loader takes it: the tree
    # is written (writeLodSplat's format) and read back through the decoder, whose node limit is lifted
    # past the reference's 20M (this tree has more nodes).
    import tempfile

    fd, path = tempfile.mkstemp(suffix=".lodsplat")
    os.close(fd)
    os.unlink(path)
    t0 = time.time()
    built.write(path)
    write_s = time.time() - t0
    file_bytes = os.path.getsize(path)
    built.close()
    t0 = time.time()
    tree = skm.LodTree.read(path)
    read_s = time.time() - t0
    os.unlink(path)
    r.upload_lod_world(tree, 1)
    r.set_extent(WIDTH, HEIGHT)
    cams = [fly_camera(f, frames) for f in range(frames)]
    out = {"workload": f"{leaves} leaves SH1 over 200x200 m, octree (<=256 leaves per cluster), cut at 1.5 px, "
                       f"1920x1080, {frames}-frame fly-through", "leaves": leaves, "nodes": tree.node_count,
           "tree_build_s": round(build_s, 2), "scene_generation_s": round(gen_s, 2),
           "lodsplat_file": {"bytes": file_bytes, "write_s": round(write_s, 2), "read_s": round(read_s, 2)},
           "keys": {}}
    names = ["lod_cut", "cull", "depth_sort", "bin_tile_sort_ranges", "raster"]
    for bits in (16, 32):
        r.apply_policy(sort_depth=bits, lod_error=1.5)
        for f in range(5):
            r.render(*cams[f])
        r.frame_totals(reset=True)
        # The fly-through five times over (one pass is only about 45 ms of frames).
        t0 = time.perf_counter()
        for _ in range(LOD_PASSES):
            for f in range(frames):
                r.draw(r.make_frame(*cams[f]))
        r.finish()
        elapsed = time.perf_counter() - t0
        tot = r.frame_totals(reset=True)
        if tot["dropped"] != 0 or tot["frames"] != LOD_PASSES * frames:
            raise SystemExit(f"LOD fly-through: {tot['dropped']} frames dropped, {tot['frames']} drawn of "
                             f"{LOD_PASSES * frames}")
        n = max(1, tot["frames"])
        out["keys"][str(bits)] = {"frames_per_s": round(LOD_PASSES * frames / elapsed, 1), "passes": LOD_PASSES,
                                  "stage_ms": {k: round(tot["stage_ms"][i] / n, 4) for i, k in enumerate(names)},
                                  "gpu_ms_per_frame": round(tot["stage_ms"][5] / n, 4),
                                  "visible_per_frame": int(tot["visible"] / n), "pairs_per_frame": int(tot["pairs"] / n)}
    r.close()
    tree.close()
    return out


def cpu_baseline(frames, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    import skm

    cloud = skm.scene_c1(N_SPLATS, SH_DEGREE)
    # The same world order as the GPU legs (the loader's spatial order), restated on the CPU.
    order = pyoracle.spatial_order(cloud["positions"], cloud["positions"].min(0), cloud["positions"].max(0))
    cloud = {k: (v[order] if isinstance(v, np.ndarray) else v) for k, v in cloud.items()}
    splats = pyoracle.pack_splats(cloud)
    sh = pyoracle.pack_sh(cloud, SH_DEGREE)
    total = 0.0
    stages = np.zeros(4)
    for f in range(frames):
        view, proj, eye = skm.orbit_camera(f, WIDTH, HEIGHT)
        cam = pyoracle.make_camera(view, proj, WIDTH, HEIGHT, eye)
        secs, vis, pairs, _ = pyoracle.render_frame_timed(cam, splats, sh, SH_DEGREE, False, (0.05, 0.05, 0.08), threads)
        stages += secs
        total += sum(secs)
    return frames / total, (stages / frames * 1000).round(1).tolist()


def pixel_error_fp64(images):
    """max |GPU - FP64| per channel of the two frames drawn on the GPU above, against the oracle's FP64
    composite of the same world, camera and order (the checker; in the CPU leg)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    import skm

    cloud = skm.scene_c1(N_SPLATS, SH_DEGREE)
    order = pyoracle.spatial_order(cloud["positions"], cloud["positions"].min(0), cloud["positions"].max(0))
    cloud = {k: (v[order] if isinstance(v, np.ndarray) else v) for k, v in cloud.items()}
    o1 = pyoracle.frame(cloud, *skm.orbit_camera(0, WIDTH, HEIGHT), WIDTH, HEIGHT)
    del cloud
    o0 = pyoracle.frame(skm.scene_c0(), *skm.camera_c0(), 1280, 720)
    return {"C0_100k_sh0_720p": float(np.abs(images["C0"] - o0["image"]).max()),
            "C1_3M_sh3_1080p_frame0": float(np.abs(images["C1"] - o1["image"]).max()),
            "bar": 1.0 / 255.0,
            "note": "max over pixels and RGB of |fp32 GPU composite - FP64 oracle composite| on the same "
                    "survivors, order and tile lists"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--batches", type=int, default=5, help="timed 64-view batches (0 skips the batch leg)")
    ap.add_argument("--serving-requests", type=int, default=1000,
                    help="requests per offered rate in the serving leg (0 skips it)")
    ap.add_argument("--serving-rates", default="150,300,600,1200,2400")
    ap.add_argument("--lod-leaves", type=int, default=20_000_000, help="LOD leg's leaves (0 skips it)")
    ap.add_argument("--d2h-probe", action="store_true",
                    help="print the device-to-host copy rate of served-request-sized blocks (bytes/s) and exit")
    ap.add_argument("--lod-frames", type=int, default=120)
    ap.add_argument("--sort-sizes", default="7000000,16000000,154000000",
                    help="radix-sort leg: key counts (empty skips it)")
    ap.add_argument("--frames-in-flight", type=int, default=FRAMES_IN_FLIGHT,
                    help="frame slots of the plain and LOD renderers (1 to 3)")
    ap.add_argument("--tuning", default="",
                    help="tile,raster_px,sort_kernel,radix_bits: skip the single-draw tuner (for counter runs, "
                         "whose serialised dispatches would time the candidates wrongly)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a known-size kernel for the HBM "
                         "byte counters of a profiler)")
    ap.add_argument("--counter-leg", choices=("batch", "loop", "serve", "serve_single", "one_slot"), default="",
                    help="run one leg's work once for a profiler's counter run and exit (no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the render path on a single MI355X "
              "(multi-GPU sharding is design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import skm

    if args.counter_leg:
        print(json.dumps(counter_leg(skm, args.counter_leg)))
        return
    if args.d2h_probe:
        print(pixels_to_host(SERVING_MEAN_BYTES))
        return
    # The copy rate is measured in a child process (torch's own runtime), before this process uses the GPU.
    d2h = None
    if args.serving_requests > 0:
        import subprocess

        probe = subprocess.run([sys.executable, os.path.abspath(__file__), "--d2h-probe"], capture_output=True,
                               text=True, timeout=300)
        if probe.returncode == 0 and probe.stdout.strip():
            d2h = float(probe.stdout.strip().splitlines()[-1])
    if args.copy_calibration_mib > 0:
        import torch

        src = torch.ones(args.copy_calibration_mib * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t0 = time.time()
    cloud = skm.scene_c1(N_SPLATS, SH_DEGREE)
    gen_s = time.time() - t0
    r = skm.Renderer(0, frames_in_flight=args.frames_in_flight)
    device = r.description()
    reorder_s = spatially_ordered(r, cloud)
    r.upload_world(cloud)
    r.set_extent(WIDTH, HEIGHT)
    r.apply_policy(sort_depth=32)
    frames = [skm.orbit_camera(f, WIDTH, HEIGHT) for f in range(60)]
    # Tuning for single draws: every candidate of the autotuner (tile size, raster block shape, radix
    # pass variant, tile digit width) times 12 single draws of the orbit; the fastest is kept.
    if args.tuning:
        t, p, k, rb = (int(x) for x in args.tuning.split(","))
        plain_best, plain_table = {"tile_size": t, "raster_px": p, "sort_kernel": k, "radix_bits": rb}, []
    else:
        plain_best, plain_table = tune_single_draws(r, frames)
    r.set_tuning(**plain_best)
    for i in range(args.warmup):
        r.draw(r.make_frame(*frames[i % 60]))
    r.finish()
    r.frame_totals(reset=True)

    t0 = time.perf_counter()
    for i in range(args.steps):
        r.draw(r.make_frame(*frames[i % 60]))
    r.finish()
    elapsed = time.perf_counter() - t0
    tot = r.frame_totals(reset=True)
    fps = args.steps / elapsed
    # Every timed frame was drawn whole: a frame whose tile pairs overflowed would have been dropped
    # (skm_finish above raises then) and is counted apart.
    if tot["dropped"] != 0 or tot["frames"] != args.steps:
        raise SystemExit(f"timed region: {tot['dropped']} frames dropped, {tot['frames']} drawn of {args.steps}")

    # The same frames one at a time (each drawn and finished before the next): kernel durations with
    # the GPU to themselves, beside the overlapped ones of the timed region. These frames also count
    # the raster's coverage tests (off in the timed region: the count costs an atomic per raster block).
    r.count_raster_tests(True)
    for i in range(20):
        r.draw(r.make_frame(*frames[i % 60]))
        r.finish()
    alone = r.frame_totals(reset=True)
    r.count_raster_tests(False)

    # One frame slot (frames do not overlap, the latency form): the raster then draws its heaviest
    # tiles as four quadrant blocks each. The 60 orbit frames, after 5 untimed ones.
    one = skm.Renderer(0, frames_in_flight=1)
    one.upload_world(cloud)
    one.set_extent(WIDTH, HEIGHT)
    one.apply_policy(sort_depth=32)
    one.set_tuning(**plain_best)
    for i in range(5):
        one.draw(one.make_frame(*frames[i]))
    one.finish()
    one.frame_totals(reset=True)
    t0 = time.perf_counter()
    for i in range(ORBIT_FRAMES):
        one.draw(one.make_frame(*frames[i]))
    one.finish()
    one_s = time.perf_counter() - t0
    one_tot = one.frame_totals(reset=True)
    one.close()
    one_raster_ms = one_tot["stage_ms"][4] / max(1, one_tot["frames"])
    one_valu = measured_one_slot_valu()
    one_slot = {"frames_per_s": round(ORBIT_FRAMES / one_s, 1), "frames": one_tot["frames"],
                "raster_ms": round(one_raster_ms, 4),
                "raster_valu_instructions_per_launch": round(one_valu) if one_valu else None,
                "raster_valu_issue_frac": (round(one_valu * 64 / (one_raster_ms * 1e-3) / VALU_LANE_OPS_PEAK, 4)
                                           if one_valu and one_raster_ms > 0 else None),
                "gpu_ms_per_frame": round(one_tot["stage_ms"][5] / max(1, one_tot["frames"]), 4),
                "note": "frames_in_flight 1, orbit frames 0-59: each frame waits for the previous; the raster "
                        "splits its heaviest tiles (whole order buckets, up to 1024 tiles) into quadrant blocks. "
                        "raster_valu_issue_frac: SQ_INSTS_VALU x 4 / (1024 SIMDs x 2.4 GHz x raster_ms), the "
                        "instruction count from the counter run of these frames (profiles/sq_counters_3M_sh3_1080p.json)"}

    n_frames = max(1, tot["frames"])
    V = tot["visible"] / n_frames
    P = tot["pairs"] / n_frames
    T = ((WIDTH + plain_best["tile_size"] - 1) // plain_best["tile_size"]) * \
        ((HEIGHT + plain_best["tile_size"] - 1) // plain_best["tile_size"])
    names = ["cull", "depth_sort", "bin_tile_sort_ranges", "raster"]
    ms = {n: tot["stage_ms"][k + 1] / n_frames for k, n in enumerate(names)}
    tile_bits = max(1, (T - 1).bit_length())
    by = stage_bytes(N_SPLATS, V, P, T, 4, -(-tile_bits // plain_best["radix_bits"]))
    stages = {}
    for n in names:
        gbs = by[n] / (ms[n] * 1e-3) / 1e9 if ms[n] > 0 else 0.0
        stages[n] = {"ms": round(ms[n], 4), "bytes": int(by[n]), "GB_s": round(gbs, 1),
                     "hbm_frac": round(gbs * 1e9 / HBM_PEAK, 4)}
    tile = plain_best["tile_size"]
    # Raster work: the pair evaluations of the tile lists (E = tile^2 x P, an upper bound), and the
    # (pixel, record) tests the kernel performed (counted on the GPU over the frames drawn alone, orbit
    # frames 0-19: records each wave walked before its pixels saturated, times its pixels).
    evals_upper = float(tile * tile) * P
    tests = alone["raster_tests"] / max(1, alone["frames"])
    raster_alone_ms = alone["stage_ms"][4] / max(1, alone["frames"])
    stages["raster"]["pair_evals_upper"] = int(evals_upper)
    stages["raster"]["coverage_tests"] = int(tests)
    stages["raster"]["Gtests_per_s"] = round(tests / (ms["raster"] * 1e-3) / 1e9, 1) if ms["raster"] > 0 else 0.0
    # The roofline's kernel is the raster, the longest single kernel of the frame in the kernel trace
    # (profiles/k_raster_3M_sh3_1080p_kernel_stats.md); it is bound by VALU issue. Its VALU instructions
    # per launch for these orbit frames come from the counter run; its duration is measured here (HIP
    # events on the frame slot's stream, averaged over the timed frames, where frames overlap and share
    # the GPU; 'alone' over the same frames drawn one at a time). frac = SQ_INSTS_VALU x 4 clocks /
    # (1024 SIMDs x 2.4 GHz x duration): the share of the chip's VALU issue slots the raster fills.
    kernel = f"k_raster<{tile}, {plain_best['raster_px']}>"
    valu = measured_valu(kernel, args.steps)
    traffic = measured_traffic(kernel)

    def valu_rate(t_ms):
        return valu * 64 / (t_ms * 1e-3) if valu is not None and t_ms > 0 else None

    rate, rate_alone = valu_rate(ms["raster"]), valu_rate(raster_alone_ms)
    roofline = {"bound": "valu", "kernel": kernel,
                "achieved": round(rate / 1e12, 3) if rate else None,
                "peak": round(VALU_LANE_OPS_PEAK / 1e12, 2), "unit": "T VALU lane-op/s",
                "frac": round(rate / VALU_LANE_OPS_PEAK, 4) if rate else None,
                "traffic": traffic,
                "valu_instructions_per_launch": round(valu) if valu is not None else None,
                "launch_ms": round(ms["raster"], 4),
                "alone": {"launch_ms": round(raster_alone_ms, 4), "frames": alone["frames"],
                          "frac": round(rate_alone / VALU_LANE_OPS_PEAK, 4) if rate_alone else None},
                "coverage_tests_per_launch": int(tests), "pair_evals_upper_per_launch": int(evals_upper),
                "note": "VALU lane operations (SQ_INSTS_VALU x 64, from the counter run of these orbit frames in "
                        "profiles/sq_counters_3M_sh3_1080p.json) over the raster's average duration measured in this "
                        f"run with {args.frames_in_flight} frames in flight; peak = 1024 SIMDs x 16 lanes x 2.4 GHz "
                        "(packed fp32 disabled). traffic = HBM bytes per launch from the FETCH_SIZE/WRITE_SIZE "
                        "counter runs in profiles/hbm_traffic_3M_sh3_1080p.md. coverage_tests_per_launch: the "
                        "(pixel, record) tests the kernel performed, counted on the GPU over orbit frames 0-19 drawn "
                        "alone (the timed frames do not count them); pair_evals_upper_per_launch = tile^2 x pairs of "
                        "the timed frames"}

    # The sorts against HBM: the depth sort of the frames drawn alone (4 passes over V survivors,
    # B = V (2 (k + 4) 4 + k)), and the standalone sort at the frame's pair count below.
    depth_alone_ms = alone["stage_ms"][2] / max(1, alone["frames"])
    sort_in_frame = {"depth_sort_alone_ms": round(depth_alone_ms, 4),
                     "depth_sort_bytes": int(stage_bytes(N_SPLATS, V, P, T, 4, 0)["depth_sort"]),
                     "hbm_frac": round(stage_bytes(N_SPLATS, V, P, T, 4, 0)["depth_sort"] / (depth_alone_ms * 1e-3)
                                       / HBM_PEAK, 4) if depth_alone_ms > 0 else None}
    # The frame as a whole against both rooflines: every kernel's VALU instructions per frame (the
    # counter run) over the measured time per frame, and the stages' algorithmic bytes over it.
    frame_valu = measured_frame_valu()
    frame_s = elapsed / args.steps
    frame_roofline = {"ms_per_frame": round(frame_s * 1e3, 4),
                      "valu_instructions_per_frame": round(frame_valu) if frame_valu else None,
                      "valu_issue_frac": round(frame_valu * 4 / (1024 * 2.4e9 * frame_s), 4) if frame_valu else None,
                      "algorithmic_bytes_per_frame": int(sum(by.values())),
                      "hbm_frac": round(sum(by.values()) / frame_s / HBM_PEAK, 4),
                      "note": "whole-chip VALU issue slots filled by the frame's kernels (SQ_INSTS_VALU x 4 / (1024 SIMDs x "
                              "2.4 GHz x time per frame)), and the stages' algorithmic bytes per frame over 8 TB/s"}
    sizes = [int(x) for x in args.sort_sizes.split(",") if x]
    runs = [(int(P), tile_bits, "tile sort at the frame's P")] + [(n, 32, "depth-key sort") for n in sizes]
    sort_leg = radix_sort(r, skm, runs) if sizes else None
    # Images for the FP64 pixel check of the CPU leg: orbit frame 0 of this world, drawn alone, and the
    # 100k SH0 scene at 1280x720.
    check_images = None
    if not args.no_cpu:
        r.set_tuning(**plain_best)
        r.render(*frames[0])
        c0_cloud = skm.scene_c0()
        c0 = skm.Renderer(0)
        c0.upload_world(c0_cloud)
        c0.set_extent(1280, 720)
        c0.apply_policy(sort_depth=32)
        c0.render(*skm.camera_c0())
        check_images = {"C1": r.read_frame(), "C0": c0.read_frame()}
        c0.close()
        del c0_cloud
    batched = batched_vs_loop(r, skm, args.batches, 2, plain_best) if args.batches > 0 else None
    r.close()
    lod = (lod_flythrough(skm, args.lod_frames, args.lod_leaves, args.frames_in_flight) if args.lod_leaves > 0
           else None)
    served = None
    if args.serving_requests > 0:
        served = serving(skm, [float(x) for x in args.serving_rates.split(",")], args.serving_requests, d2h=d2h)

    cpu = None
    pixel_err = None
    if not args.no_cpu:
        pixel_err = pixel_error_fp64(check_images)
        cpu_fps, cpu_stages = cpu_baseline(args.cpu_frames, 1)
        threads = min(16, os.cpu_count() or 1)  # the host share of one GPU
        mt_fps, mt_stages = cpu_baseline(args.cpu_frames, threads)
        cpu = {"value": round(cpu_fps, 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_frames} orbit frames of the 3M SH3 1920x1080 workload, whole frame "
                         "(cull, stable sort, bin, fp32 raster) on one host thread",
               "stage_ms": dict(zip(["cull", "sort", "bin", "raster"], cpu_stages)),
               "threaded": {"value": round(mt_fps, 4), "cores": threads,
                            "note": "cull and raster over std::thread workers; sort and bin on one thread",
                            "stage_ms": dict(zip(["cull", "sort", "bin", "raster"], mt_stages))},
               "host_threads_available": os.cpu_count()}

    out = {
        "metric": "frames/s at 3M splats SH3 1080p", "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1000, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 3M-Gaussian SH3 scene: 70% near a ground square and boxes, 30% iid N(0,5^2)), "
                "in the loader's Morton order (reorderSpatially, applied on the GPU at load)",
        "config": {"workload": "plain single-view render, 3M Gaussians SH3, 1920x1080, 60-frame orbit, 32-bit keys",
                   "frames_timed": (f"orbit frames 0-{min(args.steps, ORBIT_FRAMES) - 1}" +
                                    (f", then frame i mod 60 for steps {ORBIT_FRAMES}-{args.steps - 1}"
                                     if args.steps > ORBIT_FRAMES else "")),
                   "splats": N_SPLATS, "sh_degree": SH_DEGREE, "width": WIDTH, "height": HEIGHT,
                   "tile": plain_best["tile_size"], "raster_px": plain_best["raster_px"],
                   "sort_kernel": ["reduce_then_scan", "onesweep_lookback"][plain_best["sort_kernel"]],
                   "tile_radix_bits": plain_best["radix_bits"], "frames_in_flight": args.frames_in_flight,
                   "parallelism": "single GPU"},
        "autotune_ms": [{"tile": t, "raster_px": p, "sort_kernel": k, "radix_bits": rb, "ms_per_draw": round(ms, 4)}
                        for t, p, k, rb, ms in plain_table],
        "msplats_per_s": round(N_SPLATS * fps / 1e6, 1),
        "visible_per_frame": int(V), "pairs_per_frame": int(P),
        "gpu_ms_per_frame": round(tot["stage_ms"][5] / n_frames, 4),
        "stages": stages, "roofline": roofline, "frame_roofline": frame_roofline, "pixel_err_fp64": pixel_err,
        "depth_sort_in_frame": sort_in_frame, "one_slot": one_slot,
        "cpu_baseline": cpu, "radix_sort": sort_leg, "batched": batched,
        "serving": served, "lod": lod,
        "device": device, "scene_generation_s": round(gen_s, 2), "load_reorder_s": round(reorder_s, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
