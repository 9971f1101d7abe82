"""The serving loop on one MI355X (include/skm_server.h): resident scenes, continuous batching
under an HBM budget. A request's BGRA8 pixels are the same alone and in a mix of scenes and
resolutions, and the same as the renderer's single draw of that view."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scenes(skm_mod):
    return [skm_mod.make_cloud(100_000, 0, seed=61, layout=0, spread=2.0),
            skm_mod.make_cloud(200_000, 1, seed=62, layout=1, spread=3.0, area=6.0),
            skm_mod.make_cloud(300_000, 3, seed=63, layout=1, spread=4.0, area=8.0)]


def requests(skm_mod, count, seed, n_scenes=3):
    views = skm_mod.views_c2(count, seed=seed, radius=(3.0, 9.0),
                             resolutions=((320, 180), (640, 360), (1280, 720), (333, 217)))
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(n_scenes)),) + v for v in views]


def single_bgra(r, skm_mod, cloud, view, proj, eye, w, h):
    r.upload_world(cloud)
    r.apply_policy(sort_depth=16)
    r.set_extent(w, h)
    r.render(view, proj, eye)
    return r.read_frame(skm_mod.PIXELS_BGRA8)


def test_request_alone_and_in_a_mix_give_the_same_pixels(skm_mod, scenes, renderer):
    say("case: 3 scenes, one request alone, then inside a mix of 40")
    srv = skm_mod.Server(max_batch_views=16)
    try:
        for c in scenes:
            srv.add_scene(c)
        probe = (1,) + skm_mod.views_c2(1, seed=71, radius=(3.0, 6.0), resolutions=((640, 360),))[0]
        t = srv.submit(*probe)
        done = srv.wait_all([t])
        assert done[t]["status"] == 0 and done[t]["batch_views"] == 1
        alone = srv.image(t)
        srv.release(t)
        mix = requests(skm_mod, 40, seed=72)
        tickets = [srv.submit(*q, user_data=k) for k, q in enumerate(mix[:20])]
        tp = srv.submit(*probe, user_data=999)
        tickets += [srv.submit(*q, user_data=20 + k) for k, q in enumerate(mix[20:])]
        done = srv.wait_all(tickets + [tp])
        assert all(c["status"] == 0 for c in done.values())
        say(f"probe batched with {done[tp]['batch_views']} views; stats {srv.stats()}")
        assert np.array_equal(srv.image(tp), alone)
        assert max(c["batch_views"] for c in done.values()) > 1  # requests were batched
        # Every request equals the renderer's single draw of that view (a sample).
        for k in (0, 7, 19, 33):
            s, v, p, e, w, h = mix[k]
            want = single_bgra(renderer, skm_mod, scenes[s], v, p, e, w, h)
            tk = tickets[k] if k < 20 else tickets[k]
            assert np.array_equal(srv.image(tk), want), k
        assert np.array_equal(alone, single_bgra(renderer, skm_mod, scenes[1], *probe[1:]))
    finally:
        srv.close()


def test_one_request_per_launch_mode(skm_mod, scenes):
    say("case: batching off, one launch in flight")
    srv = skm_mod.Server(max_batch_views=1, max_in_flight=1)
    try:
        for c in scenes:
            srv.add_scene(c)
        qs = requests(skm_mod, 12, seed=73)
        tickets = [srv.submit(*q) for q in qs]
        done = srv.wait_all(tickets)
        st = srv.stats()
        assert st["launches"] == 12 and all(c["batch_views"] == 1 for c in done.values())
        assert st["completed"] == 12 and st["p99_ms"] >= st["p50_ms"] > 0
    finally:
        srv.close()


def test_budget_admission_and_rejection(skm_mod, scenes):
    say("case: a budget that holds the scenes and small launches only")
    resident = sum(c["positions"].shape[0] * (32 + 4 * {0: 0, 1: 5, 2: 12, 3: 23}[c["sh_degree"]]) for c in scenes)
    budget = resident + 200 * 1024 * 1024
    srv = skm_mod.Server(max_batch_views=64, budget_bytes=budget)
    try:
        for c in scenes:
            srv.add_scene(c)
        big = skm_mod.views_c2(1, seed=74, resolutions=((7680, 4320),))[0]
        tb = srv.submit(2, *big)
        small = requests(skm_mod, 16, seed=75)
        ts = [srv.submit(*q) for q in small]
        done = srv.wait_all(ts + [tb])
        st = srv.stats()
        say(f"stats {st}")
        assert done[tb]["status"] == -3  # SKM_ERR_MEMORY: cannot fit even alone
        assert all(done[t]["status"] == 0 for t in ts)
        assert st["resident_bytes"] + st["scratch_bytes"] <= budget
        assert st["rejected"] == 1
    finally:
        srv.close()


def test_poisson_stream_reports_latency(skm_mod, scenes):
    say("case: a short Poisson stream at 200 requests/s")
    srv = skm_mod.Server(max_batch_views=32)
    try:
        for c in scenes:
            srv.add_scene(c)
        qs = requests(skm_mod, 60, seed=76)
        rng = np.random.default_rng(77)
        t0 = time.perf_counter()
        arrival = np.cumsum(rng.exponential(1 / 200.0, len(qs)))
        tickets = []
        for q, at in zip(qs, arrival):
            while time.perf_counter() - t0 < at:
                pass
            tickets.append(srv.submit(*q))
        srv.wait_all(tickets)
        st = srv.stats()
        say(f"stats {st}")
        assert st["completed"] == 60 and 0 < st["p50_ms"] <= st["p99_ms"]
    finally:
        srv.close()


def test_launches_stay_within_the_pixel_cap(skm_mod, scenes):
    """A launch takes at most 2^25 pixels past its first request: 3840x2160 requests (8.3M pixels)
    go at most four to a launch even when 64 are allowed and twelve are waiting."""
    say("case: 12 requests at 3840x2160 for one scene, max_batch_views 64")
    srv = skm_mod.Server(max_batch_views=64, max_in_flight=1)
    try:
        srv.add_scene(scenes[0])
        views = skm_mod.views_c2(12, seed=81, radius=(3.0, 6.0), resolutions=((3840, 2160),))
        tickets = [srv.submit(0, *v) for v in views]
        done = srv.wait_all(tickets)
        assert all(c["status"] == 0 for c in done.values())
        sizes = [c["batch_views"] for c in done.values()]
        say(f"launch sizes {sorted(set(sizes))}")
        assert max(sizes) <= 4
        for t in tickets:
            srv.release(t)
    finally:
        srv.close()


def served_vs_oracle(img, o, label):
    """A served BGRA8 image against the oracle's FP64 composite: each channel within 1/255 of FP64
    (the bar of the single draw) plus the half step of the 8-bit rounding."""
    want = np.clip(o["image"][..., [2, 1, 0]], 0.0, 1.0) * 255.0
    err = np.abs(img[..., :3].astype(np.float64) - want).max()
    say(f"{label}: visible {o['count']} pairs {o['pairs']} max |served - 255 fp64| = {err:.3f}")
    assert err <= 1.5, label
    assert np.all(img[..., 3] == 255), label


def test_a_dropped_launch_is_relaunched_as_the_same_group(skm_mod):
    """Requests of one scene (1M splats SH1, 720p to 2160p views) arriving one by one while launches
    run: a launch whose tile pairs overflow the scene's pair buffers is dropped and its requests are
    launched again as exactly that group, which then fits (requests that arrived meanwhile wait for
    the next launch, whose overflow is handled the same way). Every request completes, each with the
    pixels of the same request served alone."""
    say("case: 1M SH1 scene, 5 requests 720p-2160p trickled in, fresh server x3")
    cloud = skm_mod.make_cloud(1_000_000, 1, seed=0x5EED0005, layout=1, spread=5.0, area=10.0)
    rng = np.random.default_rng(1003)
    views = skm_mod.views_c2(32, seed=2003, resolutions=((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)))
    picks = [v for v, s in zip(views, rng.integers(0, 8, 32)) if s == 1]
    assert len(picks) >= 4
    alone = []
    srv = skm_mod.Server(max_batch_views=1, max_in_flight=1)
    try:
        srv.add_scene(cloud)
        for v in picks:
            t = srv.submit(0, *v)
            assert srv.wait_all([t], timeout_s=120)[t]["status"] == 0
            alone.append(srv.image(t))
            srv.release(t)
    finally:
        srv.close()
    for gap_s in (0.0, 2e-4, 1e-3):
        srv = skm_mod.Server(max_batch_views=4, max_in_flight=3)
        try:
            srv.add_scene(cloud)
            tickets = []
            for v in picks:
                tickets.append(srv.submit(0, *v))
                time.sleep(gap_s)
            done = srv.wait_all(tickets, timeout_s=120)
            st = srv.stats()
            say(f"gap {gap_s * 1e3:.1f} ms: launches {st['launches']} rejected {st['rejected']} "
                f"statuses {[done[t]['status'] for t in tickets]}")
            assert all(done[t]["status"] == 0 for t in tickets) and st["rejected"] == 0
            for t, a in zip(tickets, alone):
                assert np.array_equal(srv.image(t), a)
                srv.release(t)
        finally:
            srv.close()


def test_served_requests_match_oracle_across_sh_and_resolutions(skm_mod, oracle):
    """Four resident scenes at SH0, SH1, SH2 and SH3, one request per scene and resolution (720p,
    1080p, 2160p), all queued together so they are batched: every served image matches the oracle's
    frame of that scene, camera and resolution (16-bit keys, the server's default)."""
    say("case: 4 scenes SH0-3 x 720p/1080p/2160p against the oracle")
    clouds = [skm_mod.make_cloud(n, deg, seed=400 + deg, layout=1, spread=3.0, area=6.0)
              for deg, n in ((0, 120_000), (1, 150_000), (2, 180_000), (3, 200_000))]
    resolutions = ((1280, 720), (1920, 1080), (3840, 2160))
    srv = skm_mod.Server(max_batch_views=8)
    try:
        for c in clouds:
            srv.add_scene(c)
        reqs = []
        for s in range(4):
            views = skm_mod.views_c2(len(resolutions), seed=500 + s, radius=(3.0, 9.0), resolutions=resolutions)
            for k, (view, proj, eye, _, _) in enumerate(views):
                w, h = resolutions[k]
                proj = skm_mod.perspective(np.pi / 3, w / h, 0.2, 100.0)
                reqs.append((s, view, proj, eye, w, h))
        tickets = [srv.submit(*q) for q in reqs]
        done = srv.wait_all(tickets)
        assert all(c["status"] == 0 for c in done.values())
        say(f"launch sizes {sorted(c['batch_views'] for c in done.values())}")
        for t, (s, view, proj, eye, w, h) in zip(tickets, reqs):
            o = oracle.frame(clouds[s], view, proj, eye, w, h, key16=True)
            served_vs_oracle(srv.image(t), o, f"scene SH{s} {w}x{h}")
            srv.release(t)
    finally:
        srv.close()


@pytest.mark.timeout(600)
def test_c4_scale_alone_equals_in_the_mix(skm_mod, oracle):
    """The serving workload's scale: eight resident scenes of 0.5, 1, 2, 3, 5, 6, 8 and 10M splats at
    SH0-3 cyclic. One request per scene alone, then the same requests inside a Poisson-free burst of
    48 mixed requests (720p to 2160p) with batching and several launches in flight: each result is
    bit-identical alone and in the mix, and the two smallest scenes' match the oracle."""
    say("case: 8 scenes of 0.5M-10M splats, probes alone and in a 48-request mix")
    sizes = (500_000, 1_000_000, 2_000_000, 3_000_000, 5_000_000, 6_000_000, 8_000_000, 10_000_000)
    srv = skm_mod.Server(max_batch_views=4, max_in_flight=3)
    clouds = []
    try:
        for k, n in enumerate(sizes):
            c = skm_mod.make_cloud(n, k % 4, seed=0x5EED0004 + k, layout=1, spread=5.0, area=10.0)
            srv.add_scene(c)
            clouds.append(c if k < 2 else None)  # the large scenes stay on the device only
            say(f"scene {k}: {n} splats SH{k % 4} resident")
        resolutions = ((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))
        probes = []
        for k in range(8):
            view, proj, eye, w, h = skm_mod.views_c2(1, seed=600 + k, resolutions=(resolutions[k % 4],))[0]
            probes.append((k, view, proj, eye, w, h))
        alone = {}
        for q in probes:
            t = srv.submit(*q)
            c = srv.wait_all([t], timeout_s=300)[t]
            assert c["status"] == 0 and c["batch_views"] == 1
            alone[q[0]] = srv.image(t)
            srv.release(t)
        rng = np.random.default_rng(601)
        mix = [(int(rng.integers(8)),) + v for v in skm_mod.views_c2(40, seed=602, resolutions=resolutions)]
        tagged = ([(None, q) for q in mix[:10]] + [(q[0], q) for q in probes[:4]] + [(None, q) for q in mix[10:30]] +
                  [(q[0], q) for q in probes[4:]] + [(None, q) for q in mix[30:]])
        tickets = [srv.submit(*q) for _, q in tagged]
        done = srv.wait_all(tickets, timeout_s=300)
        assert all(c["status"] == 0 for c in done.values())
        st = srv.stats()
        say(f"mix: launches {st['launches']} mean views per launch {st['mean_batch_views']:.2f}")
        assert st["mean_batch_views"] > 1.0
        for t, (probe, _) in zip(tickets, tagged):
            if probe is not None:
                assert np.array_equal(srv.image(t), alone[probe]), probe
        for k in (0, 1):
            _, view, proj, eye, w, h = probes[k]
            served_vs_oracle(alone[k], oracle.frame(clouds[k], view, proj, eye, w, h, key16=True),
                             f"scene {k} ({sizes[k]} splats) {w}x{h}")
        for t in tickets:
            srv.release(t)
    finally:
        srv.close()
