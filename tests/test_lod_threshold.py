"""LodThreshold (packages/splatkit-engine/include/splatkit/rendering/LodThreshold.h:9-45), the
per-frame adaptation of a tree world's selection threshold: the reference's own known answers
(packages/splatkit-engine/tests/rendering/LodThresholdTest.cpp:12-57) against the oracle's
restatement and against the renderer's rule (skm_lod_threshold_update in libskm.so, a host
function: no GPU is touched), and both against the reference's header compiled from source on
random sequences."""
import ctypes as C

import numpy as np
import pytest

LIMIT = 1000


class Rules:
    """The renderer's rule through the C ABI and the oracle's, stepped side by side."""

    def __init__(self, skm_mod, oracle, policy):
        self.lib = skm_mod.lib()
        self.t = skm_mod.LodThreshold()
        self.lib.skm_lod_threshold_reset(C.byref(self.t), policy)
        self.oracle = oracle
        self.policy = policy
        self.steps = []

    def update(self, selected, limited, limit=LIMIT):
        got = self.lib.skm_lod_threshold_update(C.byref(self.t), selected, limited, limit)
        self.steps.append((selected, limited, limit))
        want = self.oracle.lod_threshold(self.policy, self.steps)[-1]
        assert np.float32(got) == want, (self.steps, got, want)
        return got

    @property
    def pixels(self):
        return self.t.pixels

    def reset(self, policy):
        self.lib.skm_lod_threshold_reset(C.byref(self.t), policy)
        self.policy = policy
        self.steps = []


def feq(a, b):
    """EXPECT_FLOAT_EQ: within 4 ulp."""
    a, b = np.float32(a), np.float32(b)
    return abs(int(a.view(np.int32)) - int(b.view(np.int32))) <= 4


def test_starts_at_the_policy_threshold(skm_mod, oracle):
    """LodThresholdTest.cpp:12-15."""
    assert Rules(skm_mod, oracle, 0.5).pixels == np.float32(0.5)


def test_denied_refinements_raise_it_ten_percent_a_readback(skm_mod, oracle):
    """LodThresholdTest.cpp:17-21."""
    t = Rules(skm_mod, oracle, 1.0)
    assert feq(t.update(LIMIT, 3), 1.1)
    assert feq(t.update(LIMIT, 1), 1.21)


def test_rises_no_higher_than_the_ceiling_or_the_policy(skm_mod, oracle):
    """LodThresholdTest.cpp:23-30."""
    t = Rules(skm_mod, oracle, 1.0)
    for _ in range(100):
        t.update(LIMIT, 1)
    assert feq(t.pixels, 64.0)
    coarse = Rules(skm_mod, oracle, 100.0)
    coarse.update(LIMIT, 1)
    assert feq(coarse.pixels, 100.0)


def test_relaxes_toward_the_policy_once_the_cut_sits_under_ninety_percent(skm_mod, oracle):
    """LodThresholdTest.cpp:32-39."""
    t = Rules(skm_mod, oracle, 1.0)
    t.update(LIMIT, 1)
    t.update(LIMIT, 1)
    assert feq(t.update(LIMIT // 2, 0), np.float32(1.21) / np.float32(1.03))
    for _ in range(100):
        t.update(LIMIT // 2, 0)
    assert feq(t.pixels, 1.0)


def test_holds_inside_the_band_below_the_limit(skm_mod, oracle):
    """LodThresholdTest.cpp:41-46: at 900 of 1000 (limit / 10 * 9) and at the limit it holds."""
    t = Rules(skm_mod, oracle, 1.0)
    t.update(LIMIT, 1)
    assert feq(t.update(LIMIT // 10 * 9, 0), 1.1)
    assert feq(t.update(LIMIT, 0), 1.1)
    # One below the band relaxes; the band uses integer division (limit 1009 -> 900).
    assert t.update(LIMIT // 10 * 9 - 1, 0) < np.float32(1.1)
    u = Rules(skm_mod, oracle, 1.0)
    u.update(1009, 1, 1009)
    assert feq(u.update(900, 0, 1009), 1.1)
    assert u.update(899, 0, 1009) < np.float32(1.1)


def test_reset_returns_to_the_new_policy_at_once(skm_mod, oracle):
    """LodThresholdTest.cpp:48-54."""
    t = Rules(skm_mod, oracle, 1.0)
    t.update(LIMIT, 1)
    t.reset(0.25)
    assert t.pixels == np.float32(0.25) and t.t.policy_pixels == np.float32(0.25)


@pytest.mark.parametrize("policy,seed", [(1.0, 1), (0.1, 2), (16.0, 3), (100.0, 4), (1.5, 5)])
def test_random_sequences_match_the_reference_header(skm_mod, oracle, ref, policy, seed):
    """The oracle and the renderer's rule against LodThreshold.h compiled from the reference tree,
    bit for bit, over random runs of denials, full cuts and sparse cuts around the band."""
    rng = np.random.default_rng(seed)
    steps = []
    for _ in range(400):
        limit = int(rng.integers(1, 5_000_000))
        kind = rng.integers(3)
        limited = int(rng.integers(1, 1000)) if kind == 0 else 0
        selected = int(limit if kind == 1 else rng.integers(0, limit + 1))
        if rng.uniform() < 0.2:
            selected = limit // 10 * 9 - int(rng.integers(0, 2))
        steps.append((max(selected, 0), limited, limit))
    want = oracle.lod_threshold(policy, steps, f=ref.ref_lod_threshold)
    assert np.array_equal(oracle.lod_threshold(policy, steps), want)
    t = Rules(skm_mod, oracle, policy)
    got = np.array([t.update(*s) for s in steps], np.float32)
    assert np.array_equal(got, want)
