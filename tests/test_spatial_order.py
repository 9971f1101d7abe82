"""The loader's spatial order (reorderSpatially, packages/splat-core/src/sorting/SpatialOrder.cpp:48-92).

CPU: the oracle's restatement against the reference's own known answers (SpatialOrderTest.cpp) and
against the reference code itself (oracle/_ref). GPU: skm_reorder_spatially against the oracle, bit for
bit, with every attribute permuted."""
import numpy as np
import pytest


def test_morton_code_known_answers(oracle):
    """SpatialOrderTest.cpp:31-43: unit steps on x, y, z give 1, 2, 4; the far corner 2^30 - 1."""
    lo, hi = (0, 0, 0), (1024, 1024, 1024)
    assert oracle.morton_code((1, 0, 0), lo, hi) == 1
    assert oracle.morton_code((0, 1, 0), lo, hi) == 2
    assert oracle.morton_code((0, 0, 1), lo, hi) == 4
    assert oracle.morton_code((1023.5, 1023.5, 1023.5), lo, hi) == 0x3FFFFFFF
    # Below the box clamps to cell 0, above it to 1023; a flat axis quantises to 0.
    assert oracle.morton_code((-5, -5, -5), lo, hi) == 0
    assert oracle.morton_code((5000, 5000, 5000), lo, hi) == 0x3FFFFFFF
    assert oracle.morton_code((7, 700, 7), (0, 5, 0), (1024, 5, 1024)) == oracle.morton_code((7, 0, 7), lo, hi)


def random_positions(n, seed, lo=-5.0, hi=5.0):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)).astype(np.float32)


def test_spatial_order_is_a_stable_sort_by_code(oracle):
    """SpatialOrderTest.cpp:45-82: every splat kept once, codes non-decreasing, ties in input order."""
    lo, hi = np.full(3, -5, np.float32), np.full(3, 5, np.float32)
    for n, seed in ((1000, 7), (1000, 11), (50_000, 3)):
        pos = random_positions(n, seed)
        pos[n // 2:n // 2 + 40] = pos[n // 3]  # a run of equal codes
        order = oracle.spatial_order(pos, lo, hi)
        assert np.array_equal(np.sort(order), np.arange(n))
        codes = np.array([oracle.morton_code(p, lo, hi) for p in pos[order]], np.uint64)
        assert np.all(codes[1:] >= codes[:-1])
        ties = codes[1:] == codes[:-1]
        assert np.all(order[1:][ties] > order[:-1][ties])
    flat = oracle.spatial_order(random_positions(10, 3), lo, lo)
    assert np.array_equal(flat, np.arange(10))  # every code 0: the decode order stays
    assert np.array_equal(oracle.spatial_order(random_positions(1, 5), lo, hi), [0])


def test_spatial_order_matches_reference_code(oracle, ref):
    lo, hi = np.full(3, -5, np.float32), np.full(3, 5, np.float32)
    for n, seed in ((1000, 7), (20_000, 13)):
        pos = random_positions(n, seed)
        assert np.array_equal(oracle.spatial_order(pos, lo, hi), oracle.spatial_order(pos, lo, hi, ref.ref_spatial_order))
        for p in pos[:500]:
            assert oracle.morton_code(p, lo, hi) == oracle.morton_code(p, lo, hi, ref.ref_morton_code)
    pos = random_positions(5000, 21, -40.0, 90.0)
    lo2, hi2 = pos.min(0), pos.max(0)
    assert np.array_equal(oracle.spatial_order(pos, lo2, hi2), oracle.spatial_order(pos, lo2, hi2, ref.ref_spatial_order))


@pytest.mark.gpu
def test_gpu_reorder_matches_oracle(renderer, skm_mod, oracle):
    """skm_reorder_spatially: the oracle's permutation bit for bit, every attribute moved with its splat."""
    for n, deg, seed in ((300_001, 3, 31), (4097, 1, 32), (2, 0, 33)):
        cloud = skm_mod.make_cloud(n, deg, seed=seed, layout=1, spread=5.0, area=10.0)
        before = {k: v.copy() for k, v in cloud.items() if isinstance(v, np.ndarray)}
        lo, hi = before["positions"].min(0), before["positions"].max(0)
        expect = oracle.spatial_order(before["positions"], lo, hi)
        order = renderer.reorder_spatially(cloud)
        assert np.array_equal(order, expect), n
        for k, v in before.items():
            assert np.array_equal(cloud[k], v[expect]), k
    flat = skm_mod.make_cloud(1000, 0, seed=34)
    pos0 = flat["positions"].copy()
    order = renderer.reorder_spatially(flat, lo=(0, 0, 0), hi=(0, 0, 0))
    assert np.array_equal(order, np.arange(1000)) and np.array_equal(flat["positions"], pos0)
    one = skm_mod.make_cloud(1, 0, seed=35)
    assert list(renderer.reorder_spatially(one)) == [0]
