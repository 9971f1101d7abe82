"""The oracle's LOD cut (orc_lod_select) against the expectations of the reference's own selection
tests (MetalLodSelectionTest.mm) and against invariants: every leaf covered exactly once, the
budget respected, determinism, and the closed-form extremes (refine everything / nothing)."""
import numpy as np
import pytest

import lod_cases as lc


def select(oracle, tree, distance=0.0, capacity=None, pixel=1.0, cw=4.0, cull=True):
    view, proj, eye = lc.camera(distance)
    cam = oracle.make_camera(view, proj, lc.W, lc.H, eye)
    cap = min(capacity or tree.leaf_count, tree.leaf_count)
    return oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, cap, pixel, cw, cull)


@pytest.mark.parametrize("budget", [1, 7, 31, 32, 33, 64, 257, 400, 515])
def test_budget_cuts_cover_every_leaf_once(oracle, skm_mod, budget):
    """MetalLodSelectionTest.mm:80-102."""
    tree = lc.hierarchy(skm_mod)
    cut, stats = select(oracle, tree, capacity=budget, pixel=0.0)
    assert 0 < len(cut) <= budget and stats[0] == len(cut)
    assert np.array_equal(cut, select(oracle, tree, capacity=budget, pixel=0.0)[0])
    assert lc.covers_every_leaf_once(tree, cut)
    if budget == 515:
        assert len(cut) == 515


def test_distance_changes_the_cut(oracle, skm_mod):
    """MetalLodSelectionTest.mm:104-134."""
    tree = lc.hierarchy(skm_mod)
    assert len(select(oracle, tree, capacity=400, pixel=1.0)[0]) > 1
    assert len(select(oracle, tree, distance=1_000_000, capacity=400, pixel=1.0)[0]) == 1


def test_quality_refines_colour_variation(oracle, skm_mod):
    """MetalLodSelectionTest.mm:249-275: cut {1, 5, 6}, no denials, 3 clusters evaluated."""
    cut, stats = select(oracle, lc.quality_tree(skm_mod), distance=100, capacity=4, pixel=1.0, cull=False)
    assert set(cut.tolist()) == {1, 5, 6}
    assert stats[4] == 0 and stats[5] == 3


def test_packet_expansion_and_capacity_pressure(oracle, skm_mod):
    """MetalLodSelectionTest.mm:277-291."""
    tree = lc.hierarchy(skm_mod)
    cut, stats = select(oracle, tree, capacity=515, pixel=0.0, cull=False)
    assert len(cut) == 515 and stats[4] == 0 and stats[5] == len(tree.clusters)
    cut, stats = select(oracle, tree, capacity=33, pixel=0.0, cull=False)
    assert len(cut) <= 33 and stats[4] > 0


def test_empty_view_and_single_leaf_root(oracle, skm_mod):
    """MetalLodSelectionTest.mm:293-307."""
    assert len(select(oracle, lc.hierarchy(skm_mod), distance=-100, capacity=515, pixel=0.0)[0]) == 0
    import cases

    one = cases.iso_cloud([(0, 0, -2)], colors=np.array([[1, 1, 1]], np.float32), alphas=[0.5])
    one["covariances"][0, [0, 3, 5]] = 0.001
    tree = skm_mod.LodTree.build(one)
    assert list(select(oracle, tree, capacity=1)[0]) == [0]


@pytest.mark.parametrize("capacity", [20000, 32768])
def test_binary_tree_scan(oracle, skm_mod, capacity):
    """MetalLodSelectionTest.mm:309-344."""
    tree = lc.binary_tree(skm_mod)
    cut, stats = select(oracle, tree, capacity=capacity, pixel=0.0, cull=False)
    assert len(cut) == capacity and lc.covers_every_leaf_once(tree, cut)
    if capacity == 32768:
        assert stats[4] == 0 and stats[5] == 32767
    else:
        assert stats[4] > 0


def test_each_round_emits_after_the_rounds_before(oracle, skm_mod):
    """MetalLodSelectionTest.mm:346-366."""
    cut, stats = select(oracle, lc.rounds_tree(skm_mod), capacity=3, pixel=0.0, cull=False)
    assert list(cut) == [1, 3, 4] and stats[5] == 2


def test_large_leaf_packet(oracle, skm_mod):
    """MetalLodSelectionTest.mm:368-393."""
    tree = lc.packet_tree(skm_mod)
    cut, stats = select(oracle, tree, capacity=513, pixel=0.0)
    assert list(cut) == list(range(1, 514)) and stats[5] == 1 and stats[4] == 0
    cut, stats = select(oracle, tree, capacity=512, pixel=0.0)
    assert list(cut) == [0] and stats[4] == 1


def test_error_metric_closed_form(oracle, skm_mod):
    """lodErrorPixels (SplatLOD.metal:33-43) for a cluster on the axis: (e + r min(sqrt(v), 1)) f /
    (z - r) * max(opacity, 0.1) * (1 + w sqrt(v))."""
    import ctypes as C

    view, proj, eye = lc.camera(0.0)
    cam = oracle.make_camera(view, proj, lc.W, lc.H, eye)
    for e, r, v, op, z in ((0.1, 0.5, 0.04, 0.7, -10.0), (0.0, 2.0, 4.0, 0.05, -50.0)):
        c = np.zeros(16, np.uint32)
        c[:4] = np.array([0, 0, z, r], np.float32).view(np.uint32)
        c[7] = np.float32(e).view(np.uint32)
        c[8] = np.float32(v).view(np.uint32)
        c[9] = np.float32(op).view(np.uint32)
        err = C.c_float()
        outside = oracle.lib().orc_lod_cluster(C.byref(cam), oracle.up(c), 4.0, C.byref(err))
        expect = (e + r * min(v ** 0.5, 1)) * 500 / (-z - r) * max(op, 0.1) * (1 + 4 * v ** 0.5)
        assert outside == 0 and abs(err.value - expect) <= 1e-5 * expect


def test_thresholds_reduce_the_cut_monotonically(oracle, skm_mod):
    """A larger pixel threshold never selects more nodes; refining everything gives every leaf."""
    cloud = skm_mod.make_cloud(30_000, 1, seed=77, layout=1, spread=5.0, area=20.0)
    tree = skm_mod.LodTree.build(cloud, 21, 32, threads=4)
    view = skm_mod.look_at((0, 2, 12))
    proj = skm_mod.perspective(np.pi / 3, 16 / 9, 0.2, 500)
    cam = oracle.make_camera(view, proj, 1920, 1080, (0, 2, 12))
    sizes = [len(oracle.lod_select(cam, tree.clusters, tree.leaves, tree.rounds, 30_000, p, 4.0, False)[0])
             for p in (0.0, 0.5, 1.5, 4.0, 16.0, 1e30)]
    assert sizes[0] == 30_000 and sizes[-1] == 1
    assert all(a >= b for a, b in zip(sizes, sizes[1:]))
