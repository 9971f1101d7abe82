"""The LOD builder (include/skm_lod.h) against the reference's own buildLodTree and
buildLodSelectionData compiled from its sources (oracle/_ref), and against the invariants the
reference's tests state (LodTreeTest.cpp, MetalLodSelectionTest.mm)."""
import ctypes as C

import numpy as np
import pytest


def ref_tree(oracle, ref, cloud, depth, cluster_leaves):
    n = len(cloud["alphas"])
    deg = cloud["sh_degree"] if cloud.get("sh") is not None else 0
    ref.ref_lod_build.restype = C.c_void_p
    ref.ref_lod_build.argtypes = [C.c_uint32, C.c_int] + [C.POINTER(C.c_float)] * 5 + [C.c_uint32, C.c_uint32]
    h = ref.ref_lod_build(n, deg, oracle.fp(cloud["positions"]), oracle.fp(cloud["covariances"]),
                          oracle.fp(cloud["colors"]), oracle.fp(cloud["alphas"]),
                          oracle.fp(cloud["sh"]) if deg else None, depth, cluster_leaves)
    counts = np.zeros(4, np.uint64)
    ref.ref_lod_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    ref.ref_lod_counts(h, counts.ctypes.data_as(C.POINTER(C.c_uint64)))
    nodes, leaves, clusters, refs = (int(x) for x in counts)
    k = 3 * ((deg + 1) ** 2 - 1)
    out = {"positions": np.zeros((nodes, 3), np.float32), "covariances": np.zeros((nodes, 6), np.float32),
           "colors": np.zeros((nodes, 3), np.float32), "alphas": np.zeros(nodes, np.float32),
           "sh": np.zeros((nodes, max(k, 1)), np.float32), "layout": np.zeros((nodes, 6), np.uint32),
           "clusters": np.zeros((clusters, 16), np.uint32), "leaves": np.zeros(max(refs, 1), np.uint32)}
    ref.ref_lod_read.argtypes = [C.c_void_p] + [C.POINTER(C.c_float)] * 5 + [C.POINTER(C.c_uint32)] * 3
    ref.ref_lod_read(h, oracle.fp(out["positions"]), oracle.fp(out["covariances"]), oracle.fp(out["colors"]),
                     oracle.fp(out["alphas"]), oracle.fp(out["sh"]) if deg else None, oracle.up(out["layout"]),
                     oracle.up(out["clusters"]), oracle.up(out["leaves"]))
    ref.ref_lod_free.argtypes = [C.c_void_p]
    ref.ref_lod_free(h)
    out["leaves"] = out["leaves"][:refs]
    out["leaf_count"] = leaves
    return out


@pytest.mark.parametrize("n,deg,depth,leaves,layout", [
    (3000, 0, 21, 32, 0), (5000, 1, 8, 8, 1), (2000, 3, 6, 2, 0), (20000, 1, 21, 256, 1), (1, 0, 21, 32, 0),
    (7, 2, 21, 32, 0)])
def test_builder_matches_reference_builder(skm_mod, oracle, ref, n, deg, depth, leaves, layout):
    cloud = skm_mod.make_cloud(n, deg, seed=40 + n, layout=layout, spread=3.0, area=20.0)
    mine = skm_mod.LodTree.build(cloud, depth, leaves, threads=4)
    theirs = ref_tree(oracle, ref, cloud, depth, leaves)
    assert mine.node_count == len(theirs["alphas"]) and mine.leaf_count == theirs["leaf_count"] == n
    assert np.array_equal(mine.layout, theirs["layout"])
    for k in ("positions", "covariances", "colors", "alphas"):
        assert np.array_equal(mine.nodes[k], theirs[k]), k
    if deg:
        assert np.array_equal(mine.nodes["sh"], theirs["sh"])
    assert np.array_equal(mine.clusters, theirs["clusters"])
    assert np.array_equal(mine.leaves, theirs["leaves"])


def test_builder_invariants(skm_mod):
    cloud = skm_mod.make_cloud(50_000, 1, seed=9, layout=1, spread=5.0, area=40.0)
    t = skm_mod.LodTree.build(cloud, 21, 32, threads=8)
    lay = t.layout
    # Every leaf of the cloud appears once as a childless node; interiors merge >= 2 members.
    assert (lay[:, 5] == 0).sum() == 50_000
    assert np.all(lay[lay[:, 5] > 0, 5] >= 2)
    # Breadth-first: children follow their parent and are contiguous.
    idx = np.nonzero(lay[:, 5])[0]
    assert np.all(lay[idx, 4] > idx)
    # Merged opacity stays within [0, 1000] (LodTree.cpp:91); clusters cover every leaf once.
    assert t.nodes["alphas"].max() <= 1000.0
    assert t.clusters[0, 15] == 50_000  # subtree leaves of the root cluster
    assert len(np.unique(t.leaves)) == len(t.leaves) == 50_000
    # Leaves keep their original attributes (as a set).
    leaves = lay[:, 5] == 0
    a = np.sort(t.nodes["positions"][leaves].view(np.uint32).reshape(-1, 3), axis=0)
    b = np.sort(cloud["positions"].view(np.uint32).reshape(-1, 3), axis=0)
    assert np.array_equal(a, b)
    assert t.rounds >= 2


def test_moment_matching_of_two_splats(skm_mod):
    """LodTreeTest: a merge of two equal splats keeps their mean and adds the spread of the means
    to the within covariance (cov = E[cov] + Var(means))."""
    cloud = {"positions": np.array([[-1, 0, 0], [1, 0, 0]], np.float32),
             "covariances": np.array([[0.01, 0, 0, 0.02, 0, 0.03]] * 2, np.float32),
             "colors": np.array([[1, 0, 0], [0, 0, 1]], np.float32), "alphas": np.array([0.5, 0.5], np.float32),
             "sh_degree": 0, "sh": None}
    t = skm_mod.LodTree.build(cloud, 21, 32, threads=1)
    root = 0
    assert t.layout[root, 5] == 2
    np.testing.assert_allclose(t.nodes["positions"][root], [0, 0, 0], atol=1e-7)
    np.testing.assert_allclose(t.nodes["covariances"][root], [1.01, 0, 0, 0.02, 0, 0.03], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(t.nodes["colors"][root], [0.5, 0, 0.5], rtol=1e-6)
