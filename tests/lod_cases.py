"""LOD trees and cameras restated from the reference's selection tests
(packages/splatkit-ios/tests/rendering/metal/selection/MetalLodSelectionTest.mm). The camera there has
focal 500 px and tan(half FOV) 1 over 1000 px; here the projection's 90 degree FOV gives exactly
those values."""
import numpy as np

import cases

W = H = 1000
PROJ = cases.perspective(np.pi / 2, 1.0, 0.1, 1_000_000.0)


def camera(distance=0.0):
    """camera(distance) of MetalLodSelectionTest.mm:40-54: view translated to z = -distance."""
    view = cases.identity()
    view[14] = -distance
    return view, PROJ, np.array([0, 0, distance], np.float32)


def hierarchy(skm):
    """hierarchy() of MetalLodSelectionTest.mm:22-39: a 17x17x2 grid of 515 splats, depth 6."""
    i = np.arange(515)
    pos = np.stack([(i % 17) / 8.0 - 1, ((i // 17) % 17) / 8.0 - 1, -2.0 - (i // 289) * 0.25], 1).astype(np.float32)
    cloud = cases.iso_cloud(pos, sigmas=np.sqrt(np.float32(0.001)), colors=np.array([[1, 0, 0]] * 515, np.float32),
                            alphas=[0.5] * 515)
    cloud["covariances"][:, [0, 3, 5]] = 0.001
    return skm.LodTree.build(cloud, depth=6, cluster_leaves=32, threads=1)


def from_layout(skm, rows, leaf_count, cov=0.01, colour=(1, 1, 1), alpha=0.5):
    """A tree from (position, size, child_start, child_count) rows, nodes at the rows' positions."""
    n = len(rows)
    pos = np.array([r[0] for r in rows], np.float32)
    nodes = cases.iso_cloud(pos, colors=np.array([colour] * n, np.float32), alphas=[alpha] * n)
    nodes["covariances"][:] = 0
    nodes["covariances"][:, [0, 3, 5]] = cov
    layout = np.zeros((n, 6), np.uint32)
    for k, (p, size, cs, cc) in enumerate(rows):
        layout[k, :3] = np.asarray(p, np.float32).view(np.uint32)
        layout[k, 3] = np.float32(size).view(np.uint32)
        layout[k, 4], layout[k, 5] = cs, cc
    return nodes, layout, leaf_count


def quality_tree(skm):
    """MetalLodSelectionTest.mm:249-275: uniform pair under node 1, colour-varying pair under 2."""
    rows = [((0, 0, -2), 2, 1, 2), ((-0.7, 0, -2), 0.1, 3, 2), ((0.7, 0, -2), 1.8, 5, 2),
            ((-0.8, 0, -2), 0.1, 0, 0), ((-0.6, 0, -2), 0.1, 0, 0), ((0.6, 0, -2), 0.1, 0, 0),
            ((0.8, 0, -2), 0.1, 0, 0)]
    nodes, layout, leaves = from_layout(skm, rows, 4)
    nodes["colors"][5, 0] = 0
    nodes["colors"][6, 1] = 0
    return skm.LodTree.from_layout(nodes, layout, leaves)


def tiny_rows(count, children):
    return [((0, 0, -2), 0.02, *children(i)) for i in range(count)]


def rounds_tree(skm):
    """MetalLodSelectionTest.mm:346-366: root 0 holds leaf 1 and interior 2 (leaves 3, 4)."""
    rows = tiny_rows(5, lambda i: (1, 2) if i == 0 else (3, 2) if i == 2 else (0, 0))
    return skm.LodTree.from_layout(*from_layout(skm, rows, 3, cov=0.0001, colour=(0.5, 0.5, 0.5), alpha=0.1))


def packet_tree(skm):
    """MetalLodSelectionTest.mm:368-393: one root over 513 leaves."""
    rows = tiny_rows(514, lambda i: (1, 513) if i == 0 else (0, 0))
    return skm.LodTree.from_layout(*from_layout(skm, rows, 513, cov=0.0001, colour=(0.5, 0.5, 0.5), alpha=0.1))


def binary_tree(skm, count=32768):
    """MetalLodSelectionTest.mm:309-344: a complete binary tree of `count` leaves."""
    rows = tiny_rows(2 * count - 1, lambda i: (2 * i + 1, 2) if i < count - 1 else (0, 0))
    return skm.LodTree.from_layout(*from_layout(skm, rows, count, cov=0.0001, colour=(0.5, 0.5, 0.5), alpha=0.1))


def leaves_under(tree, node, out):
    lay = tree.layout
    stack = [node]
    while stack:
        k = stack.pop()
        if lay[k, 5] == 0:
            out.append(k)
        else:
            stack.extend(range(lay[k, 4], lay[k, 4] + lay[k, 5]))
    return out


def covers_every_leaf_once(tree, cut):
    rep = []
    for node in cut:
        leaves_under(tree, int(node), rep)
    return len(rep) == tree.leaf_count and len(set(rep)) == tree.leaf_count
