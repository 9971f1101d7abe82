"""The LODSPLAT codec (include/skm_lod.h: skm_lod_read_header / decode / read / validate / write)
against the reference's own LodFile.cpp compiled from its source (oracle/_ref): files written by
either writer are byte-identical, each decoder reads the other's files back to the same arrays, and
both refuse the same corruptions with the same message (LodFileTest.cpp:35-156 restated). The one
intended difference: the reference refuses headers of more than 20,000,000 nodes (LodFile.cpp:24),
this decoder reads up to 2^31."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from test_lod_build import ref_tree


def ref_build_handle(oracle, ref, cloud, depth, cluster_leaves):
    deg = cloud["sh_degree"] if cloud.get("sh") is not None else 0
    return ref.ref_lod_build(len(cloud["alphas"]), deg, oracle.fp(cloud["positions"]),
                             oracle.fp(cloud["covariances"]), oracle.fp(cloud["colors"]),
                             oracle.fp(cloud["alphas"]), oracle.fp(cloud["sh"]) if deg else None, depth,
                             cluster_leaves)


def ref_decode(ref, data, max_sh=3):
    buf = np.frombuffer(bytes(data), np.uint8)
    err = C.create_string_buffer(256)
    h = ref.ref_lod_decode(buf.ctypes.data_as(C.c_void_p), buf.nbytes, max_sh, err, 256)
    return h, err.value.decode()


def mine_decode(skm_mod, data, max_sh=3):
    try:
        return skm_mod.LodTree.decode(bytes(data), max_sh), ""
    except ValueError as e:
        return None, str(e)


def write_both(skm_mod, oracle, ref, tmp_path, cloud, depth, leaves):
    mine = skm_mod.LodTree.build(cloud, depth, leaves, threads=4)
    a = tmp_path / "mine.lodsplat"
    mine.write(str(a))
    h = ref_build_handle(oracle, ref, cloud, depth, leaves)
    b = tmp_path / "ref.lodsplat"
    err = C.create_string_buffer(256)
    assert ref.ref_lod_write(h, str(b).encode(), err, 256) == 1, err.value
    ref.ref_lod_free(h)
    return mine, a.read_bytes(), b.read_bytes()


@pytest.mark.parametrize("n,deg,depth,leaves,layout", [(3000, 0, 21, 32, 0), (5000, 1, 8, 8, 1), (2000, 3, 6, 2, 0),
                                                      (20000, 1, 21, 256, 1), (1, 0, 21, 32, 0), (7, 2, 21, 32, 0)])
def test_files_are_byte_identical_and_cross_decode(skm_mod, oracle, ref, tmp_path, n, deg, depth, leaves, layout):
    cloud = skm_mod.make_cloud(n, deg, seed=70 + n, layout=layout, spread=3.0, area=20.0)
    mine, a, b = write_both(skm_mod, oracle, ref, tmp_path, cloud, depth, leaves)
    assert a == b
    stride = 64 + 4 * 3 * ((deg + 1) ** 2 - 1)
    assert len(a) == 64 + mine.node_count * stride + len(mine.clusters) * 64 + len(mine.leaves) * 4
    # This decoder on the reference's file: the builder's arrays, bit for bit.
    back = skm_mod.LodTree.read(str(tmp_path / "ref.lodsplat"))
    theirs = ref_tree(oracle, ref, cloud, depth, leaves)
    assert np.array_equal(back.layout, theirs["layout"])
    for k in ("positions", "covariances", "colors", "alphas"):
        assert np.array_equal(back.nodes[k], theirs[k]), k
    if deg:
        assert np.array_equal(back.nodes["sh"], theirs["sh"])
    assert np.array_equal(back.clusters, theirs["clusters"]) and np.array_equal(back.leaves, theirs["leaves"])
    assert back.validate() == mine.validate()
    lo, hi = back.bounds()
    assert np.array_equal(lo, cloud["positions"].min(0)) and np.array_equal(hi, cloud["positions"].max(0))
    # The reference decoder on this writer's file, and both decoders' header readings.
    h, why = ref_decode(ref, a)
    assert h, why
    counts = np.zeros(4, np.uint64)
    ref.ref_lod_counts(h, counts.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert list(counts) == [mine.node_count, n, len(mine.clusters), len(mine.leaves)]
    depth_ref = C.c_uint32()
    assert ref.ref_lod_validate(h, C.byref(depth_ref), C.create_string_buffer(256), 256) == 1
    ref.ref_lod_free(h)
    hdr, why = skm_mod.read_lod_header(a)
    assert hdr and hdr["version"] == 2 and hdr["node_count"] == mine.node_count and hdr["depth"] == depth_ref.value
    back.close()
    mine.close()


def fixture_tree(skm_mod):
    """LodFileTest.cpp:11-24: two splats 2 m apart, SH1 of 0.25, an octree of depth 6."""
    cloud = {"positions": np.array([[-1, 0, -2], [1, 0, -2]], np.float32),
             "covariances": np.array([[1, 0, 0, 1, 0, 1]] * 2, np.float32),
             "colors": np.array([[1, 0, 0], [0, 0, 1]], np.float32), "alphas": np.array([0.5, 0.5], np.float32),
             "sh_degree": 1, "sh": np.full((2, 9), 0.25, np.float32)}
    return skm_mod.LodTree.build(cloud, 6, 32, threads=1)


def test_version_two_round_trip_caps_sh_and_never_overwrites(skm_mod, tmp_path):
    """LodFileTest.cpp:50-69 and :96-127: moment matching of the fixture, the version-2 file's size,
    SH capped at decode, and an existing file left alone."""
    t = fixture_tree(skm_mod)
    assert t.node_count == 3 and t.leaf_count == 2
    assert t.nodes["positions"][0, 0] == 0 and t.nodes["covariances"][0, 0] == 2 and t.nodes["covariances"][0, 3] == 1
    assert t.nodes["colors"][0, 0] == np.float32(0.5) and t.nodes["sh"][0, 0] == np.float32(0.25)
    assert len(t.clusters) == 1 and len(t.leaves) == 2 and t.clusters[0, 15] == 2
    path = tmp_path / "lod-file.lodsplat"
    t.write(str(path))
    data = path.read_bytes()
    assert data[8] == 2 and len(data) == 64 + 3 * 100 + 64 + 8
    back = skm_mod.LodTree.decode(data)
    for k in ("positions", "covariances", "colors", "alphas", "sh"):
        assert np.array_equal(back.nodes[k], t.nodes[k]), k
    dc = skm_mod.LodTree.decode(data, 0)
    assert dc.sh_degree == 0 and dc.nodes["sh"] is None
    with pytest.raises(skm_mod.SkmError):
        t.write(str(path))
    assert path.read_bytes() == data


def flip(data, offset, value=255):
    b = bytearray(data)
    b[offset] = value
    return bytes(b)


def put_f32(data, offset, value):
    b = bytearray(data)
    b[offset:offset + 4] = struct.pack("<f", value)
    return bytes(b)


def put_u32(data, offset, value):
    b = bytearray(data)
    b[offset:offset + 4] = struct.pack("<I", value)
    return bytes(b)


def test_corruptions_are_refused_like_the_reference(skm_mod, oracle, ref, tmp_path):
    """LodFileTest.cpp:96-156: truncations, header bytes, child ranges, a negative covariance, NaN SH,
    and the selection records (a zero extent, a wrong leaf reference, a NaN colour variance): both
    decoders refuse each, with the same message."""
    t = fixture_tree(skm_mod)
    path = tmp_path / "f.lodsplat"
    t.write(str(path))
    good = path.read_bytes()
    n, stride = t.node_count, 100
    clusters_at = 64 + n * stride
    bad = [good[:k] for k in (0, 8, 63, len(good) - 1)]
    bad += [flip(good, k) for k in (8, 12, 16, 24, 28, 64 + 16, 64 + 20)]
    bad += [put_u32(good, 64 + 16, 0),                       # node 0's children start at itself
            put_f32(good, 64 + 24, -1.0),                     # node 0's covariance xx < 0
            put_f32(good, 64 + 64, float("nan")),             # node 0's first SH coefficient
            put_f32(good, clusters_at + 16, 0.0),             # the cluster's extent x
            put_u32(good, clusters_at + 64, 0),               # the first leaf reference -> the root
            put_f32(good, clusters_at + 32, float("nan")),    # the cluster's colour variance
            put_f32(good, 64 + 60, 1001.0),                   # opacity above 1000
            put_f32(good, 64 + 48, 1.5)]                      # colour above 1
    for k, data in enumerate(bad):
        mine, why_mine = mine_decode(skm_mod, data)
        theirs, why_ref = ref_decode(ref, data)
        assert mine is None and not theirs, k
        assert why_mine == why_ref, (k, why_mine, why_ref)
    for data in (good,):
        mine, _ = mine_decode(skm_mod, data)
        theirs, _ = ref_decode(ref, data)
        assert mine is not None and theirs
        ref.ref_lod_free(theirs)


def test_node_limit_is_lifted_past_the_reference_cap(skm_mod, ref):
    """A header of 20,250,000 nodes (the 20M-leaf world's tree): the reference's readLodSplatHeader
    refuses it (kMaxNodes = 20,000,000, LodFile.cpp:24); this one accepts it and checks the file's
    length as before. Only the 64-byte header is built; the length is what a whole file would have."""
    nodes, leaves, clusters = 20_250_000, 20_000_000, 250_000
    hdr = bytearray(64)
    hdr[0:8] = b"LODSPLAT"
    struct.pack_into("<7I", hdr, 8, 2, nodes, leaves, 1, 12, clusters, 0)
    struct.pack_into("<6f", hdr, 32, -100, -5, -100, 100, 5, 100)
    struct.pack_into("<2I", hdr, 56, leaves, 64)
    size = 64 + nodes * (64 + 36) + clusters * 64 + leaves * 4
    h, why = skm_mod.read_lod_header((bytes(hdr), size))
    assert h is not None, why
    assert h["node_count"] == nodes and h["sh_degree"] == 1 and h["depth"] == 12
    h, why = skm_mod.read_lod_header((bytes(hdr), size - 1))
    assert h is None and "length" in why
    buf = np.frombuffer(bytes(hdr), np.uint8)
    err = C.create_string_buffer(256)
    out = np.zeros(7, np.uint64)
    assert ref.ref_lod_header(buf.ctypes.data_as(C.c_void_p), size, out.ctypes.data_as(C.POINTER(C.c_uint64)), err,
                              256) == 0
    assert err.value.decode() == "invalid LODSPLAT header"
    # At 20,000,000 nodes both read it.
    struct.pack_into("<2I", hdr, 12, 20_000_000, 19_800_000)
    struct.pack_into("<I", hdr, 28, 200_000)
    struct.pack_into("<I", hdr, 56, 19_800_000)
    size = 64 + 20_000_000 * 100 + 200_000 * 64 + 19_800_000 * 4
    assert skm_mod.read_lod_header((bytes(hdr), size))[0] is not None
    buf = np.frombuffer(bytes(hdr), np.uint8)
    assert ref.ref_lod_header(buf.ctypes.data_as(C.c_void_p), size, out.ctypes.data_as(C.POINTER(C.c_uint64)), err,
                              256) == 1
