"""GPU parity: every stage and the whole pipeline through the C ABI against
the CPU oracle on the same seeded inputs.

Bar: bit-exact leaves, scene box, Morton keys, sorted keys and permutation,
BVH8 nodes (slot-wise, all bytes except the child/primitive base indices);
BVH2 isomorphic to the oracle's as an ordered tree with bit-identical boxes;
SAH within 1e-6 relative (FP64 sums of the same FP32 terms in another order).
"""
import os
import sys

import numpy as np
import pytest

from nexusbvh_amd import builder as B
from nexusbvh_amd import native, scenes, stages
from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, MergeMode, PrimType
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(__file__))
import nxb_checks  # noqa: E402

pytestmark = pytest.mark.gpu

SAH_RTOL = 1e-6


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if native.device_count() == 0:
        pytest.skip("no GPU")


def oracle_stages(prims, ptype, bits, width=32, radius=8):
    n = len(prims)
    nodes, scene, bad = O.scene_bounds(prims, ptype)
    keys = O.morton(nodes, n, scene, bits)
    b, e = O.sort_window(bits)
    sk, perm = O.sort_pairs(keys, b, e)
    full = O.hploc(sk, perm, nodes, width, radius)
    return {"leaves": nodes[:n].copy(), "scene": scene, "keys": keys, "sorted": sk, "perm": perm, "bvh2": full}


# ------------------------------------------------------------------- stages
@pytest.mark.parametrize("ptype", [PrimType.TRIANGLE, PrimType.AABB])
def test_scene_bounds_bit_exact(ptype):
    tris = scenes.uniform(300_000)
    prims = tris if ptype == PrimType.TRIANGLE else scenes.boxes_of(tris)
    say("scene bounds", ptype.name, len(prims))
    d = DevicePrimitives.from_numpy(prims, ptype)
    nodes, scene, flags = stages.scene_bounds(d)
    ref_nodes, ref_scene, bad = O.scene_bounds(prims, int(ptype))
    n = len(prims)
    got = stages.nodes_to_host(nodes, n)
    assert flags == 0 and not bad
    assert got.tobytes() == ref_nodes[:n].tobytes()
    assert scene.to_array(np.float32).tobytes() == ref_scene.tobytes()


def test_scene_bounds_unaligned_and_tails():
    tris = scenes.slivers(5000, seed=9)
    for n in (1, 2, 1023, 1025, 4097):
        d = DevicePrimitives.from_numpy(tris[:n])
        nodes, scene, flags = stages.scene_bounds(d)
        ref_nodes, ref_scene, _ = O.scene_bounds(tris[:n], 1)
        assert stages.nodes_to_host(nodes, n).tobytes() == ref_nodes[:n].tobytes(), n
        assert scene.to_array(np.float32).tobytes() == ref_scene.tobytes(), n
    # primitives starting 4 bytes into a buffer take the unaligned load path
    buf = native.DeviceBuffer.from_array(np.concatenate([np.zeros(1, np.float32), tris[:3001].reshape(-1)]))
    nodes = native.DeviceBuffer(3001 * 2 * 32)
    scene = native.DeviceBuffer(24)
    flags = native.DeviceBuffer.from_array(np.zeros(1, np.uint32))
    native.check(native.lib().nxb_compute_scene_bounds(buf.address + 4, 1, 3001, nodes.address, scene.address,
                                                       flags.address, None))
    native.synchronize()
    ref_nodes, ref_scene, _ = O.scene_bounds(tris[:3001], 1)
    assert stages.nodes_to_host(nodes, 3001).tobytes() == ref_nodes[:3001].tobytes()
    assert scene.to_array(np.float32).tobytes() == ref_scene.tobytes()


@pytest.mark.parametrize("bits", [32, 64])
def test_morton_bit_exact(bits):
    for name, tris in (("uniform", scenes.uniform(300_000)), ("icosphere", scenes.icosphere(200_000, 100))):
        say("morton", bits, name)
        n = len(tris)
        ref_nodes, ref_scene, _ = O.scene_bounds(tris, 1)
        nodes = native.DeviceBuffer.from_array(ref_nodes[:n])
        scene = native.DeviceBuffer.from_array(ref_scene)
        keys, vals = stages.morton(nodes, n, scene, bits, with_values=True)
        ref = O.morton(ref_nodes, n, ref_scene, bits)
        assert np.array_equal(keys.to_array(ref.dtype, n), ref)
        assert np.array_equal(vals.to_array(np.uint32, n), np.arange(n, dtype=np.uint32))


def test_morton_flat_axis():
    tris = scenes.uniform(10_000, seed=3)
    tris[:, 2::3] = 0.25  # every z equal: 0/0 on that axis quantizes to 0
    n = len(tris)
    ref_nodes, ref_scene, _ = O.scene_bounds(tris, 1)
    for bits in (32, 64):
        keys, _ = stages.morton(native.DeviceBuffer.from_array(ref_nodes[:n]), n,
                                native.DeviceBuffer.from_array(ref_scene), bits)
        ref = O.morton(ref_nodes, n, ref_scene, bits)
        assert np.array_equal(keys.to_array(ref.dtype, n), ref)


@pytest.mark.parametrize("bits,digit_bits", [(32, 7), (32, 8), (32, 6), (32, 9), (64, 8), (64, 7), (64, 9), (64, 6)])
def test_sort_bit_exact(bits, digit_bits):
    rng = np.random.default_rng(bits + digit_bits)
    dt = np.uint32 if bits == 32 else np.uint64
    b, e = O.sort_window(bits)
    cases = [
        ("random", rng.integers(0, 2 ** (bits - 1), 1_000_003, dtype=np.uint64).astype(dt)),
        ("ties", (rng.integers(0, 16, 300_000).astype(dt) << dt(4)) | rng.integers(0, 4, 300_000).astype(dt)),
        ("tiny", rng.integers(0, 1000, 5).astype(dt)),
        ("one", np.array([7], dtype=dt)),
        ("tile+1", rng.integers(0, 2 ** 20, 4097).astype(dt)),
    ]
    for name, keys in cases:
        say("sort", bits, digit_bits, name, len(keys))
        n = len(keys)
        kd = native.DeviceBuffer.from_array(keys)
        ko, vo = stages.sort_pairs(kd, n, bits, b, e, digit_bits)
        sk, perm = O.sort_pairs(keys, b, e)
        assert np.array_equal(vo.to_array(np.uint32, n), perm), name
        assert np.array_equal(ko.to_array(dt, n), sk), name


@pytest.mark.parametrize("digit_bits", [6, 7, 8, 9])
def test_sort_full_64bit_window(digit_bits):
    """Every key bit in the window (11 passes at 6-bit digits) and keys with the top bit set."""
    rng = np.random.default_rng(100 + digit_bits)
    keys = rng.integers(0, 2 ** 64, 250_001, dtype=np.uint64)
    keys[::7] |= np.uint64(1) << np.uint64(63)
    n = len(keys)
    say("sort64 full", digit_bits, n)
    ko, vo = stages.sort_pairs(native.DeviceBuffer.from_array(keys), n, 64, 0, 64, digit_bits)
    sk, perm = O.sort_pairs(keys, 0, 64)
    assert np.array_equal(vo.to_array(np.uint32, n), perm)
    assert np.array_equal(ko.to_array(np.uint64, n), sk)


@pytest.mark.parametrize("bits,digit_bits", [(32, 6), (32, 7), (32, 8), (32, 9), (64, 6), (64, 7), (64, 8), (64, 9)])
def test_sort_wide_tiles(bits, digit_bits):
    """Inputs of 2^23 keys and more take the wide tile shape (1024-thread tiles, two LDS
    phases): a ragged last tile, runs of equal high digits (rows of one digit) mixed
    with random keys, and caller-supplied values."""
    rng = np.random.default_rng(7 * bits + digit_bits)
    dt = np.uint32 if bits == 32 else np.uint64
    b, e = O.sort_window(bits)
    n = (1 << 23) + 12_345
    keys = rng.integers(0, 2 ** (bits - 1), n, dtype=np.uint64)
    clustered = rng.random(n) < 0.5
    hi = np.uint64(bits - 12)
    keys[clustered] = (keys[clustered] & ((np.uint64(1) << hi) - np.uint64(1))) | (
        rng.integers(0, 3, int(clustered.sum()), dtype=np.uint64) << hi)
    keys = keys.astype(dt)
    vals = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    say("sort wide", bits, digit_bits, n)
    ko, vo = stages.sort_pairs(native.DeviceBuffer.from_array(keys), n, bits, b, e, digit_bits,
                               values=native.DeviceBuffer.from_array(vals))
    sk, perm = O.sort_pairs(keys, b, e)
    assert np.array_equal(vo.to_array(np.uint32, n), vals[perm])
    assert np.array_equal(ko.to_array(dt, n), sk)


@pytest.mark.parametrize("begin,end,digit_bits", [(0, 64, 8), (40, 49, 9), (20, 38, 9), (28, 35, 7)])
def test_sort_wide_windows(begin, end, digit_bits):
    """Wide 64-bit sorts (2^23 keys and more) on other key windows: every key bit (8
    passes), a single pass, two passes whose second digit lies across the two 32-bit
    words, and a 7-bit digit across them; keys with all-zero high words and all-one low
    words mixed in."""
    rng = np.random.default_rng(begin * 100 + end + digit_bits)
    n = (1 << 23) + 777
    keys = rng.integers(0, 2 ** 64, n, dtype=np.uint64)
    keys[::5] &= np.uint64(0xFFFFFFFF)  # high words of zero
    keys[1::5] |= np.uint64(0xFFFFFFFF)  # low words of ones
    say("sort wide window", begin, end, digit_bits, n)
    ko, vo = stages.sort_pairs(native.DeviceBuffer.from_array(keys), n, 64, begin, end, digit_bits)
    sk, perm = O.sort_pairs(keys, begin, end)
    assert np.array_equal(vo.to_array(np.uint32, n), perm)
    assert np.array_equal(ko.to_array(np.uint64, n), sk)


def test_sort_with_values_and_partial_window():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 2 ** 32, 200_000, dtype=np.uint64).astype(np.uint32)
    vals = rng.integers(0, 2 ** 32, 200_000, dtype=np.uint64).astype(np.uint32)
    n = len(keys)
    ko, vo = stages.sort_pairs(native.DeviceBuffer.from_array(keys), n, 32, 5, 21, 8,
                               values=native.DeviceBuffer.from_array(vals))
    window = (keys >> np.uint32(5)) & np.uint32((1 << 16) - 1)
    ref = np.argsort(window, kind="stable")
    assert np.array_equal(vo.to_array(np.uint32, n), vals[ref])
    assert np.array_equal(ko.to_array(np.uint32, n), keys[ref])


@pytest.mark.parametrize("bits,mode", [(32, MergeMode.WAVE32), (64, MergeMode.WAVE32), (64, MergeMode.WAVE64),
                                       (32, MergeMode.WAVE64)])
def test_hploc_isomorphic(bits, mode):
    width = 32 if mode == MergeMode.WAVE32 else 64
    for name, tris in (("uniform", scenes.uniform(300_000)), ("slivers", scenes.slivers(200_000, seed=5))):
        say("hploc", bits, mode.name, name)
        n = len(tris)
        ref = oracle_stages(tris, 1, bits, width)
        nodes = native.DeviceBuffer(max(1, 2 * n - 1) * 32)
        native.hip_check(native.hip().hipMemcpy(nodes.ptr, ref["leaves"].ctypes.data_as(native.c_void_p),
                                                n * 32, native.H2D), "upload leaves")
        kd = native.DeviceBuffer.from_array(ref["sorted"])
        vd = native.DeviceBuffer.from_array(ref["perm"])
        stages.hploc(kd, vd, n, bits, nodes, BuildConfig(mergeMode=mode))
        got = stages.nodes_to_host(nodes, 2 * n - 1)
        code, where = O.compare_bvh2(got, ref["bvh2"], n)
        assert code == 0, (name, code, where)
        # node numbering is deterministic: the arrays are identical
        assert got.tobytes() == ref["bvh2"][: 2 * n - 1].tobytes(), name
        nxb_checks.bvh2_invariants(got, n, scenes.boxes_of(tris), ref["scene"])


@pytest.mark.parametrize("blocks", [1, 5])
def test_hploc_blocks_per_cu(blocks):
    """Fewer persistent H-PLOC blocks per CU (each wave takes a larger share of
    the sorted positions) build the same BVH2, byte for byte."""
    tris = scenes.uniform(300_000, seed=40 + blocks)
    n = len(tris)
    say("hploc blocks per CU", blocks)
    ref = oracle_stages(tris, 1, 64, 32)
    nodes = native.DeviceBuffer.from_array(np.concatenate([ref["leaves"], ref["leaves"][: n - 1]]))
    stages.hploc(native.DeviceBuffer.from_array(ref["sorted"]), native.DeviceBuffer.from_array(ref["perm"]), n, 64,
                 nodes, BuildConfig(hplocBlocksPerCu=blocks))
    assert stages.nodes_to_host(nodes, 2 * n - 1).tobytes() == ref["bvh2"][: 2 * n - 1].tobytes()


def test_hploc_blocks_per_cu_is_checked():
    d = DevicePrimitives.from_numpy(scenes.uniform(100, seed=3))
    with pytest.raises(native.NxbError) as ei:
        B.BuildBVH8(d, 100, BuildConfig(hplocBlocksPerCu=13))
    assert ei.value.status == 1


@pytest.mark.parametrize("radius", [1, 2, 4, 16])
def test_hploc_radius(radius):
    tris = scenes.uniform(100_000, seed=radius)
    n = len(tris)
    ref = oracle_stages(tris, 1, 64, 32, radius)
    nodes = native.DeviceBuffer.from_array(np.concatenate([ref["leaves"], ref["leaves"][: n - 1]]))
    stages.hploc(native.DeviceBuffer.from_array(ref["sorted"]), native.DeviceBuffer.from_array(ref["perm"]), n, 64,
                 nodes, BuildConfig(searchRadius=radius))
    code, where = O.compare_bvh2(stages.nodes_to_host(nodes, 2 * n - 1), ref["bvh2"], n)
    assert code == 0, (code, where)


def test_hploc_radius_table():
    tris = scenes.uniform(50_000, seed=77)
    n = len(tris)
    seg_log2 = 10
    table = np.random.default_rng(3).choice(np.array([2, 4, 8, 16], dtype=np.uint8), (n >> seg_log2) + 1)
    table[5] = 0  # 0 falls back to the fixed radius
    nodes0, scene, _ = O.scene_bounds(tris, 1)
    keys = O.morton(nodes0, n, scene, 64)
    sk, perm = O.sort_pairs(keys, 1, 64)
    ref = O.hploc(sk, perm, nodes0, 32, 8, table=table, seg_log2=seg_log2)
    tbl = native.DeviceBuffer.from_array(table)
    nodes = native.DeviceBuffer.from_array(np.concatenate([nodes0[:n], nodes0[: n - 1]]))
    cfg = BuildConfig(radiusPolicy=B.RadiusPolicy.TABLE, radiusTable=tbl, radiusTableLen=len(table),
                      segmentLog2=seg_log2)
    stages.hploc(native.DeviceBuffer.from_array(sk), native.DeviceBuffer.from_array(perm), n, 64, nodes, cfg)
    code, where = O.compare_bvh2(stages.nodes_to_host(nodes, 2 * n - 1), ref, n)
    assert code == 0, (code, where)


@pytest.mark.parametrize("name,n", [("uniform", 300_000), ("slivers", 200_000), ("tiny", 2), ("one", 1),
                                    ("small", 1000)])
def test_bvh8_slotwise(name, n):
    tris = scenes.slivers(n, seed=4) if name == "slivers" else scenes.uniform(n, seed=n)
    say("bvh8", name, n)
    ref = oracle_stages(tris, 1, 64)
    ref8, refp = O.collapse_bvh8(ref["bvh2"], n)
    st, b8, prim, count = stages.collapse(native.DeviceBuffer.from_array(ref["bvh2"]), n)
    native.check(st, "collapse")
    got8 = stages.bvh8_to_host(b8, count)
    gotp = prim.to_array(np.uint32, n)
    assert count == len(ref8)
    code, where = O.compare_bvh8(got8, gotp, ref8, refp, n)
    assert code == 0, (code, where)
    nxb_checks.bvh8_invariants(got8, gotp, n)


def test_bvh8_spin_limit_trips_to_status():
    # A tiny host-set spin limit makes waiting lanes give up: the kernel ends
    # normally and the build reports NXB_ERROR_TIMEOUT.
    tris = scenes.uniform(200_000, seed=6)
    n = len(tris)
    ref = oracle_stages(tris, 1, 64)
    st, b8, prim, count = stages.collapse(native.DeviceBuffer.from_array(ref["bvh2"]), n, spin_limit=1)
    assert st == 5
    # the library stays usable afterwards
    st, b8, prim, count = stages.collapse(native.DeviceBuffer.from_array(ref["bvh2"]), n)
    assert st == 0


# ------------------------------------------------------------------ pipeline
def full_case(tris, ptype, bits, mode=MergeMode.WAVE32, radius=8):
    n = len(tris)
    width = 32 if mode == MergeMode.WAVE32 else 64
    prims = tris if ptype == PrimType.TRIANGLE else scenes.boxes_of(tris)
    ref = O.build(prims, int(ptype), bits, width, radius)
    d = DevicePrimitives.from_numpy(prims, ptype)
    cfg = BuildConfig(mortonBits=bits, mergeMode=mode, searchRadius=radius)
    m2 = B.BVHBuildMetrics()
    bvh2 = B.BuildBVH2(d, n, cfg, m2)
    h = B.ToHost(bvh2)
    code, where = O.compare_bvh2(h.nodes, ref["nodes"], n)
    assert code == 0, (code, where)
    assert h.nodes.tobytes() == ref["nodes"][: 2 * n - 1].tobytes()
    assert np.array_equal(np.array(h.bounds, dtype=np.float32).reshape(-1), ref["scene"])
    sah2_ref = O.sah_bvh2(ref["nodes"], 2 * n - 1, ref["scene"])
    assert B.sah_bvh2(bvh2) == pytest.approx(sah2_ref, rel=SAH_RTOL)
    assert m2.bvh2Cost == pytest.approx(sah2_ref, rel=SAH_RTOL)
    B.FreeHostBVH(h)
    B.FreeDeviceBVH(bvh2)
    m8 = B.BVHBuildMetrics()
    bvh8 = B.BuildBVH8(d, n, cfg, m8)
    nodes8, prim8 = B.bvh8_to_host(bvh8)
    code, where = O.compare_bvh8(nodes8, prim8, ref["bvh8"], ref["prim_idx"], n)
    assert code == 0, (code, where)
    sah8_ref = O.sah_bvh8(ref["bvh8"], len(ref["bvh8"]), ref["scene"])
    assert B.sah_bvh8(bvh8) == pytest.approx(sah8_ref, rel=SAH_RTOL)
    assert m8.bvh8Cost == pytest.approx(sah8_ref, rel=SAH_RTOL)
    assert m8.averageChildPerNode == pytest.approx((n + len(ref["bvh8"]) - 1) / len(ref["bvh8"]), rel=1e-6)
    assert m8.totalTime > 0
    B.FreeDeviceBVH(bvh8)


def test_pipeline_uniform_300k_32bit():
    say("pipeline uniform_300k K=4")
    full_case(scenes.uniform(300_000), PrimType.TRIANGLE, 32)


def test_pipeline_icosphere_1m1_both_widths():
    tris = scenes.icosphere(1_100_000, 235)
    for bits in (32, 64):
        say("pipeline icosphere_1m1", bits)
        full_case(tris, PrimType.TRIANGLE, bits)


def test_pipeline_slivers_2m9_wave64_and_wave32():
    tris = scenes.slivers(2_900_000)
    for mode in (MergeMode.WAVE64, MergeMode.WAVE32):
        say("pipeline slivers_2m9", mode.name)
        full_case(tris, PrimType.TRIANGLE, 64, mode)


def test_pipeline_blobs_28m1_full_size():
    """The 28.1M-triangle benchmark scene with 64-bit keys, against the
    oracle's full build (about 20 s on one core): BVH2 byte for byte, BVH8
    slot for slot, SAH."""
    import time
    tris = scenes.workload("blobs_28m1")
    say("pipeline blobs_28m1 (oracle build ~20 s)", len(tris))
    t0 = time.time()
    full_case(tris, PrimType.TRIANGLE, 64)
    say("blobs_28m1 done in", round(time.time() - t0, 1), "s")


def test_pipeline_blobs_28m1_32bit_keys():
    """The same scene with 32-bit keys (the reference's published setting):
    the wide sort tiles with 7-bit digits and the first pass's tile counts
    taken from the Morton kernel, against the oracle's full build."""
    tris = scenes.workload("blobs_28m1")
    say("pipeline blobs_28m1 K=4 (oracle build ~20 s)", len(tris))
    full_case(tris, PrimType.TRIANGLE, 32)


def test_pipeline_aabb_primitives():
    say("pipeline AABB")
    full_case(scenes.slivers(150_000, seed=12), PrimType.AABB, 64)


@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 32, 33, 64, 65, 257, 4096, 4097])
def test_pipeline_tiny(n):
    full_case(scenes.uniform(n, seed=1000 + n), PrimType.TRIANGLE, 64)


def test_pipeline_duplicates_and_flat():
    dup = np.tile(scenes.uniform(7, seed=2), (1000, 1))
    full_case(dup, PrimType.TRIANGLE, 32)
    flat = scenes.uniform(20_000, seed=8)
    flat[:, 1::3] = 0.75
    full_case(flat, PrimType.TRIANGLE, 64)


def test_pipeline_golden_fixture():
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "small_scene.json")) as f:
        g = json.load(f)
    tris = np.array(g["triangles"], dtype=np.float32)
    n = len(tris)
    from test_oracle import _oracle_canon
    d = DevicePrimitives.from_numpy(tris)
    for case in g["cases"]:
        mode = MergeMode.WAVE32 if case["width"] == 32 else MergeMode.WAVE64
        cfg = BuildConfig(mortonBits=case["bits"], mergeMode=mode, searchRadius=case["radius"])
        bvh2 = B.BuildBVH2(d, n, cfg)
        h = B.ToHost(bvh2)
        assert _oracle_canon(h.nodes, n) == case["bvh2_canon"]
        assert B.sah_bvh2(bvh2) == pytest.approx(case["sah2"], rel=SAH_RTOL)
        B.FreeDeviceBVH(bvh2)
        bvh8 = B.BuildBVH8(d, n, cfg)
        assert bvh8.node_count == case["bvh8_nodes"]
        assert B.sah_bvh8(bvh8) == pytest.approx(case["sah8"], rel=SAH_RTOL)
        B.FreeDeviceBVH(bvh8)


def test_structure_is_deterministic_across_runs():
    tris = scenes.slivers(500_000, seed=21)
    d = DevicePrimitives.from_numpy(tris)
    a = B.BuildBVH2(d, len(tris))
    b = B.BuildBVH2(d, len(tris))
    ha, hb = B.ToHost(a), B.ToHost(b)
    code, where = O.compare_bvh2(ha.nodes, hb.nodes, len(tris))
    assert code == 0
    B.FreeDeviceBVH(a)
    B.FreeDeviceBVH(b)


# ------------------------------------------------------------- error paths
def test_nonfinite_input_is_reported():
    tris = scenes.uniform(10_000, seed=5)
    tris[123, 7] = np.nan
    d = DevicePrimitives.from_numpy(tris)
    with pytest.raises(native.NxbError) as ei:
        B.BuildBVH8(d, len(tris))
    assert ei.value.status == 3
    tris[123, 7] = np.inf
    d = DevicePrimitives.from_numpy(tris)
    with pytest.raises(native.NxbError) as ei:
        B.BuildBVH2(d, len(tris))
    assert ei.value.status == 3


@pytest.mark.parametrize("first_levels", [1, 3])
def test_bvh8_levels_added_after_readback(monkeypatch, first_levels):
    """The BVH8 collapse runs level by level; with only the first 1 or 3 levels
    enqueued before the level counters are read back, the rest are added
    afterwards and the result still matches the oracle slot for slot."""
    tris = scenes.uniform(60_000, seed=21)
    n = len(tris)
    ref = O.build(tris, 1, 32, 32, 8)
    d = DevicePrimitives.from_numpy(tris)
    monkeypatch.setenv("NXB_BVH8_FIRST_LEVELS", str(first_levels))
    say("bvh8 first levels", first_levels)
    bvh = B.BuildBVH8(d, n, BuildConfig(mortonBits=32))
    nodes, prim = B.bvh8_to_host(bvh)
    B.FreeDeviceBVH(bvh)
    code, where = O.compare_bvh8(nodes, prim, ref["bvh8"], ref["prim_idx"], n)
    assert code == 0, (code, where)


def _bvh8_depth(nodes: np.ndarray) -> int:
    """Levels of a BVH8 (inner children of node k at child_base.., one per imask bit)."""
    d = np.zeros(len(nodes), dtype=np.int64)
    d[0] = 1
    for k in range(len(nodes)):
        c = bin(int(nodes["imask"][k])).count("1")
        if c:
            b = int(nodes["child_base"][k])
            d[b:b + c] = d[k] + 1
    return int(d.max())


@pytest.mark.parametrize("name,n,bits", [("uniform", 200_000, 64), ("slivers", 100_000, 32), ("small", 9, 64),
                                       ("two", 2, 32), ("one", 1, 64)])
def test_bvh8_auction_slots_match_the_oracle(name, n, bits):
    """slotAssignment = AUCTION: slot for slot the oracle's auction over the
    same BVH2 (the level-synchronous collapse kernels)."""
    tris = scenes.slivers(n, seed=41) if name == "slivers" else scenes.uniform(n, seed=40 + n)
    say("bvh8 auction", name, n, bits)
    ref = O.build(tris, 1, bits, 32, 8)
    a8, ap = O.collapse_bvh8(ref["nodes"], n, 1)
    d = DevicePrimitives.from_numpy(tris)
    bvh8 = B.BuildBVH8(d, n, BuildConfig(mortonBits=bits, slotAssignment=B.SlotAssignment.AUCTION))
    try:
        nodes8, prim8 = B.bvh8_to_host(bvh8)
    finally:
        B.FreeDeviceBVH(bvh8)
    code, where = O.compare_bvh8(nodes8, prim8, a8, ap, n)
    assert code == 0, (code, where)
    nxb_checks.bvh8_invariants(nodes8, prim8, n)
    if n > 8:
        assert nodes8.tobytes() != ref["bvh8"].tobytes()  # the greedy tree differs


def test_slot_assignment_is_checked():
    d = DevicePrimitives.from_numpy(scenes.uniform(100, seed=3))
    with pytest.raises(native.NxbError) as ei:
        B.BuildBVH8(d, 100, BuildConfig(slotAssignment=2))
    assert ei.value.status == 1


def test_bvh8_level_count_follows_the_last_tree():
    """The first batch of BVH8 level launches is sized from the device's last
    tree of a similar primitive count. A much deeper tree of the same size
    after a shallow one still comes out whole (levels added after the
    readback), and so does the shallow one again after it."""
    shallow = scenes.uniform(60_000, seed=21)
    deep = scenes.uniform(60_000, seed=22)
    lo, hi = float(deep.min()), float(deep.max())
    deep = (lo + (hi - lo) * ((deep - lo) / (hi - lo)) ** 4).astype(np.float32)  # clustered near one corner
    depths = []
    for name, tris in (("shallow", shallow), ("deep", deep), ("shallow", shallow)):
        n = len(tris)
        ref = O.build(tris, 1, 32, 32, 8)
        depths.append(_bvh8_depth(ref["bvh8"]))
        say("bvh8 depth history", name, depths[-1])
        d = DevicePrimitives.from_numpy(tris)
        bvh = B.BuildBVH8(d, n, BuildConfig(mortonBits=32))
        nodes, prim = B.bvh8_to_host(bvh)
        B.FreeDeviceBVH(bvh)
        code, where = O.compare_bvh8(nodes, prim, ref["bvh8"], ref["prim_idx"], n)
        assert code == 0, (name, code, where)
    assert depths[1] > depths[0] + 1  # the deep tree needs levels beyond the first batch


def test_bvh8_to_host_and_dump(tmp_path):
    """ToHost(BVH8) (nodes and primIdx through nxb_bvh8_to_host) equals a
    direct copy of the device arrays and the oracle's BVH8, and survives a
    versioned dump."""
    from nexusbvh_amd import dump
    tris = scenes.uniform(40_000, seed=23)
    n = len(tris)
    d = DevicePrimitives.from_numpy(tris)
    bvh = B.BuildBVH8(d, n, BuildConfig(mortonBits=32))
    try:
        h = B.ToHost(bvh)
        nodes, prim = B.bvh8_to_host(bvh)
    finally:
        B.FreeDeviceBVH(bvh)
    assert h.nodeCount == len(nodes) and h.primCount == n
    assert h.nodes.tobytes() == nodes.tobytes() and np.array_equal(h.primIdx, prim)
    ref = O.build(tris, 1, 32, 32, 8)
    code, where = O.compare_bvh8(h.nodes, h.primIdx, ref["bvh8"], ref["prim_idx"], n)
    assert code == 0, (code, where)
    p = str(tmp_path / "bvh8.npz")
    dump.save(p, h, workload="uniform_40k")
    back, meta = dump.load(p)
    assert dump.first_difference(h, back) is None and meta["workload"] == "uniform_40k"


def test_allocation_failure_is_reported(monkeypatch):
    tris = scenes.uniform(1000, seed=5)
    d = DevicePrimitives.from_numpy(tris)
    monkeypatch.setenv("NXB_FAIL_ALLOC", "1")
    with pytest.raises(native.NxbError) as ei:
        B.BuildBVH8(d, len(tris))
    assert ei.value.status == 2
    monkeypatch.delenv("NXB_FAIL_ALLOC")
    bvh = B.BuildBVH8(d, len(tris))
    B.FreeDeviceBVH(bvh)


def test_cpp_dropin_benchmark_runs():
    """The C++ user of NXB/BVHBuilder.h (BenchmarkBuild over BuildBVH8<Triangle>)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nexusbvh_amd", "lib",
                       "nxb_benchmark")
    r = subprocess.run([exe, "200000", "2", "3", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "BVH8, 200000 primitives, 32-bit Morton, 2 warm-up + 3 measured builds" in out, out
    rows = {}
    for line in out.splitlines():
        for name in ("scene bounds", "Morton codes", "radix sort", "BVH2 H-PLOC", "BVH8 collapse", "whole build"):
            if line.startswith(name):
                rows[name] = [float(x.rstrip("%")) for x in line[len(name):].split()]
    assert len(rows) == 6, out
    for name, (mean_ms, min_ms, mprims, gbs, frac) in rows.items():
        # mean/min ms, Mprim/s = n / mean, GB/s = algorithmic bytes / mean, share of 8 TB/s
        assert 0 < min_ms <= mean_ms * 1.0001, (name, rows[name])
        assert mprims == pytest.approx(200000 / (mean_ms * 1e3), rel=5e-3, abs=0.06), (name, rows[name])
        assert frac == pytest.approx(100 * gbs / 8000, rel=2e-3, abs=0.06), (name, rows[name])
    # 32-bit keys: bounds 36 + 32 B, Morton 24 + 4 + 4 B, sort 2 * 8 * 4 + 4 B per triangle
    for name, per in (("scene bounds", 68), ("Morton codes", 32), ("radix sort", 68)):
        mean_ms, gbs = rows[name][0], rows[name][3]
        assert gbs == pytest.approx(per * 200000 / (mean_ms * 1e6), rel=5e-3, abs=0.06), (name, rows[name])
    assert "SAH: BVH2" in out and "BVH8" in out
    assert "root 399998 is internal" in out


def _indexed(tris):
    v = tris.reshape(-1, 3)
    uniq, inv = np.unique(v, axis=0, return_inverse=True)
    return uniq.astype(np.float32), inv.reshape(-1, 3).astype(np.uint32)


@pytest.mark.parametrize("bits", [32, 64])
def test_indexed_mesh_equals_triangle_soup(bits):
    """An indexed mesh (vertices shared between triangles) builds exactly the
    BVH2 (byte for byte) and BVH8 (slot for slot) of its expanded triangle
    array."""
    tris = scenes.icosphere(120_000, freq=80)
    verts, idx = _indexed(tris)
    assert len(verts) < len(tris) * 3 // 4  # vertices really are shared
    n = len(tris)
    mesh = B.DeviceIndexedMesh.from_numpy(verts, idx)
    soup = DevicePrimitives.from_numpy(tris)
    cfg = BuildConfig(mortonBits=bits)
    a8 = B.BuildBVH8Indexed(mesh, cfg)
    b8 = B.BuildBVH8(soup, n, cfg)
    a2 = B.BuildBVH2Indexed(mesh, cfg)
    b2 = B.BuildBVH2(soup, n, cfg)
    try:
        na, pa = B.bvh8_to_host(a8)
        nb, pb = B.bvh8_to_host(b8)
        h2a, h2b = B.ToHost(a2), B.ToHost(b2)
    finally:
        for x in (a8, b8, a2, b2):
            B.FreeDeviceBVH(x)
    say("indexed", n, "tris", len(verts), "verts", bits)
    # BVH8 node ids within a level follow the order blocks reach the level
    # counter, so the trees are compared slot by slot from the root
    code, where = O.compare_bvh8(na, pa, nb, pb, n)
    assert code == 0, (code, where)
    assert h2a.nodes.tobytes() == h2b.nodes.tobytes() and h2a.bounds == h2b.bounds


def test_indexed_mesh_bad_index_is_reported():
    tris = scenes.uniform(5_000, seed=24)
    verts, idx = _indexed(tris)
    idx[1234, 1] = len(verts) + 7
    mesh = B.DeviceIndexedMesh.from_numpy(verts, idx)
    with pytest.raises(native.NxbError) as ei:
        B.BuildBVH8Indexed(mesh)
    assert ei.value.status == 1  # NXB_ERROR_INVALID_VALUE
    idx[1234, 1] = 0
    ok = B.BuildBVH8Indexed(B.DeviceIndexedMesh.from_numpy(verts, idx))
    B.FreeDeviceBVH(ok)


def test_async_stage_then_build_on_another_stream():
    """nxb_radix_sort_pairs and nxb_compute_scene_bounds without a workspace
    of the caller's return before their kernels finish, using the library's
    scratch buffer; a build enqueued right after on another stream uses the
    same buffer. Both results must still be exact: the sort against numpy's
    stable sort, the scene box against numpy, the build against the oracle."""
    import ctypes
    rng = np.random.default_rng(77)
    m = 1 << 22
    keys = rng.integers(0, 1 << 62, m, dtype=np.uint64)
    kin = native.DeviceBuffer.from_array(keys)
    stages.sort_pairs(kin, m, 64, 1, 63, 8)  # sizes the scratch buffer for m keys
    native.synchronize()
    tris = scenes.uniform(100_000, seed=78)
    d = DevicePrimitives.from_numpy(tris)
    ref = O.build(tris, 1, 64, 32, 8)
    big = scenes.uniform(1 << 20, seed=79)
    dbig = DevicePrimitives.from_numpy(big)
    a, b = native.Stream(), native.Stream()
    for trial in range(3):
        say("async stage + build, trial", trial)
        kout = native.DeviceBuffer(m * 8)
        vout = native.DeviceBuffer(m * 4)
        native.check(native.lib().nxb_radix_sort_pairs(kin.address, None, kout.address, vout.address, m, 64, 1, 63,
                                                       8, None, 0, int(a)), "radix_sort_pairs")
        nodes = native.DeviceBuffer((2 * len(big) - 1) * 32)
        scene = native.DeviceBuffer(24)
        flags = native.DeviceBuffer.from_array(np.zeros(1, dtype=np.uint32))
        native.check(native.lib().nxb_compute_scene_bounds(dbig.address, int(dbig.prim_type), len(big), nodes.address,
                                                           scene.address, flags.address, int(a)), "scene_bounds")
        bvh = B.BuildBVH8(d, len(tris), BuildConfig(), stream=int(b))
        nodes8, prim8 = B.bvh8_to_host(bvh)
        B.FreeDeviceBVH(bvh)
        a.synchronize()
        code, where = O.compare_bvh8(nodes8, prim8, ref["bvh8"], ref["prim_idx"], len(tris))
        assert code == 0, (trial, code, where)
        perm = np.argsort((keys >> np.uint64(1)) & np.uint64((1 << 62) - 1), kind="stable")
        assert np.array_equal(vout.to_array(np.uint32, m), perm.astype(np.uint32)), trial
        assert np.array_equal(kout.to_array(np.uint64, m), keys[perm]), trial
        v = big.reshape(-1, 3)
        want = np.concatenate([v.min(axis=0), v.max(axis=0)]).astype(np.float32)
        assert np.array_equal(scene.to_array(np.float32, 6), want), trial
        assert int(flags.to_array(np.uint32)[0]) == 0


def test_python_benchmark_build_reports_rates(capsys):
    """BenchmarkBuild (Python mirror) returns the mean metrics and reports
    per-stage Mprim/s and algorithmic GB/s."""
    tris = scenes.uniform(150_000, seed=80)
    d = DevicePrimitives.from_numpy(tris)
    mean = B.BenchmarkBuild(B.BuildBVH8, 1, 3, d, len(tris), BuildConfig(mortonBits=64))
    out = capsys.readouterr().out
    assert "BVH8, 150000 primitives, 64-bit Morton, 1 warm-up + 3 measured builds" in out, out
    sort_row = [ln for ln in out.splitlines() if ln.startswith("radix sort")][0].split()
    mean_ms, gbs = float(sort_row[2]), float(sort_row[5])
    assert mean_ms == pytest.approx(mean.radixSortTime, abs=1e-5)
    # 64-bit keys: 8 passes of (8 + 4) B read and written + 8 B of digit counting
    assert gbs == pytest.approx(200 * 150000 / (mean_ms * 1e6), rel=5e-3)
    assert mean.totalTime > 0 and mean.bvh8Cost > 0 and mean.averageChildPerNode > 1
