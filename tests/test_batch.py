"""Batched builds: every mesh of a batch against the CPU oracle built on that
mesh alone (slot-wise BVH8 comparison and primIdx, the bar of the single-mesh
parity tests), graph and stage-by-stage runs, and the argument checks.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def nx():
    from nexusbvh_amd import native
    if native.device_count() == 0:
        pytest.skip("no GPU")
    from nexusbvh_amd import batch
    return batch


def _batch(meshes, seed, lo, hi):
    from nexusbvh_amd import scenes
    tris, offsets = scenes.batch(meshes, seed, lo, hi)
    return tris, np.asarray(offsets, dtype=np.uint32)


def _check_meshes(plan, tris, offsets, width=32, radius=8, which=None, slots=0):
    for m in (range(len(offsets) - 1) if which is None else which):
        a, b = int(offsets[m]), int(offsets[m + 1])
        ref = O.build(tris[a:b], 1, 32, width, radius)
        if slots:
            ref["bvh8"], ref["prim_idx"] = O.collapse_bvh8(ref["nodes"], b - a, slots)
        nodes, prim = plan.mesh_to_host(m)
        code, where = O.compare_bvh8(nodes, prim, ref["bvh8"], ref["prim_idx"], b - a)
        assert code == 0, (m, b - a, code, where)
        box = plan.mesh(m).bounds
        got = [box.bmin.x, box.bmin.y, box.bmin.z, box.bmax.x, box.bmax.y, box.bmax.z]
        assert np.array_equal(np.array(got, np.float32), ref["scene"]), m


def test_batch_meshes_match_the_oracle(nx):
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = _batch(60, 11, 2, 6000)
    say("batch", len(offsets) - 1, len(tris))
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    plan.run()
    _check_meshes(plan, tris, offsets)
    assert plan.node_total == sum(plan.mesh(m).node_count for m in range(plan.meshes))


def test_batch_2048_full_size(nx):
    """The 2,048-mesh benchmark batch (19.7M triangles) in one plan: every
    mesh slot for slot against the oracle built on that mesh alone."""
    import numpy as np
    from nexusbvh_amd import scenes
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = scenes.batch(2048)
    offsets = np.asarray(offsets, dtype=np.uint32)
    say("batch_2048", len(offsets) - 1, len(tris))
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    plan.run()
    _check_meshes(plan, tris, offsets)
    say("batch_2048 all meshes equal")


def test_batch_many_tiny_meshes(nx):
    """Hundreds of meshes of 2-8 triangles: a 1024-primitive bounds tile then
    spans more meshes than it reduces in LDS (the rest take global atomics),
    and most waves cross mesh boundaries."""
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = _batch(400, 21, 2, 8)
    say("batch tiny", len(offsets) - 1, len(tris))
    assert max(np.diff(offsets)) <= 8 and len(tris) > 1024
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    plan.run()
    _check_meshes(plan, tris, offsets)


def test_batch_graph_runs_repeat_and_equal_stage_runs(nx):
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = _batch(100, 12, 64, 20000)
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    out = []
    for k in range(3):
        say("graph run", k)
        t = plan.run(stage_times=(k == 1))
        if t is not None:
            assert len(t) == 6 and all(x >= 0 for x in t)
        out.append([plan.mesh_to_host(m) for m in (0, 37, 99)])
    # node ids and primIdx slots follow the collapse's allocation order, which
    # can differ between runs; structure, boxes and the primitive of every
    # leaf slot are identical (slot-wise compare)
    for run in (out[1], out[2]):
        for m, a, b in zip((0, 37, 99), out[0], run):
            n = int(offsets[m + 1] - offsets[m])
            assert np.array_equal(np.sort(a[1]), np.arange(n, dtype=np.uint32)), m
            assert np.array_equal(np.sort(b[1]), np.arange(n, dtype=np.uint32)), m
            assert O.compare_bvh8(a[0], a[1], b[0], b[1], n)[0] == 0, m
    _check_meshes(plan, tris, offsets, which=(0, 5, 37, 99))


@pytest.mark.parametrize("mode,radius", [(1, 8), (0, 2), (0, 16)])
def test_batch_modes_match_the_oracle(nx, mode, radius):
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives
    tris, offsets = _batch(24, 13 + radius, 2, 9000)
    say("batch mode", mode, radius)
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets,
                        BuildConfig(mortonBits=32, mergeMode=mode, searchRadius=radius, sortDigitBits=8))
    plan.run()
    _check_meshes(plan, tris, offsets, width=64 if mode == 1 else 32, radius=radius)


def test_batch_hploc_blocks_per_cu(nx):
    """A plan with fewer persistent H-PLOC blocks per CU (an autotuner knob)
    builds the same trees."""
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives
    tris, offsets = _batch(30, 23, 2, 8000)
    say("batch hploc blocks per CU")
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets, BuildConfig(mortonBits=32, hplocBlocksPerCu=4))
    plan.run()
    _check_meshes(plan, tris, offsets)


@pytest.mark.parametrize("count", [2, 3, 5000])
def test_batch_of_one_mesh(nx, count):
    """A plan of one mesh (one tree in the first block of the top-levels
    kernel, a 29-bit key window) builds what nxb_build_bvh8 builds for it."""
    from nexusbvh_amd import scenes
    from nexusbvh_amd.builder import DevicePrimitives
    tris = scenes.batch_mesh(0, count, seed=41)
    offsets = np.array([0, count], dtype=np.uint32)
    say("batch of one mesh", count)
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    plan.run()
    _check_meshes(plan, tris, offsets)
    assert plan.node_total == plan.mesh(0).node_count


@pytest.mark.parametrize("digit_bits", [6, 9])
def test_batch_digit_widths(nx, digit_bits):
    """The Morton kernel counts the digits of every sort pass for each digit
    width (40 meshes: a 34-bit key window, 6 passes of 6-bit digits or 4 of
    9-bit ones)."""
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives
    tris, offsets = _batch(40, 31 + digit_bits, 2, 7000)
    say("batch digit bits", digit_bits)
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets,
                        BuildConfig(mortonBits=32, sortDigitBits=digit_bits))
    plan.run()
    _check_meshes(plan, tris, offsets)


def test_batch_mesh_node_regions(nx):
    """Level-synchronous plans write mesh m's nodes into its own region of
    nxb_bvh8_capacity(count_m) nodes, the regions back to back in mesh order:
    every view starts at its region, fits in it, and the node counts match
    the oracle's trees. Tiny and large meshes are mixed so that the waves of
    the collapse hold runs of several meshes."""
    from nexusbvh_amd import native
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = _batch(120, 29, 2, 3000)
    say("batch node regions", len(offsets) - 1, len(tris))
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    plan.run()
    cap = native.lib().nxb_bvh8_capacity
    base = plan.mesh(0).nodes
    first = 0
    total = 0
    for m in range(plan.meshes):
        count = int(offsets[m + 1] - offsets[m])
        view = plan.mesh(m)
        assert view.nodes == base + 80 * first, m
        assert 1 <= view.node_count <= cap(count), (m, view.node_count, cap(count))
        ref = O.build(tris[int(offsets[m]):int(offsets[m + 1])], 1, 32, 32, 8)
        assert view.node_count == len(ref["bvh8"]), m
        first += cap(count)
        total += view.node_count
    assert plan.node_total == total
    _check_meshes(plan, tris, offsets, which=(0, 1, 59, 119))


def test_batch_persistent_bvh8_kernel(nx):
    """bvh8_lanes != 0 selects the persistent BVH8 kernel instead of the
    level-synchronous one; the meshes match the oracle either way."""
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = _batch(30, 17, 2, 5000)
    say("batch persistent bvh8")
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets, bvh8_lanes=4096)
    plan.run()
    _check_meshes(plan, tris, offsets)


@pytest.mark.parametrize("lanes", [0, 4096])
def test_batch_auction_slots(nx, lanes):
    """slotAssignment = AUCTION in a batch, with the level-synchronous (0)
    and the persistent (4096 lanes) collapse kernels."""
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, SlotAssignment
    tris, offsets = _batch(25, 19, 2, 4000)
    say("batch auction", lanes)
    cfg = BuildConfig(mortonBits=32, slotAssignment=SlotAssignment.AUCTION)
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets, cfg, bvh8_lanes=lanes)
    plan.run()
    _check_meshes(plan, tris, offsets, slots=1)


def test_batch_aabb_primitives(nx):
    from nexusbvh_amd import scenes
    from nexusbvh_amd.builder import DevicePrimitives, PrimType
    tris, offsets = _batch(20, 14, 2, 3000)
    boxes = scenes.boxes_of(tris)
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(boxes, PrimType.AABB), offsets)
    plan.run()
    for m in range(0, 20, 3):
        a, b = int(offsets[m]), int(offsets[m + 1])
        ref = O.build(boxes[a:b], 0, 32, 32, 8)
        nodes, prim = plan.mesh_to_host(m)
        assert O.compare_bvh8(nodes, prim, ref["bvh8"], ref["prim_idx"], b - a)[0] == 0


def test_batch_argument_checks(nx):
    from nexusbvh_amd import native
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives
    tris, offsets = _batch(4, 15, 64, 200)
    d = DevicePrimitives.from_numpy(tris)
    with pytest.raises(native.NxbError) as e:
        nx.BatchPlan(d, offsets, BuildConfig(mortonBits=64))
    assert "NOT_SUPPORTED" in str(e.value)
    bad = offsets.copy()
    bad[1] = 1            # a mesh of one primitive
    with pytest.raises(native.NxbError):
        nx.BatchPlan(d, bad)
    bad = offsets.copy()
    bad[0] = 1
    with pytest.raises(native.NxbError):
        nx.BatchPlan(d, bad)


def test_batch_nonfinite_is_reported(nx):
    from nexusbvh_amd import native
    from nexusbvh_amd.builder import DevicePrimitives
    tris, offsets = _batch(6, 16, 64, 500)
    tris = tris.copy()
    tris[int(offsets[3]) + 1, 4] = np.nan
    plan = nx.BatchPlan(DevicePrimitives.from_numpy(tris), offsets)
    with pytest.raises(native.NxbError) as e:
        plan.run()
    assert "NONFINITE" in str(e.value)


def _spiral(n):
    """Triangles clustered geometrically towards one corner: a deep BVH8
    (depth 11 at 512 triangles, against 7 for uniform ones)."""
    i = np.arange(n)
    c = np.stack([0.5 ** (i / 17.0), 0.5 ** (i / 19.0), 0.5 ** (i / 23.0)], axis=1).astype(np.float32)
    h = np.float32(1e-4)
    return np.concatenate([c, c + np.array([h, 0, 0], np.float32), c + np.array([0, h, 0], np.float32)],
                          axis=1).astype(np.float32)


def test_batch_levels_follow_deeper_geometry(nx):
    """A plan fits its BVH8 level count to the forest of its first run; when
    the primitives in its buffer are replaced by geometry whose trees are
    deeper, the next run still builds every mesh exactly (it records the
    bound for any forest again and builds once more)."""
    import ctypes
    from nexusbvh_amd import native, scenes
    from nexusbvh_amd.builder import DevicePrimitives
    n = 512
    shallow = np.concatenate([scenes.uniform(n, seed=31), scenes.uniform(n, seed=32)])
    deep = np.concatenate([scenes.uniform(n, seed=31), _spiral(n)])
    offsets = np.array([0, n, 2 * n], dtype=np.uint32)
    d = DevicePrimitives.from_numpy(shallow)
    plan = nx.BatchPlan(d, offsets)
    for tris in (shallow, shallow, deep, deep, shallow):
        src = np.ascontiguousarray(tris, dtype=np.float32)
        native.hip_check(native.hip().hipMemcpy(ctypes.c_void_p(d.address), src.ctypes.data_as(ctypes.c_void_p),
                                                src.nbytes, 1), "hipMemcpy H2D")
        plan.run()
        _check_meshes(plan, tris, offsets)
