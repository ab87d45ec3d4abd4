"""nxb_build_bvh2_noalloc: the allocation-free BVH2 build, through the C ABI,
against nxb_build_bvh2 (byte-identical nodes and the same scene box), on two
calls in a row (the second with a smaller input, so the shared scratch array
holds the first call's contents), at the size limit, and the limit itself."""
import ctypes

import numpy as np
import pytest

from nexusbvh_amd import builder as B
from nexusbvh_amd import native, scenes
from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, MergeMode, PrimType
from nexusbvh_amd.native import BVH2, DeviceBuffer, lib

pytestmark = pytest.mark.gpu
NOALLOC_MAX = 1 << 19  # NXB_NOALLOC_MAX_PRIMS


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if native.device_count() == 0:
        pytest.skip("no GPU")


def noalloc(d, n, cfg):
    nodes = DeviceBuffer(max(1, 2 * n - 1) * 32)
    out = BVH2()
    st = lib().nxb_build_bvh2_noalloc(d.address, int(d.prim_type), n, ctypes.byref(cfg.to_c()), nodes.address,
                                      ctypes.byref(out), None)
    return st, out, nodes


def reference(d, n, cfg):
    bvh = B.BuildBVH2(d, n, cfg)
    h = B.ToHost(bvh)
    try:
        (lx, ly, lz), (hx, hy, hz) = h.bounds
        return h.nodes.copy(), (lx, ly, lz, hx, hy, hz)
    finally:
        B.FreeHostBVH(h)
        B.FreeDeviceBVH(bvh)


def check_same(d, n, cfg, name):
    say("noalloc", name, n)
    want, box = reference(d, n, cfg)
    st, out, nodes = noalloc(d, n, cfg)
    assert st == 0, (name, native.STATUS.get(st, st))
    assert out.nodes == nodes.address and out.node_count == 2 * n - 1 and out.prim_count == n
    got = nodes.to_array(np.uint8, (2 * n - 1) * 32).view(B.BVH2_NODE_DTYPE)
    assert got.tobytes() == want.tobytes(), name
    assert (out.bounds.bmin.x, out.bounds.bmin.y, out.bounds.bmin.z, out.bounds.bmax.x, out.bounds.bmax.y,
            out.bounds.bmax.z) == box


def test_noalloc_equals_build_bvh2_two_calls_second_smaller():
    big = scenes.uniform(400_000, seed=91)
    small = scenes.slivers(70_000, seed=92)
    db, ds = DevicePrimitives.from_numpy(big), DevicePrimitives.from_numpy(small)
    for cfg in (BuildConfig(), BuildConfig(mortonBits=32, mergeMode=MergeMode.WAVE64)):
        check_same(db, len(big), cfg, "uniform 400k")
        check_same(ds, len(small), cfg, "slivers 70k")


def test_noalloc_at_the_size_limit_and_aabb_input():
    tris = scenes.uniform(NOALLOC_MAX, seed=93)
    check_same(DevicePrimitives.from_numpy(tris), NOALLOC_MAX, BuildConfig(), "uniform at the limit")
    boxes = scenes.boxes_of(tris[:1000])
    check_same(DevicePrimitives.from_numpy(boxes, PrimType.AABB), 1000, BuildConfig(mortonBits=32), "aabb 1000")
    for n in (1, 2, 3):
        check_same(DevicePrimitives.from_numpy(tris[:n]), n, BuildConfig(), f"tiny {n}")


def test_noalloc_rejects_sizes_beyond_its_scratch():
    st, out, _ = noalloc(DevicePrimitives.from_numpy(scenes.uniform(16, seed=94)), NOALLOC_MAX + 1,
                         BuildConfig())
    assert native.STATUS[st] == "NXB_ERROR_TOO_LARGE"
    assert out.nodes is None and out.node_count == 0
