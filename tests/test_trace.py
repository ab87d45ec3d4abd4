"""GPU parity of the BVH8 ray traversal (nxb_trace_closest) against the
oracle's brute-force closest hit over all triangles.

Bar: for every ray the same triangle index and bit-identical t, u and v (the
kernel and the oracle evaluate Moller-Trumbore in FP32 in the same order
without contraction, and resolve equal distances to the smaller index), so a
single box culled by the traversal that held the closest hit fails the test.
"""
import numpy as np
import pytest

from nexusbvh_amd import builder as B
from nexusbvh_amd import native, scenes
from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, MergeMode
from nexusbvh_amd.trace import INVALID, TraceClosest, camera_rays, make_rays, random_rays
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if native.device_count() == 0:
        pytest.skip("no GPU")


def _bounds(tris):
    v = tris.reshape(-1, 3)
    return v.min(axis=0), v.max(axis=0)


def _check(tris, rays, cfg=None, name=""):
    d = DevicePrimitives.from_numpy(tris)
    bvh = B.BuildBVH8(d, len(tris), cfg or BuildConfig(mortonBits=32))
    try:
        got, stats = TraceClosest(bvh, d, rays, with_stats=True)
        node_count = bvh.node_count
    finally:
        B.FreeDeviceBVH(bvh)
    ref = O.trace_bruteforce(tris, rays)
    bad = np.nonzero((got["prim"] != ref["prim"]) | (got["t"].view(np.uint32) != ref["t"].view(np.uint32)) |
                     (got["u"].view(np.uint32) != ref["u"].view(np.uint32)) |
                     (got["v"].view(np.uint32) != ref["v"].view(np.uint32)))[0]
    hits = int((ref["prim"] != INVALID).sum())
    say("trace", name, len(tris), "tris", len(rays), "rays", hits, "hits", "mean nodes", float(stats[:, 0].mean()))
    assert bad.size == 0, (name, int(bad[0]), got[bad[0]], ref[bad[0]])
    assert stats[:, 0].max() <= node_count and (len(tris) == 0 or stats[:, 0].min() >= 1)
    return got, stats


def test_trace_uniform_random_rays():
    tris = scenes.uniform(20_000, seed=31, half_extent=0.02)
    _check(tris, random_rays(_bounds(tris), 4096, seed=1), name="uniform")


def test_trace_slivers_random_rays():
    tris = scenes.slivers(20_000, seed=32)
    _check(tris, random_rays(_bounds(tris), 4096, seed=2), name="slivers")


def test_trace_axis_aligned_rays():
    # directions with zero components take the reciprocal fallback
    tris = scenes.uniform(8_000, seed=33, half_extent=0.03)
    lo, hi = _bounds(tris)
    rng = np.random.default_rng(3)
    axes = np.eye(3, dtype=np.float32)
    dirs = np.concatenate([axes, -axes])[rng.integers(0, 6, 3000)]
    origins = (lo + rng.random((3000, 3)) * (hi - lo)).astype(np.float32)
    origins[dirs[:, 0] > 0, 0] = lo[0] - 0.1
    origins[dirs[:, 0] < 0, 0] = hi[0] + 0.1
    _check(tris, make_rays(origins, dirs), name="axis")


def test_trace_wide_keys_and_wave64_tree():
    # another tree over the same triangles: the closest hits do not change
    tris = scenes.uniform(12_000, seed=34, half_extent=0.02)
    rays = random_rays(_bounds(tris), 2048, seed=4)
    a, _ = _check(tris, rays, BuildConfig(mortonBits=32), name="k4")
    b, _ = _check(tris, rays, BuildConfig(mortonBits=64, mergeMode=MergeMode.WAVE64, searchRadius=2), name="k8")
    assert np.array_equal(a, b)


def test_trace_auction_slots():
    # slot order changes the traversal order, never the closest hit
    tris = scenes.uniform(15_000, seed=37, half_extent=0.02)
    rays = random_rays(_bounds(tris), 2048, seed=8)
    a, _ = _check(tris, rays, BuildConfig(mortonBits=32), name="greedy")
    b, _ = _check(tris, rays, BuildConfig(mortonBits=32, slotAssignment=B.SlotAssignment.AUCTION), name="auction")
    assert np.array_equal(a, b)


def test_trace_ray_limits():
    tris = scenes.uniform(5_000, seed=35, half_extent=0.02)
    rays = random_rays(_bounds(tris), 2048, seed=5)
    rays["tmin"] = 0.05
    rays["tmax"] = np.random.default_rng(6).uniform(0.1, 1.0, len(rays)).astype(np.float32)
    _check(tris, rays, name="limits")


@pytest.mark.parametrize("n", [1, 2, 3, 9])
def test_trace_tiny_scenes(n):
    tris = scenes.uniform(n, seed=36 + n, half_extent=0.2)
    _check(tris, random_rays(_bounds(tris), 1024, seed=n), name=f"n={n}")


def test_trace_camera_rays_icosphere_1m1():
    tris = scenes.workload("icosphere_1m1")
    rays = camera_rays(_bounds(tris), 16, 16, seed=7)
    got, stats = _check(tris, rays, name="icosphere camera")
    assert (got["prim"] != INVALID).sum() > 0


def _shared_vertices(tris):
    v = tris.reshape(-1, 3)
    verts, inv = np.unique(v, axis=0, return_inverse=True)
    return verts.astype(np.float32), inv.reshape(-1, 3).astype(np.uint32)


def test_trace_indexed_mesh():
    """nxb_trace_closest_indexed on the BVH8 of an indexed mesh gives the
    brute-force hits of the expanded triangles, as the soup path does."""
    tris = scenes.icosphere(20_000, freq=40)
    verts, idx = _shared_vertices(tris)
    assert len(verts) < len(tris) * 3 // 4
    mesh = B.DeviceIndexedMesh.from_numpy(verts, idx)
    rays = camera_rays(_bounds(tris), 48, 48, seed=9)
    bvh = B.BuildBVH8Indexed(mesh, BuildConfig(mortonBits=32))
    try:
        got, stats = TraceClosest(bvh, mesh, rays, with_stats=True)
    finally:
        B.FreeDeviceBVH(bvh)
    ref = O.trace_bruteforce(tris, rays)
    say("trace indexed", len(tris), "tris", len(verts), "verts", int((ref["prim"] != INVALID).sum()), "hits")
    assert (ref["prim"] != INVALID).sum() > 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    soup, _ = _check(tris, rays, BuildConfig(mortonBits=32), name="soup")
    assert np.array_equal(got, soup)


def test_trace_indexed_out_of_range_index_is_never_hit():
    tris = scenes.icosphere(5_000, freq=20)
    verts, idx = _shared_vertices(tris)
    rays = camera_rays(_bounds(tris), 32, 32, seed=10)
    ref = O.trace_bruteforce(tris, rays)
    k = int(np.bincount(ref["prim"][ref["prim"] != INVALID]).argmax())  # the most-hit triangle
    mesh = B.DeviceIndexedMesh.from_numpy(verts, idx)
    bvh = B.BuildBVH8Indexed(mesh, BuildConfig(mortonBits=32))
    bad = idx.copy()
    bad[k, 1] = len(verts)  # past the vertex buffer, after the build
    try:
        got = TraceClosest(bvh, B.DeviceIndexedMesh.from_numpy(verts, bad), rays)
    finally:
        B.FreeDeviceBVH(bvh)
    gone = tris.copy()
    gone[k] = np.nan  # a NaN triangle is never hit by the brute force
    assert np.array_equal(got.view(np.uint32), O.trace_bruteforce(gone, rays).view(np.uint32))
    assert not (got["prim"] == k).any()


def _grid_rounding_scene():
    """Three triangles whose BVH8 root has p.x = -1000 and an x grid step of 8.
    Triangle 0's min x is -2^-20: fl(-2^-20 - (-1000)) rounds to exactly 1000,
    a grid point, so its decoded box starts at x = 0, right of the triangle.
    Rays along +y at -2^-20 < x < 0 hit triangle 0 inside that sliver."""
    e = np.float32(2.0 ** -20)
    tris = np.array([[-e, 1, -1, -e, 1, 1, 5, 1, 0],
                     [-1000, 50, 50, -999, 50, 50, -1000, 51, 50],
                     [1000, -50, -50, 999, -50, -50, 1000, -51, -50]], dtype=np.float32)
    xs = np.array([-2.0 ** -21, -2.0 ** -22, -3 * 2.0 ** -23, -2.0 ** -30], dtype=np.float32)
    zs = np.array([-0.5, 0.0, 0.25], dtype=np.float32)
    org = np.array([[x, -10, z] for x in xs for z in zs], dtype=np.float32)
    dirs = np.tile(np.array([0, 1, 0], dtype=np.float32), (len(org), 1))
    return tris, make_rays(org, dirs)


def test_trace_child_face_rounded_onto_the_grid():
    """A child face that the builder's fl(c - p) moves onto a grid point (and
    so inward) must still be hit: the widening has to scale with |p|, not
    only with the decoded coordinate (which is 0 here)."""
    tris, rays = _grid_rounding_scene()
    got, _ = _check(tris, rays, cfg=BuildConfig(mortonBits=32), name="grid_rounding")
    assert np.all(got["prim"] == 0)
