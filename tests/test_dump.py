"""Versioned BVH dumps (host only): round trips, format check, diffs."""
import numpy as np
import pytest

from nexusbvh_amd import dump
from nexusbvh_amd.builder import BVH2_NODE_DTYPE, BVH8_NODE_DTYPE, HostBVH2, HostBVH8


def _bvh8(seed, nodes=37, prims=100):
    rng = np.random.default_rng(seed)
    raw = rng.integers(0, 256, size=nodes * BVH8_NODE_DTYPE.itemsize, dtype=np.uint8)
    return HostBVH8(raw.view(BVH8_NODE_DTYPE).copy(), rng.permutation(prims).astype(np.uint32), nodes, prims,
                    ((0.0, -1.0, 2.0), (1.0, 3.0, 4.5)))


def _bvh2(seed, prims=50):
    rng = np.random.default_rng(seed)
    raw = rng.integers(0, 256, size=(2 * prims - 1) * BVH2_NODE_DTYPE.itemsize, dtype=np.uint8)
    return HostBVH2(raw.view(BVH2_NODE_DTYPE).copy(), 2 * prims - 1, prims, ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))


def test_bvh8_round_trip(tmp_path):
    a = _bvh8(1)
    p = str(tmp_path / "a.npz")
    dump.save(p, a, workload="unit", radius=8)
    b, meta = dump.load(p)
    assert isinstance(b, HostBVH8)
    assert b.nodes.tobytes() == a.nodes.tobytes() and np.array_equal(b.primIdx, a.primIdx)
    assert b.bounds == a.bounds and meta == {"workload": "unit", "radius": 8}
    assert dump.first_difference(a, b) is None


def test_bvh2_round_trip(tmp_path):
    a = _bvh2(2)
    p = str(tmp_path / "a.npz")
    dump.save(p, a)
    b, meta = dump.load(p)
    assert isinstance(b, HostBVH2) and b.primCount == 50 and meta == {}
    assert b.nodes.tobytes() == a.nodes.tobytes() and dump.first_difference(a, b) is None


def test_first_difference():
    a = _bvh8(3)
    b = HostBVH8(a.nodes.copy(), a.primIdx.copy(), a.nodeCount, a.primCount, a.bounds)
    b.nodes["meta"][17, 3] ^= 1
    assert dump.first_difference(a, b) == 17
    c = HostBVH8(a.nodes[:20].copy(), a.primIdx.copy(), 20, a.primCount, a.bounds)
    assert dump.first_difference(a, c) == 20
    d = HostBVH8(a.nodes.copy(), a.primIdx[::-1].copy(), a.nodeCount, a.primCount, a.bounds)
    assert dump.first_difference(a, d) == -1


def test_unknown_format_is_refused(tmp_path):
    p = str(tmp_path / "x.npz")
    np.savez(p, format=np.asarray("nxb-bvh8-v0"), nodes=np.zeros(80, np.uint8), bounds=np.zeros(6, np.float32))
    with pytest.raises(ValueError):
        dump.load(p)
