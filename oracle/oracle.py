"""Python access to the CPU oracle (oracle/lib/libnxb_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker / timed CPU baseline, never by the GPU product
path. See nxb_oracle.cpp for the restated reference functions and their
reference citations.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "lib", "libnxb_oracle.so")
_lib = None

BVH2_NODE_DTYPE = np.dtype([("bmin", "<f4", (3,)), ("bmax", "<f4", (3,)), ("left", "<u4"), ("right", "<u4")])
BVH8_NODE_DTYPE = np.dtype([("p", "<f4", (3,)), ("e", "u1", (3,)), ("imask", "u1"), ("child_base", "<u4"),
                            ("prim_base", "<u4"), ("meta", "u1", (8,)), ("qlox", "u1", (8,)), ("qloy", "u1", (8,)),
                            ("qloz", "u1", (8,)), ("qhix", "u1", (8,)), ("qhiy", "u1", (8,)), ("qhiz", "u1", (8,))])
INVALID = 0xFFFFFFFF


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"{_PATH} is not built")
        l = ctypes.CDLL(_PATH)
        l.nxbo_scene_bounds.argtypes = [c_void_p, c_int, c_uint32, c_void_p, c_void_p]
        l.nxbo_trace_bruteforce.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p]
        l.nxbo_trace_bruteforce.restype = None
        l.nxbo_scene_bounds.restype = c_int
        l.nxbo_morton.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p]
        l.nxbo_morton_point.argtypes = [c_float, c_float, c_float, c_uint32]
        l.nxbo_morton_point.restype = c_uint64
        l.nxbo_sort_pairs.argtypes = [c_void_p, c_uint32, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p]
        l.nxbo_find_nn.argtypes = [c_uint32, c_uint32, c_uint32, c_void_p, c_void_p]
        l.nxbo_hploc.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_uint32, c_uint32, c_void_p,
                                 c_uint32, c_uint32]
        l.nxbo_hploc.restype = c_uint32
        l.nxbo_collapse_bvh8.argtypes = [c_void_p, c_uint32, c_void_p, c_void_p]
        l.nxbo_collapse_bvh8.restype = c_uint32
        l.nxbo_collapse_bvh8_slots.argtypes = [c_void_p, c_uint32, c_void_p, c_void_p, c_uint32]
        l.nxbo_collapse_bvh8_slots.restype = c_uint32
        l.nxbo_auction_assign.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p]
        l.nxbo_auction_assign.restype = c_uint32
        l.nxbo_sah_bvh2.argtypes = [c_void_p, c_uint32, c_void_p]
        l.nxbo_sah_bvh2.restype = c_double
        l.nxbo_sah_bvh8.argtypes = [c_void_p, c_uint32, c_void_p]
        l.nxbo_sah_bvh8.restype = c_double
        l.nxbo_compare_bvh2.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_uint32, POINTER(c_uint64)]
        l.nxbo_compare_bvh2.restype = c_int
        l.nxbo_compare_bvh8.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, c_uint32,
                                        POINTER(c_uint64)]
        l.nxbo_compare_bvh8.restype = c_int
        l.nxbo_build.argtypes = [c_void_p, c_int, c_uint32, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]
        l.nxbo_build.restype = c_uint32
        _lib = l
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def bvh8_capacity(n: int) -> int:
    return 0 if n == 0 else 1 + (4 * n - 2) // 7


def scene_bounds(prims: np.ndarray, prim_type: int):
    """prims: (n, 9) triangles (type 1) or (n, 6) boxes (type 0).
    Returns (leaf nodes[n], scene box float32[6], nonfinite flag)."""
    p = np.ascontiguousarray(prims, dtype=np.float32)
    n = p.shape[0]
    nodes = np.zeros(max(1, 2 * n - 1), dtype=BVH2_NODE_DTYPE)
    scene = np.zeros(6, dtype=np.float32)
    bad = lib().nxbo_scene_bounds(_ptr(p), prim_type, n, _ptr(nodes), _ptr(scene))
    return nodes, scene, bool(bad)


def morton(nodes: np.ndarray, n: int, scene: np.ndarray, bits: int) -> np.ndarray:
    keys = np.zeros(n, dtype=np.uint32 if bits == 32 else np.uint64)
    lib().nxbo_morton(_ptr(nodes), n, _ptr(scene), bits, _ptr(keys))
    return keys


def morton_point(x: float, y: float, z: float, bits: int) -> int:
    return int(lib().nxbo_morton_point(x, y, z, bits))


def sort_pairs(keys: np.ndarray, begin_bit: int, end_bit: int):
    """Stable sort of (key, iota) on bits [begin_bit, end_bit): (sorted keys, permutation)."""
    bits = 32 if keys.dtype == np.uint32 else 64
    k = np.ascontiguousarray(keys)
    out = np.empty_like(k)
    vals = np.empty(k.shape[0], dtype=np.uint32)
    lib().nxbo_sort_pairs(_ptr(k), k.shape[0], bits, begin_bit, end_bit, _ptr(out), _ptr(vals))
    return out, vals


def sort_window(bits: int) -> tuple[int, int]:
    """Sorted key bits of the pipeline (Setup.cu:74-78)."""
    return (2, 32) if bits == 32 else (1, 64)


def find_nn(boxes: np.ndarray, width: int = 32, radius: int = 8) -> np.ndarray:
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    out = np.zeros(64, dtype=np.uint32)
    lib().nxbo_find_nn(b.shape[0], width, radius, _ptr(b), _ptr(out))
    return out[:width]


def hploc(sorted_keys: np.ndarray, sorted_vals: np.ndarray, leaves: np.ndarray, width: int = 32, radius: int = 8,
          table: np.ndarray | None = None, seg_log2: int = 12) -> np.ndarray:
    """H-PLOC over sorted keys. leaves: node array whose first n entries are the leaves.
    Returns the full node array (2n-1)."""
    n = sorted_keys.shape[0]
    bits = 32 if sorted_keys.dtype == np.uint32 else 64
    nodes = np.zeros(max(1, 2 * n - 1), dtype=BVH2_NODE_DTYPE)
    nodes[:n] = leaves[:n]
    k = np.ascontiguousarray(sorted_keys)
    v = np.ascontiguousarray(sorted_vals, dtype=np.uint32)
    t = None if table is None else np.ascontiguousarray(table, dtype=np.uint8)
    lib().nxbo_hploc(_ptr(k), bits, _ptr(v), n, _ptr(nodes), width, radius, None if t is None else _ptr(t),
                     seg_log2, 0 if t is None else t.shape[0])
    return nodes


def collapse_bvh8(nodes: np.ndarray, n: int, slot_assignment: int = 0):
    """BVH8 of a BVH2 (WideConverter.cu:225-414); slot_assignment 0 greedy, 1 auction."""
    out = np.zeros(max(1, bvh8_capacity(n)), dtype=BVH8_NODE_DTYPE)
    prim = np.zeros(max(1, n), dtype=np.uint32)
    if slot_assignment == 0:
        c = lib().nxbo_collapse_bvh8(_ptr(nodes), n, _ptr(out), _ptr(prim))
    else:
        c = lib().nxbo_collapse_bvh8_slots(_ptr(nodes), n, _ptr(out), _ptr(prim), slot_assignment)
    return out[:c], prim[:n]


def auction_assign(parent6, children6, with_prices: bool = False):
    """Auction slot of each child (WideConverter.cu:52-102): returns slot[c]
    for c < len(children6) (and the 8 final slot prices); boxes are
    (min xyz, max xyz)."""
    p = np.ascontiguousarray(parent6, dtype=np.float32).reshape(6)
    c = np.ascontiguousarray(children6, dtype=np.float32).reshape(-1, 6)
    prices = np.zeros(8, dtype=np.float32)
    a = lib().nxbo_auction_assign(_ptr(p), _ptr(c), c.shape[0], _ptr(prices))
    slot = np.full(c.shape[0], 0xF, dtype=np.int64)
    for s in range(8):
        k = (a >> (4 * s)) & 0xF
        if k != 0xF:
            slot[k] = s
    return (slot, prices) if with_prices else slot


def sah_bvh2(nodes: np.ndarray, count: int, scene: np.ndarray) -> float:
    return float(lib().nxbo_sah_bvh2(_ptr(nodes), count, _ptr(np.ascontiguousarray(scene, dtype=np.float32))))


def sah_bvh8(nodes: np.ndarray, count: int, scene: np.ndarray) -> float:
    return float(lib().nxbo_sah_bvh8(_ptr(nodes), count, _ptr(np.ascontiguousarray(scene, dtype=np.float32))))


def compare_bvh2(a: np.ndarray, b: np.ndarray, n: int) -> tuple[int, int]:
    """Ordered-tree comparison from both roots (2n-2). (0, visits) when equal."""
    where = c_uint64(0)
    root = 2 * n - 2
    code = lib().nxbo_compare_bvh2(_ptr(a), root, _ptr(b), root, 2 * n - 1, ctypes.byref(where))
    return code, where.value


def compare_bvh8(a: np.ndarray, pa: np.ndarray, b: np.ndarray, pb: np.ndarray, n: int) -> tuple[int, int]:
    where = c_uint64(0)
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    code = lib().nxbo_compare_bvh8(_ptr(a), _ptr(pa), a.shape[0], _ptr(b), _ptr(pb), b.shape[0], n,
                                   ctypes.byref(where))
    return code, where.value


def trace_bruteforce(tris: np.ndarray, rays: np.ndarray) -> np.ndarray:
    """Closest hits by brute force (rays: 8 float32 per ray, origin/tmin/dir/tmax;
    returns a structured t/prim/u/v array)."""
    t = np.ascontiguousarray(tris, dtype=np.float32)
    r = np.ascontiguousarray(rays).view(np.float32).reshape(-1)
    count = r.size // 8
    out = np.zeros(count, dtype=np.dtype([("t", "<f4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")]))
    lib().nxbo_trace_bruteforce(_ptr(t), t.size // 9, _ptr(r), count, _ptr(out))
    return out


def build(prims: np.ndarray, prim_type: int = 1, bits: int = 64, width: int = 32, radius: int = 8,
          with_bvh8: bool = True):
    """Full CPU pipeline. Returns dict with nodes, scene, bvh8, prim_idx, stage_ms[5]."""
    p = np.ascontiguousarray(prims, dtype=np.float32)
    n = p.shape[0]
    nodes = np.zeros(max(1, 2 * n - 1), dtype=BVH2_NODE_DTYPE)
    scene = np.zeros(6, dtype=np.float32)
    b8 = np.zeros(max(1, bvh8_capacity(n)), dtype=BVH8_NODE_DTYPE) if with_bvh8 else None
    prim = np.zeros(max(1, n), dtype=np.uint32)
    ms = np.zeros(5, dtype=np.float64)
    c = lib().nxbo_build(_ptr(p), prim_type, n, bits, width, radius, _ptr(nodes), _ptr(scene),
                         None if b8 is None else _ptr(b8), _ptr(prim), _ptr(ms))
    return {"nodes": nodes, "scene": scene, "bvh8": None if b8 is None else b8[:c], "prim_idx": prim[:n],
            "stage_ms": ms}
