"""Stage-level access to libnxb.so (one call per pipeline stage), device
buffers in and out. Each wraps one stage entry point of include/nxb.h."""
from __future__ import annotations

import ctypes

import numpy as np

from .builder import BVH2_NODE_DTYPE, BVH8_NODE_DTYPE, BuildConfig, DevicePrimitives
from .native import DeviceBuffer, check, lib, synchronize

INVALID = 0xFFFFFFFF


def bvh8_capacity(n: int) -> int:
    return int(lib().nxb_bvh8_capacity(n))


def scene_bounds(prims: DevicePrimitives):
    """Leaf nodes (buffer sized for 2n-1 nodes), scene box buffer, flags word."""
    n = prims.count
    nodes = DeviceBuffer(max(1, 2 * n - 1) * 32)
    scene = DeviceBuffer(24)
    flags = DeviceBuffer.from_array(np.zeros(1, dtype=np.uint32))
    check(lib().nxb_compute_scene_bounds(prims.address, int(prims.prim_type), n, nodes.address, scene.address,
                                         flags.address, None), "scene_bounds")
    synchronize()
    return nodes, scene, int(flags.to_array(np.uint32)[0])


def morton(nodes: DeviceBuffer, n: int, scene: DeviceBuffer, bits: int, with_values: bool = False):
    keys = DeviceBuffer(max(1, n) * (bits // 8))
    vals = DeviceBuffer(max(1, n) * 4) if with_values else None
    check(lib().nxb_compute_morton(nodes.address, n, scene.address, bits, keys.address,
                                   vals.address if vals else None, None), "morton")
    synchronize()
    return keys, vals


def sort_pairs(keys: DeviceBuffer, n: int, bits: int, begin_bit: int, end_bit: int, digit_bits: int = 8,
               values: DeviceBuffer | None = None):
    ko = DeviceBuffer(max(1, n) * (bits // 8))
    vo = DeviceBuffer(max(1, n) * 4)
    check(lib().nxb_radix_sort_pairs(keys.address, values.address if values else None, ko.address, vo.address, n,
                                     bits, begin_bit, end_bit, digit_bits, None, 0, None), "radix_sort_pairs")
    synchronize()
    return ko, vo


def hploc(sorted_keys: DeviceBuffer, sorted_values: DeviceBuffer, n: int, bits: int, nodes: DeviceBuffer,
          config: BuildConfig | None = None) -> None:
    cfg = (config or BuildConfig()).to_c()
    check(lib().nxb_build_hploc(sorted_keys.address, sorted_values.address, n, bits, ctypes.byref(cfg),
                                nodes.address, None, 0, None), "build_hploc")


def collapse(nodes: DeviceBuffer, n: int, spin_limit: int = 0):
    cap = max(1, bvh8_capacity(n))
    b8 = DeviceBuffer(cap * 80)
    prim = DeviceBuffer(max(1, n) * 4)
    count = ctypes.c_uint32(0)
    st = lib().nxb_collapse_bvh8(nodes.address, n, b8.address, prim.address, ctypes.byref(count), spin_limit,
                                 None, 0, None)
    return st, b8, prim, int(count.value)


def nodes_to_host(buf: DeviceBuffer, count: int) -> np.ndarray:
    return buf.to_array(np.uint8, count * 32).view(BVH2_NODE_DTYPE)


def bvh8_to_host(buf: DeviceBuffer, count: int) -> np.ndarray:
    return buf.to_array(np.uint8, count * 80).view(BVH8_NODE_DTYPE)
