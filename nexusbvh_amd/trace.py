"""Closest-hit ray traversal of a device BVH8 (no counterpart in the
reference, which builds BVH8s but does not trace them).

`TraceClosest` runs `nxb_trace_closest` (csrc/trace.hip) through the C ABI:
one GPU lane per ray, Moller-Trumbore in FP32 without contraction, equal
distances resolved to the smaller triangle index. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .builder import BVH8, DeviceIndexedMesh, DevicePrimitives, PrimType
from .native import DeviceBuffer, check, lib, synchronize

RAY_DTYPE = np.dtype([("origin", "<f4", (3,)), ("tmin", "<f4"), ("dir", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])
INVALID = 0xFFFFFFFF


def make_rays(origins: np.ndarray, dirs: np.ndarray, tmin: float = 0.0, tmax: float = np.inf) -> np.ndarray:
    n = len(origins)
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["origin"] = origins
    r["dir"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def camera_rays(bounds, width: int, height: int, fov_deg: float = 40.0, seed: int = 0) -> np.ndarray:
    """Pinhole primary rays looking at the centre of `bounds` (lo, hi) from
    outside it, along a seeded oblique direction."""
    lo, hi = np.asarray(bounds[0], np.float32), np.asarray(bounds[1], np.float32)
    centre = (lo + hi) * 0.5
    radius = float(np.linalg.norm(hi - lo)) * 0.5
    rng = np.random.default_rng(seed)
    view = rng.normal(size=3)
    view /= np.linalg.norm(view)
    eye = centre - view * radius * 2.5
    up = np.array([0.0, 1.0, 0.0]) if abs(view[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    right = np.cross(view, up)
    right /= np.linalg.norm(right)
    up = np.cross(right, view)
    half = np.tan(np.radians(fov_deg) * 0.5)
    ys, xs = np.meshgrid((np.arange(height) + 0.5) / height * 2 - 1, (np.arange(width) + 0.5) / width * 2 - 1,
                         indexing="ij")
    d = view[None, :] + (xs.reshape(-1, 1) * half) * right[None, :] + (ys.reshape(-1, 1) * half) * up[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(np.repeat(eye[None, :], len(d), axis=0).astype(np.float32), d.astype(np.float32))


def random_rays(bounds, count: int, seed: int = 0) -> np.ndarray:
    """Origins uniform in `bounds` grown by half its size, directions uniform
    on the sphere."""
    lo, hi = np.asarray(bounds[0], np.float64), np.asarray(bounds[1], np.float64)
    ext = hi - lo
    rng = np.random.default_rng(seed)
    o = lo - ext * 0.5 + rng.random((count, 3)) * ext * 2.0
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o.astype(np.float32), d.astype(np.float32))


def TraceClosest(bvh: BVH8, tris, rays, with_stats: bool = False):
    """Closest hits of `rays` (RAY_DTYPE array or a DeviceBuffer of them)
    against the triangles `bvh` was built from (DevicePrimitives of
    triangles, or the DeviceIndexedMesh of an indexed build). Returns a
    HIT_DTYPE array, and with with_stats also (nodes visited, triangles
    tested) per ray."""
    indexed = isinstance(tris, DeviceIndexedMesh)
    if not indexed and tris.prim_type != PrimType.TRIANGLE:
        raise ValueError("traversal needs triangle primitives")
    if isinstance(rays, DeviceBuffer):
        rbuf, count = rays, rays.nbytes // RAY_DTYPE.itemsize
    else:
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        rbuf, count = DeviceBuffer.from_array(rays.view(np.uint8)), len(rays)
    hits = DeviceBuffer(max(1, count) * HIT_DTYPE.itemsize)
    stats = DeviceBuffer(max(1, count) * 8) if with_stats else None
    if indexed:
        check(lib().nxb_trace_closest_indexed(ctypes.byref(bvh), tris.vertices.address, tris.vertexCount,
                                              tris.indices.address, rbuf.address, count, hits.address,
                                              stats.address if stats else None, None), "TraceClosest")
    else:
        check(lib().nxb_trace_closest(ctypes.byref(bvh), tris.address, rbuf.address, count, hits.address,
                                      stats.address if stats else None, None), "TraceClosest")
    synchronize()
    h = hits.to_array(np.uint8, count * HIT_DTYPE.itemsize).view(HIT_DTYPE)
    if not with_stats:
        return h
    return h, stats.to_array(np.uint32, 2 * count).reshape(count, 2)
