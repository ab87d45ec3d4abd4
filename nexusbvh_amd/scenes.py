"""Synthetic workloads (libnxb_scenes.so): numpy float32 arrays of shape (n, 9)
(triangles v0, v1, v2) for the five measured configurations.

  uniform_300k     300,000 triangles, centroid U[0,1)^3, vertices +-0.005
  icosphere_1m1    1,100,000 triangles of a frequency-235 geodesic icosphere,
                   radius 1, radial 4-octave value-noise displacement 0.02
  slivers_2m9      2,900,000 hair-like slivers (length U(0.02,0.08), width 1e-4)
  batch_2048       2,048 meshes, log-uniform counts in [64, 65536] (~19M total)
  blobs_28m1       28,100,000 triangles: 64 Gaussian blobs + 0.01 ground-plane grid
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_float, c_uint32, c_uint64, c_void_p

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_PKG, "lib", "libnxb_scenes.so")
_lib = None

DEFAULT_SEED = 0x5EED2024


def _l():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"{_PATH} is not built")
        l = ctypes.CDLL(_PATH)
        l.nxbs_uniform.argtypes = [c_uint32, c_uint64, c_float, c_void_p]
        l.nxbs_icosphere.argtypes = [c_uint32, c_uint32, c_uint64, c_float, c_void_p]
        l.nxbs_icosphere.restype = c_uint32
        l.nxbs_slivers.argtypes = [c_uint32, c_uint64, c_float, c_void_p]
        l.nxbs_batch_counts.argtypes = [c_uint32, c_uint64, c_uint32, c_uint32, c_void_p]
        l.nxbs_batch_counts.restype = c_uint64
        l.nxbs_batch_mesh.argtypes = [c_uint32, c_uint32, c_uint64, c_void_p]
        l.nxbs_batch_mesh.restype = c_uint32
        l.nxbs_blobs_plane.argtypes = [c_uint32, c_uint32, c_uint32, c_float, c_uint64, c_void_p]
        l.nxbs_blobs_plane.restype = c_uint32
        _lib = l
    return _lib


def _buf(n: int) -> np.ndarray:
    return np.empty((n, 9), dtype=np.float32)


def uniform(n: int, seed: int = DEFAULT_SEED, half_extent: float = 0.005) -> np.ndarray:
    out = _buf(n)
    _l().nxbs_uniform(n, seed, half_extent, out.ctypes.data)
    return out


def icosphere(n: int, freq: int = 235, seed: int = DEFAULT_SEED, amplitude: float = 0.02) -> np.ndarray:
    total = 20 * freq * freq
    n = min(n, total)
    out = _buf(n)
    got = _l().nxbs_icosphere(n, freq, seed, amplitude, out.ctypes.data)
    return out[:got]


def slivers(n: int, seed: int = DEFAULT_SEED, width: float = 1e-4) -> np.ndarray:
    out = _buf(n)
    _l().nxbs_slivers(n, seed, width, out.ctypes.data)
    return out


def batch_counts(meshes: int = 2048, seed: int = DEFAULT_SEED, lo: int = 64, hi: int = 65536) -> np.ndarray:
    c = np.empty(meshes, dtype=np.uint32)
    _l().nxbs_batch_counts(meshes, seed, lo, hi, c.ctypes.data)
    return c


def batch_mesh(mesh: int, count: int, seed: int = DEFAULT_SEED) -> np.ndarray:
    out = _buf(count)
    _l().nxbs_batch_mesh(mesh, count, seed, out.ctypes.data)
    return out


def batch(meshes: int = 2048, seed: int = DEFAULT_SEED, lo: int = 64, hi: int = 65536):
    """All meshes of the batched workload concatenated: (triangles, offsets[meshes+1])."""
    counts = batch_counts(meshes, seed, lo, hi)
    offsets = np.zeros(meshes + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(counts, dtype=np.uint64)
    tris = np.empty((int(offsets[-1]), 9), dtype=np.float32)
    for m in range(meshes):
        a, b = int(offsets[m]), int(offsets[m + 1])
        _l().nxbs_batch_mesh(m, b - a, seed, tris[a:b].ctypes.data)
    return tris, offsets


def blobs_plane(n: int = 28_100_000, seed: int = DEFAULT_SEED, blobs: int = 64, side: int = 2053,
                cell: float = 0.01) -> np.ndarray:
    out = _buf(n)
    _l().nxbs_blobs_plane(n, blobs, side, cell, seed, out.ctypes.data)
    return out


WORKLOADS = {
    "uniform_300k": lambda seed=DEFAULT_SEED: uniform(300_000, seed),
    "icosphere_1m1": lambda seed=DEFAULT_SEED: icosphere(1_100_000, 235, seed),
    "slivers_2m9": lambda seed=DEFAULT_SEED: slivers(2_900_000, seed),
    "blobs_28m1": lambda seed=DEFAULT_SEED: blobs_plane(28_100_000, seed),
}


def workload(name: str, seed: int = DEFAULT_SEED) -> np.ndarray:
    return WORKLOADS[name](seed)


def boxes_of(tris: np.ndarray) -> np.ndarray:
    """(n, 6) AABB primitives [min xyz, max xyz] of (n, 9) triangles."""
    v = tris.reshape(-1, 3, 3)
    return np.concatenate([v.min(axis=1), v.max(axis=1)], axis=1).astype(np.float32)
