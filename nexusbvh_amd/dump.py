"""Versioned on-disk dumps of host BVHs, for diffs across runs and machines
(no counterpart in the reference).

A dump is an uncompressed .npz holding the node array, primIdx (BVH8), the
root bounds and a `format` string ("nxb-bvh2-v1" / "nxb-bvh8-v1") that names
the node layout (include/nxb.h: 32-byte BVH2 nodes, 80-byte BVH8 nodes);
`load` refuses other formats. Loading uses numpy without pickles.
"""
from __future__ import annotations

import numpy as np

from .builder import BVH2_NODE_DTYPE, BVH8_NODE_DTYPE, HostBVH2, HostBVH8

FORMAT_BVH2 = "nxb-bvh2-v1"
FORMAT_BVH8 = "nxb-bvh8-v1"


def save(path: str, bvh, **meta) -> None:
    """Writes a HostBVH2 or HostBVH8; string/number keyword arguments are
    stored alongside (e.g. workload, build configuration)."""
    extra = {f"meta_{k}": np.asarray(v) for k, v in meta.items()}
    bounds = np.asarray(bvh.bounds, dtype=np.float32).reshape(6)
    if isinstance(bvh, HostBVH8):
        np.savez(path, format=np.asarray(FORMAT_BVH8), nodes=bvh.nodes.view(np.uint8), prim_idx=bvh.primIdx,
                 bounds=bounds, **extra)
    elif isinstance(bvh, HostBVH2):
        np.savez(path, format=np.asarray(FORMAT_BVH2), nodes=bvh.nodes.view(np.uint8),
                 prim_count=np.asarray(bvh.primCount), bounds=bounds, **extra)
    else:
        raise TypeError("expected HostBVH2 or HostBVH8")


def load(path: str):
    """Reads a dump back into a HostBVH2 or HostBVH8 (and its metadata)."""
    with np.load(path, allow_pickle=False) as z:
        fmt = str(z["format"])
        bounds = tuple(float(x) for x in z["bounds"])
        b = (bounds[:3], bounds[3:])
        meta = {k[5:]: z[k].item() for k in z.files if k.startswith("meta_")}
        if fmt == FORMAT_BVH8:
            nodes = z["nodes"].view(BVH8_NODE_DTYPE).copy()
            prims = z["prim_idx"].astype(np.uint32)
            return HostBVH8(nodes, prims, len(nodes), len(prims), b), meta
        if fmt == FORMAT_BVH2:
            nodes = z["nodes"].view(BVH2_NODE_DTYPE).copy()
            return HostBVH2(nodes, len(nodes), int(z["prim_count"]), b), meta
        raise ValueError(f"unknown BVH dump format {fmt!r}")


def first_difference(a, b) -> int | None:
    """Index of the first node whose bytes differ (or the shorter length when
    one array is a prefix of the other), -1 for primIdx differences only,
    None when equal."""
    na, nb = a.nodes.view(np.uint8).reshape(len(a.nodes), -1), b.nodes.view(np.uint8).reshape(len(b.nodes), -1)
    m = min(len(na), len(nb))
    diff = np.nonzero((na[:m] != nb[:m]).any(axis=1))[0]
    if diff.size:
        return int(diff[0])
    if len(na) != len(nb):
        return m
    if isinstance(a, HostBVH8) and not np.array_equal(a.primIdx, b.primIdx):
        return -1
    return None
