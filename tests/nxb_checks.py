"""Size-independent structural checks used by the GPU tests (vectorized numpy)."""
import numpy as np

INV = 0xFFFFFFFF


def bvh2_invariants(nodes, n, prim_boxes, scene):
    """Leaves in input order; every node but the root is a child exactly once;
    each internal box is the exact union of its children; root box = scene."""
    nd = nodes[: 2 * n - 1]
    assert np.all(nd["left"][:n] == INV) and np.array_equal(nd["right"][:n], np.arange(n, dtype=np.uint32))
    assert np.array_equal(nd["bmin"][:n], prim_boxes[:, :3]) and np.array_equal(nd["bmax"][:n], prim_boxes[:, 3:])
    if n == 1:
        return
    inner = nd[n:]
    l = inner["left"].astype(np.int64)
    r = inner["right"].astype(np.int64)
    assert l.max() < 2 * n - 1 and r.max() < 2 * n - 1
    cnt = np.bincount(np.concatenate([l, r]), minlength=2 * n - 1)
    assert cnt[2 * n - 2] == 0 and np.all(cnt[: 2 * n - 2] == 1)
    assert np.array_equal(inner["bmin"], np.minimum(nd["bmin"][l], nd["bmin"][r]))
    assert np.array_equal(inner["bmax"], np.maximum(nd["bmax"][l], nd["bmax"][r]))
    root = nd[2 * n - 2]
    assert np.array_equal(np.concatenate([root["bmin"], root["bmax"]]), np.asarray(scene, dtype=np.float32))


def bvh8_invariants(b8, prim_idx, n):
    """Every primitive in exactly one leaf slot; inner slots point to distinct
    nodes 1..count-1 (each once); meta bytes consistent with imask."""
    count = len(b8)
    meta = b8["meta"].astype(np.int64)                   # (count, 8)
    imask = b8["imask"].astype(np.int64)[:, None]
    slot = np.arange(8)[None, :]
    inner = ((imask >> slot) & 1).astype(bool)
    used = meta != 0
    assert np.all(~inner | used)
    assert np.all(meta[inner] == (0x20 | (24 + np.broadcast_to(slot, meta.shape)[inner])))
    leaf = used & ~inner
    leaf_rank = np.cumsum(leaf, axis=1) - leaf
    assert np.all((meta[leaf] & 31) == leaf_rank[leaf])
    assert np.all(meta[leaf] >> 5 == 1)
    # inner children: child_base + rank among inner slots
    inner_rank = np.cumsum(inner, axis=1) - inner
    child = (b8["child_base"].astype(np.int64)[:, None] + inner_rank)[inner]
    assert child.size == count - 1
    assert np.all(child >= 1) and np.all(child < count)
    assert np.all(np.bincount(child, minlength=count)[1:] == 1)
    # leaves: prim_base + rank
    pidx = (b8["prim_base"].astype(np.int64)[:, None] + leaf_rank)[leaf]
    assert pidx.size == n
    assert np.all(np.bincount(pidx, minlength=n) == 1)
    prims = prim_idx[pidx]
    assert np.all(np.bincount(prims.astype(np.int64), minlength=n) == 1)
    # children per node 2..8 (1 only for a single-primitive BVH)
    per = used.sum(axis=1)
    assert per.max() <= 8 and (per.min() >= 2 or n == 1)
