"""CPU checker of the learned-radius kernels (test infrastructure only).

The reference has no learned component; this module restates, in numpy, what
the gfx950 kernels of nexusbvh_amd/csrc/learned.hip compute, so the tests can
compare them:

  * mlp_forward / mlp_loss_grad: the 16 -> 64 -> 64 -> 4 ReLU MLP with the
    kernels' BF16 rounding points (operands rounded to BF16, FP32
    accumulation, FP32 biases) or, with emulate=False, in exact float64;
  * segment_features: the 16 per-segment features;
  * segment_sah: the per-segment SAH attribution of a BVH2;
  * radius_table: arg-max class -> radius with the confidence fallback.

Pins (tests/test_learned.py): the float64 gradient is checked against
central finite differences of the float64 loss; BF16 rounding against
hand-derived round-to-nearest-even cases; segment_sah sums to the internal
part of the oracle's total SAH; segment_features against hand-built
segments with known extents, areas and key prefixes.
"""
from __future__ import annotations

import numpy as np

FEATURES, HIDDEN, CLASSES = 16, 64, 4
PARAMS = FEATURES * HIDDEN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN * CLASSES + CLASSES
RADII = (2, 4, 8, 16)
_SHAPES = [("W1", (FEATURES, HIDDEN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
           ("W3", (HIDDEN, CLASSES)), ("b3", (CLASSES,))]


def unpack(p: np.ndarray) -> dict:
    """Parameter vector -> {W1, b1, W2, b2, W3, b3} (row-major [in][out])."""
    out, o = {}, 0
    for name, shape in _SHAPES:
        size = int(np.prod(shape))
        out[name] = p[o:o + size].reshape(shape)
        o += size
    assert o == PARAMS
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in _SHAPES])


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in _SHAPES:
        if name.startswith("W"):
            d[name] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def bf16(x) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _ident(x):
    return np.asarray(x, dtype=np.float64)


def mlp_forward(p: np.ndarray, X: np.ndarray, emulate: bool = True):
    """Logits (count x 4) and the activations (x, h1, h2)."""
    r = bf16 if emulate else _ident
    f = np.float32 if emulate else np.float64
    w = {k: np.asarray(v, dtype=f) for k, v in unpack(np.asarray(p, dtype=f)).items()}
    x = r(X).astype(f)
    h1 = r(np.maximum(x @ r(w["W1"]).astype(f) + w["b1"], 0)).astype(f)
    h2 = r(np.maximum(h1 @ r(w["W2"]).astype(f) + w["b2"], 0)).astype(f)
    z3 = h2 @ r(w["W3"]).astype(f) + w["b3"]
    return z3, (x, h1, h2, w)


def mlp_loss_grad(p: np.ndarray, X: np.ndarray, y: np.ndarray, emulate: bool = True):
    """Mean cross-entropy and its gradient (same layout as p)."""
    r = bf16 if emulate else _ident
    f = np.float32 if emulate else np.float64
    z3, (x, h1, h2, w) = mlp_forward(p, X, emulate)
    count = X.shape[0]
    m = z3.max(axis=1, keepdims=True)
    e = np.exp(z3 - m)
    s = e.sum(axis=1, keepdims=True)
    prob = e / s
    loss = float(np.mean((m[:, 0] - z3[np.arange(count), y]) + np.log(s[:, 0])))
    onehot = np.zeros_like(prob)
    onehot[np.arange(count), y] = 1
    d3 = ((prob - onehot) / count).astype(f)
    d3r = r(d3).astype(f)
    g = {"W3": h2.T @ d3r, "b3": d3.sum(axis=0)}
    dz2 = (d3r @ r(w["W3"]).astype(f).T) * (h2 > 0)
    g["b2"] = dz2.sum(axis=0)
    dz2r = r(dz2).astype(f)
    g["W2"] = h1.T @ dz2r
    dz1 = (dz2r @ r(w["W2"]).astype(f).T) * (h1 > 0)
    g["b1"] = dz1.sum(axis=0)
    g["W1"] = x.T @ r(dz1).astype(f)
    return loss, pack(g)


def radius_table(p: np.ndarray, X: np.ndarray, min_confidence: float):
    """(radius per segment, number of radius-8 fallbacks)."""
    z3, _ = mlp_forward(p, X, True)
    cls = np.argmax(z3, axis=1)
    m = z3.max(axis=1, keepdims=True)
    conf = 1.0 / np.exp(z3 - m).sum(axis=1)
    rad = np.array(RADII, dtype=np.uint8)[cls]
    fb = ~(conf >= min_confidence)
    rad[fb] = 8
    return rad, int(fb.sum())


# ---- segment features / SAH -------------------------------------------------

def _split_depth(x: np.ndarray, key_bits: int) -> np.ndarray:
    """Octree depth of the highest set bit of x (x > 0); 64 for x == 0."""
    top = 62 if key_bits == 64 else 29
    x = x.astype(np.uint64)
    hb = np.zeros(x.shape, dtype=np.int64)
    v = x.copy()
    for sh in (32, 16, 8, 4, 2, 1):
        m = v >= (np.uint64(1) << np.uint64(sh))
        hb[m] += sh
        v[m] >>= np.uint64(sh)
    d = np.maximum(top - hb, 0) // 3
    d[x == 0] = 64
    return d


def _depth_bin(d: np.ndarray, d0: int) -> np.ndarray:
    r = np.maximum(d - d0, 0)
    b = np.where(r < 4, r, np.where(r < 6, 4, np.where(r < 9, 5, 6)))
    b[d == 64] = 7
    return b


def segment_features(keys: np.ndarray, perm: np.ndarray, leaf_boxes: np.ndarray, seg_log2: int,
                     key_bits: int) -> np.ndarray:
    """(segments x 16) float64 features; leaf_boxes is (n x 6) lx ly lz hx hy hz."""
    n = keys.shape[0]
    segs = ((n - 1) >> seg_log2) + 1
    out = np.zeros((segs, FEATURES))
    f32 = np.float32
    for s in range(segs):
        lo, hi = s << seg_log2, min(n, (s + 1) << seg_log2)
        b = leaf_boxes[perm[lo:hi]].astype(f32)
        ext = (b[:, 3:] - b[:, :3]).astype(f32)
        a = (ext[:, 0] * ext[:, 1] + ext[:, 1] * ext[:, 2] + ext[:, 2] * ext[:, 0]).astype(f32)
        la = np.log2(np.maximum(a, f32(1e-30))).astype(np.float64)
        srt = np.sort(ext, axis=1)
        e0, e1 = srt[:, 2], srt[:, 1]
        sb_lo, sb_hi = b[:, :3].min(axis=0), b[:, 3:].max(axis=0)
        se = np.sort((sb_hi - sb_lo).astype(np.float64))
        seg_a = max(se[0] * se[1] + se[1] * se[2] + se[2] * se[0], 1e-30)
        cnt = hi - lo
        k = keys[lo:hi].astype(np.uint64)
        d0 = int(_split_depth(np.array([k[0] ^ k[-1]]), key_bits)[0])
        f = out[s]
        f[0] = np.log2(cnt) / 16
        f[1] = np.log2(max(se[1], 1e-30) / max(se[2], 1e-30)) / 24
        f[2] = np.log2(max(se[0], 1e-30) / max(se[2], 1e-30)) / 24
        f[3] = np.log2(seg_a / max(float(a.astype(np.float64).sum()), 1e-30)) / 16
        f[4] = (la.mean() - np.log2(seg_a)) / 48
        f[5] = np.sqrt(max((la * la).mean() - la.mean() ** 2, 0.0)) / 16
        f[6] = np.mean(e0 > 10 * e1)
        f[7] = np.mean(np.log2(np.maximum(e0, 1e-30) / np.maximum(e1, 1e-30))) / 24
        if cnt > 1:
            bins = _depth_bin(_split_depth(k[:-1] ^ k[1:], key_bits), d0)
            f[8:16] = np.bincount(bins, minlength=8)[:8] / (cnt - 1)
    return out


def segment_sah(boxes: np.ndarray, n: int, seg_log2: int, scene: np.ndarray) -> np.ndarray:
    """Per-segment SAH of the internal nodes n..2n-2 of a BVH2 (boxes: (2n-1) x 6):
    node id contributes float32(3 * (area / scene area)) to segment (id - n + 1) >> seg_log2."""
    f32 = np.float32

    def area(bx):
        d = (bx[..., 3:] - bx[..., :3]).astype(f32)
        return (d[..., 0] * d[..., 1] + d[..., 1] * d[..., 2] + d[..., 2] * d[..., 0]).astype(f32)

    sa = area(np.asarray(scene, dtype=f32).reshape(1, 6))[0]
    ids = np.arange(n, 2 * n - 1)
    terms = (f32(3.0) * (area(boxes[n:2 * n - 1].astype(f32)) / sa)).astype(np.float64)
    segs = ((n - 1) >> seg_log2) + 1
    return np.bincount((ids - n + 1) >> seg_log2, weights=terms, minlength=segs)
