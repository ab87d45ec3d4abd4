"""Learned per-segment H-PLOC search radius: data, training and use.

The sorted primitive array is cut into segments of 2^segment_log2 positions;
H-PLOC takes each merge task's radius from the segment of its first sorted
position. A 16 -> 64 -> 64 -> 4 MLP (gfx950 kernels of csrc/learned.hip,
BF16 on MFMA with FP32 accumulation) maps 16 segment features to one of the
radii 2, 4, 8, 16; below a confidence threshold the segment uses radius 8.

Training data are built by the method itself: every scene of the generator
family (Gaussian blobs over a ground-plane grid, as the 28.1M workload) is
sorted once and built with H-PLOC at each radius; the exact per-segment SAH
of every build (internal node id n-1+p is charged to the segment of sorted
position p) and the H-PLOC time are recorded. A segment's label is the
radius minimizing its relative SAH excess over the segment's best radius
plus time_weight times the relative H-PLOC time of that radius, so the
network spends search width only where it lowers the SAH enough. Merge
tasks wider than a segment (the upper tree levels) use the build's
search_radius. Training runs on the GPU (fused forward/backward kernel + Adam).

Everything goes through the C ABI (libnxb.so); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import json
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import scenes, stages
from .builder import BuildConfig, DevicePrimitives, RadiusPolicy
from .native import DeviceBuffer, Event, check, copy_device, lib, synchronize, zero

FEATURES, HIDDEN, CLASSES, PARAMS = 16, 64, 4, 5508
RADII = (2, 4, 8, 16)
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS_PATH = os.path.join(DATA_DIR, "radius_mlp.f32")
META_PATH = os.path.join(DATA_DIR, "radius_mlp.json")


# ---- kernels through the C ABI ----------------------------------------------

def segment_count(n: int, seg_log2: int) -> int:
    return ((n - 1) >> seg_log2) + 1 if n else 0


def segment_features(sorted_keys: DeviceBuffer, sorted_vals: DeviceBuffer, leaves: DeviceBuffer, n: int,
                     bits: int, seg_log2: int) -> DeviceBuffer:
    out = DeviceBuffer(max(1, segment_count(n, seg_log2)) * FEATURES * 4)
    check(lib().nxb_segment_features(sorted_keys.address, sorted_vals.address, leaves.address, n, bits, seg_log2,
                                     out.address, None), "segment_features")
    return out


def segment_sah(nodes: DeviceBuffer, n: int, scene: DeviceBuffer, seg_log2: int) -> np.ndarray:
    segs = segment_count(n, seg_log2)
    acc = DeviceBuffer(max(1, segs) * 8)
    zero(acc)
    check(lib().nxb_segment_sah(nodes.address, n, scene.address, seg_log2, acc.address, None), "segment_sah")
    synchronize()
    return acc.to_array(np.float64, segs)


def mlp_gradients(params: DeviceBuffer, feats: DeviceBuffer, labels: DeviceBuffer, count: int,
                  grads: DeviceBuffer, loss: DeviceBuffer) -> None:
    zero(grads)
    zero(loss)
    check(lib().nxb_mlp_gradients(params.address, feats.address, labels.address, count, grads.address,
                                  loss.address, None), "mlp_gradients")


def mlp_radius_table(params: DeviceBuffer, feats: DeviceBuffer, count: int, min_confidence: float,
                     want_logits: bool = False):
    """(radius table, logits or None, fallback count)."""
    table = DeviceBuffer(max(1, count))
    logits = DeviceBuffer(max(1, count) * CLASSES * 4) if want_logits else None
    fb = DeviceBuffer.from_array(np.zeros(1, dtype=np.uint32))
    check(lib().nxb_mlp_radius_table(params.address, feats.address, count, ctypes.c_float(min_confidence),
                                     table.address, logits.address if logits else None, fb.address, None),
          "mlp_radius_table")
    synchronize()
    return (table.to_array(np.uint8, count),
            logits.to_array(np.float32, count * CLASSES).reshape(count, CLASSES) if logits else None,
            int(fb.to_array(np.uint32)[0]))


# ---- training data -------------------------------------------------------------

@dataclass
class SceneRecord:
    n: int
    features: np.ndarray              # segments x 16
    seg_cost: np.ndarray              # radii x segments (internal-node SAH per segment)
    leaf_cost: float                  # SAH of the leaves (same for every radius)
    hploc_ms: np.ndarray              # radii


def family_scene(index: int, seed: int = 9000, n_lo: int = 60_000, n_hi: int = 300_000) -> np.ndarray:
    """Scene `index` of the training family: blobs over a ground-plane grid with
    the 28.1M workload's proportions (30% plane) at a random smaller size."""
    rng = np.random.default_rng(seed + index)
    n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
    side = max(4, int(np.sqrt(0.3 * n / 2)))
    blobs = int(rng.integers(1, 9))
    return scenes.blobs_plane(n, seed + index, blobs=blobs, side=side, cell=20.53 / side)


def record_scene(tris: np.ndarray, seg_log2: int = 12, bits: int = 64) -> SceneRecord:
    n = len(tris)
    prims = DevicePrimitives.from_numpy(tris)
    nodes, scene, flags = stages.scene_bounds(prims)
    if flags:
        raise RuntimeError("non-finite scene")
    keys, _ = stages.morton(nodes, n, scene, bits)
    b, e = (2, 32) if bits == 32 else (1, 64)
    ko, vo = stages.sort_pairs(keys, n, bits, b, e, 8)
    feats = segment_features(ko, vo, nodes, n, bits, seg_log2)
    synchronize()
    segs = segment_count(n, seg_log2)
    costs = np.zeros((len(RADII), segs))
    times = np.zeros(len(RADII))
    work = DeviceBuffer(n * 4)
    t0, t1 = Event(), Event()
    for k, r in enumerate(RADII):
        copy_device(work, vo, n * 4)
        t0.record()
        stages.hploc(ko, work, n, bits, nodes, BuildConfig(mortonBits=bits, searchRadius=r))
        t1.record()
        times[k] = t0.elapsed_ms(t1)
        costs[k] = segment_sah(nodes, n, scene, seg_log2)
    # leaf part of the SAH (identical for all radii): total minus internal
    total = ctypes.c_double(0)
    from .native import BVH2
    h = BVH2()
    h.nodes = nodes.address
    h.node_count = 2 * n - 1
    h.prim_count = n
    sc = scene.to_array(np.float32, 6)
    h.bounds.bmin.x, h.bounds.bmin.y, h.bounds.bmin.z = (float(v) for v in sc[:3])
    h.bounds.bmax.x, h.bounds.bmax.y, h.bounds.bmax.z = (float(v) for v in sc[3:])
    check(lib().nxb_sah_bvh2(ctypes.byref(h), ctypes.byref(total), None), "sah_bvh2")
    leaf = float(total.value) - float(costs[-1].sum())
    return SceneRecord(n, feats.to_array(np.float32, segs * FEATURES).reshape(segs, FEATURES), costs, leaf, times)


def labels_of(seg_cost: np.ndarray, tolerance: float, time_weight: float = 0.0,
              rel_time: np.ndarray | None = None) -> np.ndarray:
    """Radius class per segment. Without a time weight: the smallest radius
    whose segment SAH is within (1 + tolerance) of the segment's best radius.
    With one: the radius minimizing  SAH / best - 1 + time_weight * (t_r / t_2 - 1),
    t_r being the H-PLOC time at radius r (ties to the smaller radius)."""
    best = seg_cost.min(axis=0)
    if time_weight <= 0.0:
        ok = seg_cost <= best * (1.0 + tolerance)
        return np.argmax(ok, axis=0).astype(np.uint8)
    t = np.asarray(rel_time, dtype=np.float64).reshape(-1, 1)
    score = seg_cost / np.maximum(best, 1e-30) - 1.0 + time_weight * (t / t[0] - 1.0)
    return np.argmin(score, axis=0).astype(np.uint8)


# ---- training -------------------------------------------------------------------

def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / FEATURES), FEATURES * HIDDEN), np.zeros(HIDDEN),
             rng.normal(0, np.sqrt(2 / HIDDEN), HIDDEN * HIDDEN), np.zeros(HIDDEN),
             rng.normal(0, np.sqrt(2 / HIDDEN), HIDDEN * CLASSES), np.zeros(CLASSES)]
    return np.concatenate(parts).astype(np.float32)


@dataclass
class TrainResult:
    params: np.ndarray
    losses: list = field(default_factory=list)
    train_acc: float = 0.0
    val_acc: float = 0.0


def accuracy(params: np.ndarray, X: np.ndarray, y: np.ndarray) -> float:
    p = DeviceBuffer.from_array(params)
    f = DeviceBuffer.from_array(X.astype(np.float32))
    table, _, _ = mlp_radius_table(p, f, len(X), 0.0)
    cls = np.log2(table).astype(np.int64) - 1
    return float(np.mean(cls == y))


def train(X: np.ndarray, y: np.ndarray, steps: int = 3000, batch: int = 8192, lr: float = 3e-3, seed: int = 0,
          class_balance: bool = True, log_every: int = 500) -> TrainResult:
    """Adam on mean cross-entropy; minibatches drawn with replacement (with
    class balancing, each class equally often)."""
    rng = np.random.default_rng(seed)
    params = DeviceBuffer.from_array(init_params(seed))
    m = DeviceBuffer(PARAMS * 4)
    v = DeviceBuffer(PARAMS * 4)
    zero(m)
    zero(v)
    grads = DeviceBuffer(PARAMS * 4)
    loss = DeviceBuffer(4)
    Xf = X.astype(np.float32)
    by_class = [np.flatnonzero(y == c) for c in range(CLASSES)]
    present = [c for c in range(CLASSES) if len(by_class[c])]
    res = TrainResult(params=None)
    bsz = min(batch, len(X))
    for step in range(1, steps + 1):
        if class_balance:
            cls = rng.choice(present, bsz)
            idx = np.array([by_class[c][rng.integers(len(by_class[c]))] for c in cls])
        else:
            idx = rng.integers(0, len(X), bsz)
        fb = DeviceBuffer.from_array(Xf[idx])
        lb = DeviceBuffer.from_array(y[idx].astype(np.uint8))
        mlp_gradients(params, fb, lb, bsz, grads, loss)
        check(lib().nxb_mlp_adam_step(params.address, grads.address, m.address, v.address, step,
                                      ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                      ctypes.c_float(1e-8), None), "adam")
        if step % log_every == 0 or step == 1:
            synchronize()
            res.losses.append((step, float(loss.to_array(np.float32)[0])))
    synchronize()
    res.params = params.to_array(np.float32, PARAMS)
    return res


def save(params: np.ndarray, meta: dict, path: str = WEIGHTS_PATH, meta_path: str = META_PATH) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    params.astype("<f4").tofile(path)
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def load(path: str = WEIGHTS_PATH) -> np.ndarray:
    p = np.fromfile(path, dtype="<f4")
    if p.size != PARAMS:
        raise ValueError(f"{path}: {p.size} floats, expected {PARAMS}")
    return p


# ---- evaluation -----------------------------------------------------------------

def build_eval(tris_dev: DevicePrimitives, n: int, cfg: BuildConfig, reps: int = 3):
    """(BVH2 SAH, BVH8 SAH, mean BVH2 stage ms over reps) of BuildBVH8 with cfg."""
    from . import builder as B
    m = B.BVHBuildMetrics()
    B.FreeDeviceBVH(B.BuildBVH8(tris_dev, n, cfg, m))
    sah2, sah8 = m.bvh2Cost, m.bvh8Cost
    t = []
    for _ in range(reps):
        mt = B.BVHBuildMetrics()
        c2 = BuildConfig(**{**cfg.__dict__, "metricsFlags": 1})
        B.FreeDeviceBVH(B.BuildBVH8(tris_dev, n, c2, mt))
        t.append(mt.computeSceneBoundsTime + mt.computeMortonCodesTime + mt.radixSortTime + mt.bvhBuildTime +
                 mt.bvh8ConversionTime)
    return sah2, sah8, float(np.median(t))


def policy_configs(params: DeviceBuffer, bits: int = 64, seg_log2: int = 12, top_radius: int = 2,
                   min_confidence: float = 0.5) -> dict:
    """Build configurations compared in the evaluation: fixed radii and the
    learned policy (segment radius from the MLP, top_radius above segments)."""
    return {
        "r2": BuildConfig(mortonBits=bits, searchRadius=2),
        "r8": BuildConfig(mortonBits=bits, searchRadius=8),
        "r16": BuildConfig(mortonBits=bits, searchRadius=16),
        "learned": BuildConfig(mortonBits=bits, searchRadius=top_radius, radiusPolicy=RadiusPolicy.LEARNED,
                               segmentLog2=seg_log2, mlpParams=params, mlpMinConfidence=min_confidence),
    }


def compare_policies(tris, params: np.ndarray, min_confidence: float = 0.5, seg_log2: int = 12,
                     bits: int = 64, reps: int = 3, top_radius: int = 2, names=None) -> dict:
    """SAH and build time of fixed radii and the learned policy on one scene
    (host triangles or DevicePrimitives)."""
    d = tris if isinstance(tris, DevicePrimitives) else DevicePrimitives.from_numpy(tris)
    n = d.count
    pbuf = DeviceBuffer.from_array(params.astype(np.float32))
    out = {}
    for name, cfg in policy_configs(pbuf, bits, seg_log2, top_radius, min_confidence).items():
        if names and name not in names:
            continue
        s2, s8, ms = build_eval(d, n, cfg, reps)
        out[name] = {"bvh2_sah": s2, "bvh8_sah": s8, "build_ms": ms}
        if name == "learned":
            segs, fb = ctypes.c_uint32(0), ctypes.c_uint32(0)
            lib().nxb_last_learned_stats(ctypes.byref(segs), ctypes.byref(fb))
            out[name]["segments"] = int(segs.value)
            out[name]["fallbacks"] = int(fb.value)
            out[name]["fallback_rate"] = fb.value / max(1, segs.value)
    return out


def record_family(scenes_count: int = 2000, seg_log2: int = 12, seed: int = 9000, log=print):
    """(features, per-radius segment SAH, owning scene per segment, H-PLOC ms
    per scene and radius) over the first scenes_count family scenes."""
    t0 = time.time()
    feats, costs, times, owners = [], [], [], []
    for i in range(scenes_count):
        rec = record_scene(family_scene(i, seed), seg_log2)
        feats.append(rec.features)
        costs.append(rec.seg_cost)
        times.append(rec.hploc_ms)
        owners.append(np.full(rec.features.shape[0], i))
        if (i + 1) % 200 == 0:
            log(f"recorded {i + 1} scenes, {sum(len(f) for f in feats)} segments, {time.time() - t0:.0f}s")
    return np.concatenate(feats), np.concatenate(costs, axis=1), np.concatenate(owners), np.array(times)


def fit(X: np.ndarray, C: np.ndarray, owner: np.ndarray, tolerance: float, steps: int = 4000, seed: int = 9000,
        time_weight: float = 0.0, rel_time: np.ndarray | None = None):
    """Labels (labels_of), training on 90% of the scenes; returns
    (params, metadata)."""
    y = labels_of(C, tolerance, time_weight, rel_time)
    val = owner % 10 == 0
    res = train(X[~val], y[~val], steps=steps, seed=seed)
    res.train_acc = accuracy(res.params, X[~val], y[~val])
    res.val_acc = accuracy(res.params, X[val], y[val])
    meta = {"segments": int(len(X)), "scenes": int(owner.max() + 1), "tolerance": tolerance,
            "time_weight": time_weight, "rel_time": None if rel_time is None else list(map(float, rel_time)),
            "steps": steps,
            "seed": seed, "label_histogram": np.bincount(y, minlength=CLASSES).tolist(),
            "train_accuracy": res.train_acc, "validation_accuracy": res.val_acc, "loss_curve": res.losses}
    return res.params, meta


def run_training(scenes_count: int = 2000, seg_log2: int = 12, tolerance: float = 0.002, steps: int = 3000,
                 seed: int = 9000, time_weight: float = 0.15, log=print) -> dict:
    """Data generation on the GPU, training, and the weight files."""
    t0 = time.time()
    X, C, owner, T = record_family(scenes_count, seg_log2, seed, log)
    params, meta = fit(X, C, owner, tolerance, steps, seed, time_weight, T.mean(axis=0))
    meta.update({"segment_log2": seg_log2, "hploc_ms_mean_by_radius": T.mean(axis=0).tolist(),
                 "seconds": time.time() - t0})
    save(params, meta)
    log(json.dumps(meta))
    return meta


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train the learned search-radius MLP on the GPU")
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--tolerance", type=float, default=0.002)
    ap.add_argument("--time-weight", type=float, default=0.15)
    ap.add_argument("--segment-log2", type=int, default=12)
    a = ap.parse_args()
    run_training(a.scenes, a.segment_log2, a.tolerance, a.steps, time_weight=a.time_weight)
