"""Batched BVH8 builds of many meshes (include/nxb.h nxb_batch_plan_*) and
their autotuner.

A plan holds the device workspace and outputs for one set of meshes stored
back to back (mesh i = primitives [offsets[i], offsets[i+1])) and records
all stages for all meshes into one hipGraph; run() is one graph launch.

The autotuner picks, for a given batch, the sort digit width, the H-PLOC
launch size (persistent blocks per CU) and the BVH8 collapse
(level-synchronous, or persistent with a lane count) by build time at the
reference quality settings (radius 8, reference merge schedule), then the
search radius and merge schedule: the fastest whose summed per-mesh BVH8 SAH
stays within `quality_tolerance` of the reference settings' (exact SAH
evaluator).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from .builder import BVH8_NODE_DTYPE, BuildConfig, DevicePrimitives, MergeMode
from .native import BVH8, check, copy_to_host, lib


class BatchPlan:
    """nxb_batch_plan: one graph that builds the BVH8 of every mesh."""

    def __init__(self, prims: DevicePrimitives, offsets: np.ndarray, config: BuildConfig | None = None,
                 bvh8_lanes: int = 0):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        self.meshes = len(self.offsets) - 1
        self.prims = prims   # keeps the primitive buffer alive
        cfg = (config or BuildConfig(mortonBits=32)).to_c()
        self._h = ctypes.c_void_p(0)
        check(lib().nxb_batch_plan_create(prims.address, int(prims.prim_type), self.meshes,
                                          self.offsets.ctypes.data, ctypes.byref(cfg), int(bvh8_lanes),
                                          ctypes.byref(self._h)), "batch_plan_create")

    def run(self, stage_times: bool = False):
        """Builds all meshes; with stage_times, returns the stage times (ms):
        bounds, morton, sort, hploc, bvh8, total (stages enqueued one by one)."""
        if stage_times:
            t = (ctypes.c_float * 6)()
            check(lib().nxb_batch_plan_run(self._h, None, t), "batch_plan_run")
            return list(t)
        check(lib().nxb_batch_plan_run(self._h, None, None), "batch_plan_run")
        return None

    def mesh(self, i: int) -> BVH8:
        b = BVH8()
        check(lib().nxb_batch_plan_mesh(self._h, int(i), ctypes.byref(b)), "batch_plan_mesh")
        return b

    def mesh_to_host(self, i: int):
        b = self.mesh(i)
        nodes = copy_to_host(b.nodes or 0, b.node_count * 80, np.uint8).view(BVH8_NODE_DTYPE)
        prims = copy_to_host(b.prim_idx or 0, b.prim_count * 4, np.uint32)
        return nodes, prims

    @property
    def node_total(self) -> int:
        return int(lib().nxb_batch_plan_node_total(self._h))

    def mesh_sah(self, i: int) -> float:
        c = ctypes.c_double(0)
        b = self.mesh(i)
        check(lib().nxb_sah_bvh8(ctypes.byref(b), ctypes.byref(c), None), "sah_bvh8")
        return c.value

    def close(self) -> None:
        if self._h and self._h.value:
            lib().nxb_batch_plan_destroy(self._h)
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def time_plan(plan: BatchPlan, reps: int = 5, warmup: int = 1) -> float:
    """Median wall time (ms) of one graph launch + result readback."""
    for _ in range(warmup):
        plan.run()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.run()
        t.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(t))


@dataclass
class TuneResult:
    config: BuildConfig
    bvh8_lanes: int
    ms: float
    sah: float
    reference_sah: float
    trials: list = field(default_factory=list)


def faster_than(ms: float, best_ms: float, time_tolerance: float) -> bool:
    """ms beats best_ms by more than the timing tolerance."""
    return ms < best_ms * (1.0 - time_tolerance)


def better_setting(ms: float, sah: float, best_ms: float, best_sah: float, time_tolerance: float) -> bool:
    """A quality setting replaces the current choice if it is faster beyond
    the timing tolerance, or as fast within it and of lower SAH."""
    if faster_than(ms, best_ms, time_tolerance):
        return True
    return ms <= best_ms * (1.0 + time_tolerance) and sah < best_sah


def autotune(prims: DevicePrimitives, offsets: np.ndarray, quality_tolerance: float = 0.01,
             digit_bits=(8, 7), bvh8_lanes=(0, 32768, 131072), hploc_blocks=(12, 8), radii=(2, 4, 8),
             modes=(MergeMode.WAVE32, MergeMode.WAVE64), reps: int = 5, time_tolerance: float = 0.03,
             log=None) -> TuneResult:
    """Fastest (digit width, H-PLOC blocks per CU, BVH8 lanes, radius, merge
    schedule) for this batch within quality_tolerance of the reference
    settings' summed BVH8 SAH (bvh8_lanes 0 = the level-synchronous BVH8
    collapse, else the persistent kernel with that many lanes).

    Times within time_tolerance of each other count as equal, so that timing
    noise does not decide: among equally fast speed settings the earlier one
    in the candidate lists stays (the lists start with the plan's defaults:
    8-bit digits, which take a pass fewer than 7-bit ones over the 39-bit
    keys of 2,048 meshes), and among equally fast quality settings
    the one with the lower SAH wins. The default of 3% covers the spread of
    one setting between separately created plans on the 2,048-mesh batch
    (up to 2.5% between two plans, each timed as the median of 5 launches)."""
    trials = []

    def trial(cfg: BuildConfig, lanes: int, quality: bool):
        plan = BatchPlan(prims, offsets, cfg, lanes)
        ms = time_plan(plan, reps)
        sah = float(sum(plan.mesh_sah(i) for i in range(plan.meshes))) if quality else float("nan")
        plan.close()
        rec = {"digit_bits": cfg.sortDigitBits, "hploc_blocks_per_cu": cfg.hplocBlocksPerCu, "lanes": lanes,
               "radius": cfg.searchRadius, "merge_mode": int(cfg.mergeMode), "ms": ms, "sah": sah}
        trials.append(rec)
        if log:
            log(rec)
        return ms, sah

    # speed knobs at the reference quality settings
    best = None
    for d in digit_bits:
        for hb in hploc_blocks:
            for lanes in bvh8_lanes:
                ms, _ = trial(BuildConfig(mortonBits=32, sortDigitBits=d, hplocBlocksPerCu=hb), lanes, False)
                if best is None or faster_than(ms, best[0], time_tolerance):
                    best = (ms, d, hb, lanes)
    _, d, hb, lanes = best
    ref_ms, ref_sah = trial(BuildConfig(mortonBits=32, sortDigitBits=d, hplocBlocksPerCu=hb), lanes, True)
    choice = (ref_ms, ref_sah, BuildConfig(mortonBits=32, sortDigitBits=d, hplocBlocksPerCu=hb))
    for mode in modes:
        for r in radii:
            if r == 8 and mode == MergeMode.WAVE32:
                continue
            cfg = BuildConfig(mortonBits=32, sortDigitBits=d, hplocBlocksPerCu=hb, searchRadius=r, mergeMode=mode)
            ms, sah = trial(cfg, lanes, True)
            if sah <= ref_sah * (1.0 + quality_tolerance) and better_setting(ms, sah, choice[0], choice[1],
                                                                             time_tolerance):
                choice = (ms, sah, cfg)
    return TuneResult(choice[2], lanes, choice[0], choice[1], ref_sah, trials)


if __name__ == "__main__":
    # Times the plan of the benchmark batch (synthetic meshes of 64 to 65,536
    # triangles): python -m nexusbvh_amd.batch [--meshes M] [--runs K]
    import argparse
    import json

    from . import scenes

    ap = argparse.ArgumentParser(description="time one batch plan of synthetic meshes")
    ap.add_argument("--meshes", type=int, default=2048)
    ap.add_argument("--runs", type=int, default=10)
    args = ap.parse_args()
    tris, offsets = scenes.batch(args.meshes)
    dev = DevicePrimitives.from_numpy(tris)
    plan = BatchPlan(dev, np.asarray(offsets, dtype=np.uint32), BuildConfig(mortonBits=32))
    graph_ms = time_plan(plan, reps=args.runs, warmup=2)
    stages = plan.run(stage_times=True)
    names = ("bounds", "morton", "sort", "hploc", "bvh8", "total")
    print(json.dumps({"meshes": plan.meshes, "triangles": int(len(tris)), "graph_ms": round(graph_ms, 4),
                      "stage_ms": {k: round(v, 4) for k, v in zip(names, stages)}}), flush=True)
    plan.close()
