#!/usr/bin/env python3
"""Build-throughput benchmark of the BVH2 + BVH8 construction pipeline on one MI355X.

A step is one full NXB BuildBVH8 call (scene bounds -> Morton -> radix sort ->
H-PLOC BVH2 -> BVH8) through the C ABI, on triangles already resident in HBM,
including the output allocation, the blocking readback the API performs and
the release of the result. `value` = triangles * steps / wall time (Mtri/s).

Per-stage times are HIP events the library records on the build stream inside
the timed region (metrics flag "times only": no SAH kernels); the roofline of
the dominant stage uses the algorithmic bytes per triangle of that stage.

Default workload: the 1.1M-triangle displaced icosphere with 32-bit Morton,
radius 8 and the reference merge schedule. A second, shorter measurement of
the 28.1M-triangle blob/ground-plane scene (64-bit Morton) reports the stage
bandwidths of the memory-bound stages at full size.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_bytes(stage: str, n: int, key_bytes: int, m8: int, prim_floats: int = 9) -> float:
    """Algorithmic bytes of one stage launch (per-triangle figures x n)."""
    k = key_bytes
    passes = 4 if k == 4 else 8
    per = {
        "bounds": 4 * prim_floats + 32,             # R primitive, W leaf node
        "morton": 24 + k + 4,                        # R leaf box, W key + iota
        "sort": 2 * (k + 4) * passes + k,            # per pass R/W key+value (+ histogram read)
        "hploc": k + 76,
        "bvh8": 92 + 80.0 * m8 / max(1, n),
    }[stage]
    return per * n


def measured_traffic(workload: str, bits: int, stage: str):
    """HBM bytes per launch of a stage from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/pmc_traffic.json), scaled by the copy-kernel
    calibration; None when that workload was not profiled."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            doc = json.load(f)
        return doc[f"{workload}/{bits}"][stage]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_nproc() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def measure(d, n, cfg, steps, warmup):
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import native
    for _ in range(warmup):
        B.FreeDeviceBVH(B.BuildBVH8(d, n, cfg))
    native.synchronize()
    agg = B.BVHBuildMetrics()
    m8 = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        m = B.BVHBuildMetrics()
        bvh = B.BuildBVH8(d, n, cfg, m)
        m8 = bvh.node_count
        B.FreeDeviceBVH(bvh)
        agg += m
    native.synchronize()
    t1 = time.perf_counter()
    avg = agg / float(steps)
    return (t1 - t0) * 1000.0 / steps, avg, m8


def trace_section(d, n, cfg):
    """Closest-hit traversal of the measured BVH8: 1024x1024 pinhole primary
    rays and as many incoherent rays (random origins around the scene, random
    directions), each timed with HIP events over 5 launches after 2 warm-up
    launches; plus the nodes visited and triangles tested per ray. The same
    for the BVH8 whose child slots the auction assigned, with the collapse
    time of that build (mean of 5)."""
    import ctypes
    import dataclasses
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import native
    from nexusbvh_amd.trace import HIT_DTYPE, INVALID, camera_rays, random_rays

    def trace(bvh, rays):
        count = len(rays)
        rbuf = native.DeviceBuffer.from_array(rays.view(np.uint8))
        hbuf = native.DeviceBuffer(count * HIT_DTYPE.itemsize)
        sbuf = native.DeviceBuffer(count * 8)

        def launch(stats):
            native.check(native.lib().nxb_trace_closest(ctypes.byref(bvh), d.address, rbuf.address, count,
                                                        hbuf.address, sbuf.address if stats else None, None),
                         "nxb_trace_closest")
        launch(True)
        native.synchronize()
        st = sbuf.to_array(np.uint32, 2 * count).reshape(count, 2)
        hits = hbuf.to_array(np.uint8, count * HIT_DTYPE.itemsize).view(HIT_DTYPE)
        launch(False)
        e0, e1 = native.Event(), native.Event()
        e0.record()
        for _ in range(5):
            launch(False)
        e1.record()
        ms = e0.elapsed_ms(e1) / 5
        return {"rays": count, "ms": round(ms, 4), "Mrays/s": round(count / (ms * 1e-3) / 1e6, 1),
                "hit_fraction": round(float((hits["prim"] != INVALID).mean()), 4),
                "nodes_per_ray": round(float(st[:, 0].mean()), 3),
                "tris_per_ray": round(float(st[:, 1].mean()), 3)}

    res = {}
    for mode in (B.SlotAssignment.GREEDY, B.SlotAssignment.AUCTION):
        mcfg = dataclasses.replace(cfg, slotAssignment=mode)
        bvh = B.BuildBVH8(d, n, mcfg)
        lo = (bvh.bounds.bmin.x, bvh.bounds.bmin.y, bvh.bounds.bmin.z)
        hi = (bvh.bounds.bmax.x, bvh.bounds.bmax.y, bvh.bounds.bmax.z)
        out = res if mode == B.SlotAssignment.GREEDY else res.setdefault("auction", {})
        out["primary"] = trace(bvh, camera_rays((lo, hi), 1024, 1024, seed=7))
        out["incoherent"] = trace(bvh, random_rays((lo, hi), 1024 * 1024, seed=8))
        B.FreeDeviceBVH(bvh)
        if mode == B.SlotAssignment.AUCTION:
            total = 0.0
            for _ in range(5):
                m = B.BVHBuildMetrics()
                B.FreeDeviceBVH(B.BuildBVH8(d, n, mcfg, m))
                total += m.bvh8ConversionTime
            out["bvh8_ms"] = round(total / 5, 4)
    res["rays"] = "primary: 1024x1024 pinhole view of the scene; incoherent: origins uniform in the box grown by half, directions uniform"
    return res


def stage_table(avg, n, key_bytes, m8, prim_floats=9):
    stages = {"bounds": avg.computeSceneBoundsTime, "morton": avg.computeMortonCodesTime,
              "sort": avg.radixSortTime, "hploc": avg.bvhBuildTime, "bvh8": avg.bvh8ConversionTime}
    out = {}
    for k, ms in stages.items():
        b = stage_bytes(k, n, key_bytes, m8, prim_floats)
        out[k] = {"ms": round(ms, 4), "GB/s": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None,
                  "frac_hbm": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None,
                  "Mtri/s": round(n / (ms * 1e-3) / 1e6, 1) if ms > 0 else None}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="icosphere_1m1")
    ap.add_argument("--bits", type=int, default=32)
    ap.add_argument("--mode", default="wave32", choices=["wave32", "wave64"])
    ap.add_argument("--radius", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-size", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-trace", action="store_true")
    ap.add_argument("--min-confidence", type=float, default=0.0,
                    help="learned radius: arg-max probability below which a segment uses radius 8")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a plain copy kernel of known size (2^26 key/value pairs: 512 MiB read, "
                         "512 MiB written) for scaling the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    from nexusbvh_amd import builder as B
    from nexusbvh_amd import native, scenes

    if native.device_count() == 0:
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    native.lib()

    if args.calibrate:
        from nexusbvh_amd import stages
        m = 1 << 26
        keys = native.DeviceBuffer.from_array(np.arange(m, dtype=np.uint32))
        vals = native.DeviceBuffer.from_array(np.arange(m, dtype=np.uint32))
        for _ in range(3):  # an empty bit window makes the sort a plain pair copy (k_copy_pairs)
            stages.sort_pairs(keys, m, 32, 0, 0, 8, values=vals)
        native.synchronize()
        print(json.dumps({"calibration": "k_copy_pairs", "read_bytes": 8 * m, "write_bytes": 8 * m}))
        return 0

    tris = scenes.workload(args.workload)
    n = len(tris)
    d = B.DevicePrimitives.from_numpy(tris)
    mode = B.MergeMode.WAVE32 if args.mode == "wave32" else B.MergeMode.WAVE64
    cfg = B.BuildConfig(mortonBits=args.bits, mergeMode=mode, searchRadius=args.radius, metricsFlags=1)
    ms, avg, m8 = measure(d, n, cfg, args.steps, args.warmup)
    kb = args.bits // 8
    stages = stage_table(avg, n, kb, m8)
    dom = max(("bounds", "morton", "sort", "hploc", "bvh8"), key=lambda k: stages[k]["ms"])
    dom_ms = stages[dom]["ms"]
    dom_bytes = stage_bytes(dom, n, kb, m8)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    value = n / (ms * 1e-3) / 1e6

    # quality of the measured build (full metrics, one extra build)
    mq = B.BVHBuildMetrics()
    B.FreeDeviceBVH(B.BuildBVH8(d, n, B.BuildConfig(mortonBits=args.bits, mergeMode=mode,
                                                    searchRadius=args.radius), mq))

    out = {
        "metric": "build Mtri/s (BVH2+BVH8, 1xMI355X) + SAH ratio learned-radius vs plain r=8",
        "value": round(value, 2),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "triangles": n, "morton_bits": args.bits, "merge_mode": args.mode,
                   "search_radius": args.radius, "primitive": "triangle"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, args.bits, dom),
                     "algorithmic_bytes": int(dom_bytes), "avg_ms": dom_ms},
        "stages": stages,
        "stage_sum_ms": round(avg.totalTime, 4),
        "bvh2_sah": mq.bvh2Cost,
        "bvh8_sah": mq.bvh8Cost,
        "bvh8_nodes": m8,
        "avg_children": round(mq.averageChildPerNode, 4),
    }

    # the same workload given as AABB primitives (the boxes of the triangles)
    boxes = scenes.boxes_of(tris)
    da = B.DevicePrimitives.from_numpy(boxes, B.PrimType.AABB)
    del boxes
    msa, avga, m8a = measure(da, n, cfg, 10, 2)
    out["aabb_input"] = {"workload": args.workload + " as AABBs", "ms_per_build": round(msa, 4),
                         "Mtri/s": round(n / (msa * 1e-3) / 1e6, 1),
                         "stages": stage_table(avga, n, kb, m8a, prim_floats=6)}
    del da

    if not args.no_trace:
        out["traversal"] = trace_section(d, n, B.BuildConfig(mortonBits=args.bits, mergeMode=mode,
                                                               searchRadius=args.radius))

    if not args.no_full_size:
        from nexusbvh_amd import learned as NL
        big = scenes.workload("blobs_28m1")
        nb = len(big)
        db = B.DevicePrimitives.from_numpy(big)
        del big
        cb = B.BuildConfig(mortonBits=64, metricsFlags=1)
        msb, avgb, m8b = measure(db, nb, cb, 5, 2)
        out["full_size"] = {"workload": "blobs_28m1", "triangles": nb, "morton_bits": 64,
                            "ms_per_build": round(msb, 3), "Mtri/s": round(nb / (msb * 1e-3) / 1e6, 1),
                            "stages": stage_table(avgb, nb, 8, m8b)}
        # learned radius vs fixed radius 8 and 16 on the same GPU (exact SAH, stage-time totals)
        cmp = NL.compare_policies(db, NL.load(), args.min_confidence, top_radius=8, reps=3,
                                  names=["r8", "r16", "learned"])
        lr, r8, r16 = cmp["learned"], cmp["r8"], cmp["r16"]
        out["learned_radius"] = {
            "workload": "blobs_28m1", "weights": "nexusbvh_amd/data/radius_mlp.f32",
            "min_confidence": args.min_confidence, "policies": cmp,
            "sah_ratio_learned_vs_r8": {"bvh2": round(lr["bvh2_sah"] / r8["bvh2_sah"], 4),
                                        "bvh8": round(lr["bvh8_sah"] / r8["bvh8_sah"], 4)},
            "sah_ratio_learned_vs_r16": {"bvh2": round(lr["bvh2_sah"] / r16["bvh2_sah"], 4),
                                         "bvh8": round(lr["bvh8_sah"] / r16["bvh8_sah"], 4)},
            "time_ratio_learned_vs_r8": round(lr["build_ms"] / r8["build_ms"], 4),
            "fallback_rate": lr["fallback_rate"]}
        out["sah_ratio_learned_vs_r8"] = out["learned_radius"]["sah_ratio_learned_vs_r8"]
        del db

    if not args.no_batch:
        # 2048 meshes (log-uniform 64..65536 triangles): one graph launch vs per-mesh calls
        from nexusbvh_amd.batch import BatchPlan, autotune, time_plan
        btris, boff = scenes.batch(2048)
        boff = np.asarray(boff, dtype=np.uint32)
        bn = len(btris)
        bd = B.DevicePrimitives.from_numpy(btris)
        del btris
        plan = BatchPlan(bd, boff, B.BuildConfig(mortonBits=32))
        g_ms = time_plan(plan, 5)
        st = plan.run(stage_times=True)
        plan.close()
        bcfg = B.BuildConfig(mortonBits=32)
        views = [bd.view(int(boff[m]), int(boff[m + 1] - boff[m])) for m in range(len(boff) - 1)]
        B.FreeDeviceBVH(B.BuildBVH8(views[0], views[0].count, bcfg))
        native.synchronize()
        t0 = time.perf_counter()
        for v in views:
            B.FreeDeviceBVH(B.BuildBVH8(v, v.count, bcfg))
        native.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3
        tuned = autotune(bd, boff)
        out["batched"] = {
            "workload": "batch_2048", "meshes": int(len(boff) - 1), "triangles": int(bn), "morton_bits": 32,
            "graph_ms": round(g_ms, 3), "graph_Mtri/s": round(bn / g_ms / 1e3, 1),
            "stage_ms": dict(zip(("bounds", "morton", "sort", "hploc", "bvh8", "total"), (round(x, 3) for x in st))),
            "sequential_ms": round(seq_ms, 3), "sequential_Mtri/s": round(bn / seq_ms / 1e3, 1),
            "speedup_vs_sequential": round(seq_ms / g_ms, 2),
            "autotuned": {"sort_digit_bits": tuned.config.sortDigitBits, "search_radius": tuned.config.searchRadius,
                          "merge_mode": int(tuned.config.mergeMode), "bvh8_lanes": tuned.bvh8_lanes,
                          "graph_ms": round(tuned.ms, 3), "graph_Mtri/s": round(bn / tuned.ms / 1e3, 1),
                          "bvh8_sah_ratio_vs_reference_settings": round(tuned.sah / tuned.reference_sah, 5)}}
        del bd, views

    if rank == 0 and not args.no_cpu_baseline:
        from oracle import oracle as O
        samples = []
        t_end = time.perf_counter() + 12.0
        while True:
            t0 = time.perf_counter()
            O.build(tris, 1, args.bits, 32 if mode == B.MergeMode.WAVE32 else 64, args.radius)
            samples.append(time.perf_counter() - t0)
            if time.perf_counter() > t_end or len(samples) >= 5:
                break
        best = sorted(samples)[len(samples) // 2]
        out["cpu_baseline"] = {"value": round(n / best / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
                               "sample": f"{len(samples)} full single-threaded oracle builds (bounds, Morton, "
                                         f"std::sort, sequential H-PLOC, BVH8) of {args.workload}, median",
                               "host_cpu": cpu_model(), "host_nproc": cpu_nproc()}
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
