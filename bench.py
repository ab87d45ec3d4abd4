#!/usr/bin/env python3
"""Build-throughput benchmark of the BVH2 + BVH8 construction pipeline on one MI355X.

A step is one full NXB BuildBVH8 call (scene bounds -> Morton -> radix sort ->
H-PLOC BVH2 -> BVH8) through the C ABI, on triangles already resident in HBM,
including the output allocation, the blocking readback the API performs and
the release of the result. `value` = triangles * steps / wall time (Mtri/s).

Headline workload: the 28.1M-triangle blob/ground-plane scene (64 Gaussian
blobs plus a ground-plane grid) with 64-bit Morton keys, radius 8 and the
reference merge schedule. Per-stage times are HIP events the library records
on the build stream inside the timed region (metrics flag "times only": no SAH
kernels). The H-PLOC stage is one kernel launch (k_hploc) between two of those
events, so its time is that kernel's launch duration; the roofline divides
its algorithmic bytes (8 + 108 per triangle: a BVH8 build's H-PLOC writes a
64-byte wide record per internal node) by it.

Secondary sections: the same scene with 32-bit keys and as AABB input, the
sort against rocPRIM's radix_sort_pairs on the scene's 64-bit keys, the
learned search radius against fixed radii 8 and 16 (at the library's default
confidence threshold and with the fallback off), the 1.1M icosphere with 32-
and 64-bit keys, the 2.9M sliver scene in both merge modes, the 300K uniform
scene, ray traversal, batched builds, and the single-threaded CPU oracle on
the headline scene.

`--pmc-summary DIR` turns rocprofv3 FETCH_SIZE / WRITE_SIZE counter output
(DIR/fetch, DIR/write; each a run of `bench.py --only-headline`) and a copy
kernel's counters (DIR/copy_fetch, DIR/copy_write; runs of `--calibrate`)
into profiles/pmc_traffic.json.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STAGE_KEYS = ("bounds", "morton", "sort", "hploc", "bvh8")
# kernels of each stage, for the counter summary
STAGE_KERNELS = {"bounds": ("k_scene_bounds",), "morton": ("k_morton_hist",),
                 "sort": ("k_sort_upsweep", "k_sort_rowscan", "k_sort_downsweep", "k_copy_pairs"),
                 "hploc": ("k_hploc",), "bvh8": ("k_bvh8_top", "k_bvh8_level")}
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def stage_bytes(stage: str, n: int, key_bytes: int, m8: int, prim_bytes: int = 36) -> float:
    """Algorithmic bytes of one stage of one BVH8 build (per-triangle figures x n)."""
    from nexusbvh_amd.builder import stage_bytes_per_prim
    return stage_bytes_per_prim(STAGE_KEYS.index(stage), prim_bytes, key_bytes, m8 / max(1, n), True) * n


def measured_traffic(workload: str, bits: int, stage: str):
    """HBM bytes per launch of a stage from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/pmc_traffic.json), scaled by the copy-kernel
    calibration; None when that workload was not profiled."""
    try:
        with open(TRAFFIC_PATH) as f:
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
    """(wall ms per build, mean stage metrics, BVH8 node count) over `steps`
    builds after `warmup` untimed ones."""
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
    return (t1 - t0) * 1000.0 / steps, agg / float(steps), m8


def bvh2_section(name, d, n, bits, cfg, steps=5, warmup=2):
    """BuildBVH2 (the reference's BVH2 API) on the same scene: wall ms per
    build and its four stages, beside the reference's "BVH2-only Mtri/s"."""
    import dataclasses
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import native
    from nexusbvh_amd.builder import stage_bytes_per_prim
    cfg = dataclasses.replace(cfg, mortonBits=bits)
    for _ in range(warmup):
        B.FreeDeviceBVH(B.BuildBVH2(d, n, cfg))
    native.synchronize()
    agg = B.BVHBuildMetrics()
    t0 = time.perf_counter()
    for _ in range(steps):
        m = B.BVHBuildMetrics()
        B.FreeDeviceBVH(B.BuildBVH2(d, n, cfg, m))
        agg += m
    native.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    avg = agg / float(steps)
    stages = {}
    for i, (k, t) in enumerate((("bounds", avg.computeSceneBoundsTime), ("morton", avg.computeMortonCodesTime),
                                ("sort", avg.radixSortTime), ("hploc", avg.bvhBuildTime))):
        b = stage_bytes_per_prim(i, 36, bits // 8, 0.0, False) * n
        stages[k] = {"ms": round(t, 4), "GB/s": round(b / (t * 1e-3) / 1e9, 1) if t > 0 else None,
                     "Mtri/s": round(n / (t * 1e-3) / 1e6, 1) if t > 0 else None}
    four = avg.computeSceneBoundsTime + avg.computeMortonCodesTime + avg.radixSortTime + avg.bvhBuildTime
    return {"workload": name, "triangles": n, "morton_bits": bits, "ms_per_build": round(ms, 4),
            "Mtri/s": round(n / (ms * 1e-3) / 1e6, 1), "stage_sum_ms": round(four, 4),
            "stage_sum_Mtri/s": round(n / (four * 1e-3) / 1e6, 1), "stages": stages}


def stage_table(avg, n, key_bytes, m8, prim_bytes=36):
    vals = {"bounds": avg.computeSceneBoundsTime, "morton": avg.computeMortonCodesTime,
            "sort": avg.radixSortTime, "hploc": avg.bvhBuildTime, "bvh8": avg.bvh8ConversionTime}
    out = {}
    for k, ms in vals.items():
        b = stage_bytes(k, n, key_bytes, m8, prim_bytes)
        out[k] = {"ms": round(ms, 4), "GB/s": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None,
                  "frac_hbm": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None,
                  "Mtri/s": round(n / (ms * 1e-3) / 1e6, 1) if ms > 0 else None}
    return out


def quality(d, n, cfg):
    """BVH2/BVH8 SAH and children per node of one build with full metrics."""
    import dataclasses
    from nexusbvh_amd import builder as B
    m = B.BVHBuildMetrics()
    B.FreeDeviceBVH(B.BuildBVH8(d, n, dataclasses.replace(cfg, metricsFlags=0), m))
    return {"bvh2_sah": m.bvh2Cost, "bvh8_sah": m.bvh8Cost, "avg_children": round(m.averageChildPerNode, 4)}


def workload_section(name, d, n, bits, mode_name, radius=8, steps=10, warmup=2, prim_bytes=36):
    from nexusbvh_amd.builder import BuildConfig, MergeMode
    mode = MergeMode.WAVE32 if mode_name == "wave32" else MergeMode.WAVE64
    cfg = BuildConfig(mortonBits=bits, mergeMode=mode, searchRadius=radius, metricsFlags=1)
    ms, avg, m8 = measure(d, n, cfg, steps, warmup)
    out = {"workload": name, "triangles": n, "morton_bits": bits, "merge_mode": mode_name, "search_radius": radius,
           "ms_per_build": round(ms, 4), "Mtri/s": round(n / (ms * 1e-3) / 1e6, 1),
           "stage_sum_ms": round(avg.totalTime, 4), "stages": stage_table(avg, n, bits // 8, m8, prim_bytes),
           "bvh8_nodes": m8}
    out.update(quality(d, n, cfg))
    return out


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


def learned_section(db, confidences):
    """Learned radius against fixed radii 8 and 16 on the headline scene
    (exact SAH; build time = sum of the five stage times, median of 3)."""
    from nexusbvh_amd import learned as NL
    params = NL.load()
    fixed = NL.compare_policies(db, params, 0.5, top_radius=8, reps=3, names=["r8", "r16"])
    r8, r16 = fixed["r8"], fixed["r16"]
    out = {"workload": "blobs_28m1", "weights": "nexusbvh_amd/data/radius_mlp.f32", "r8": r8, "r16": r16}
    for c in confidences:
        lr = NL.compare_policies(db, params, c, top_radius=8, reps=3, names=["learned"])["learned"]
        out[f"learned_min_confidence_{c}"] = {
            **lr,
            "sah_ratio_vs_r8": {"bvh2": round(lr["bvh2_sah"] / r8["bvh2_sah"], 4),
                                "bvh8": round(lr["bvh8_sah"] / r8["bvh8_sah"], 4)},
            "sah_ratio_vs_r16": {"bvh2": round(lr["bvh2_sah"] / r16["bvh2_sah"], 4),
                                 "bvh8": round(lr["bvh8_sah"] / r16["bvh8_sah"], 4)},
            "time_ratio_vs_r8": round(lr["build_ms"] / r8["build_ms"], 4)}
    return out


def sort_yardstick_section(d, n, reps=10):
    """The pipeline's key/value sort against rocPRIM's radix_sort_pairs on the
    same 64-bit Morton keys of the scene (sort window bits [1, 63)), each
    timed with HIP events over `reps` runs after 2 warm-up runs. Ours is the
    stage entry point nxb_radix_sort_pairs: its own digit-histogram pass plus
    its 9-bit passes, values the iota (in a build the histograms come with
    the Morton kernel); rocPRIM gets the iota as a value array. The two
    outputs must be equal (both sorts are stable). GB/s and the share of
    8 TB/s are against the stage's 200 algorithmic bytes per key."""
    import ctypes
    from nexusbvh_amd import native, stages
    nodes, scene, _ = stages.scene_bounds(d)
    keys, _ = stages.morton(nodes, n, scene, 64)
    del nodes
    ys = ctypes.CDLL(os.path.join(ROOT, "nexusbvh_amd", "lib", "libnxb_yardstick.so"))
    f = ys.nxb_ys_sort_pairs_u64
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t),
                  ctypes.c_void_p]
    ko1, vo1 = native.DeviceBuffer(n * 8), native.DeviceBuffer(n * 4)
    ko2, vo2 = native.DeviceBuffer(n * 8), native.DeviceBuffer(n * 4)
    iota = native.DeviceBuffer.from_array(np.arange(n, dtype=np.uint32))
    tb = ctypes.c_size_t(0)
    if f(keys.address, ko2.address, iota.address, vo2.address, n, 1, 63, None, ctypes.byref(tb), None) != 0:
        raise RuntimeError("rocprim::radix_sort_pairs temporary-size query failed")
    temp = native.DeviceBuffer(max(1, tb.value))

    def ours():
        native.check(native.lib().nxb_radix_sort_pairs(keys.address, None, ko1.address, vo1.address, n, 64, 1, 63,
                                                       0, None, 0, None), "nxb_radix_sort_pairs")

    def rocprim():
        size = ctypes.c_size_t(tb.value)
        rc = f(keys.address, ko2.address, iota.address, vo2.address, n, 1, 63, temp.address, ctypes.byref(size), None)
        if rc != 0:
            raise RuntimeError(f"rocprim::radix_sort_pairs failed ({rc})")

    def timed(fn):
        for _ in range(2):
            fn()
        native.synchronize()
        e0, e1 = native.Event(), native.Event()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        return e0.elapsed_ms(e1) / reps

    res = {}
    for name, fn in (("nxb", ours), ("rocprim", rocprim)):
        ms = timed(fn)
        gbs = 200.0 * n / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
    native.synchronize()
    same = (np.array_equal(ko1.to_array(np.uint64, n), ko2.to_array(np.uint64, n)) and
            np.array_equal(vo1.to_array(np.uint32, n), vo2.to_array(np.uint32, n)))
    if not same:
        raise RuntimeError("the pipeline's sort and rocPRIM's disagree")
    res.update({"keys": n, "key_bits": 64, "sort_window": [1, 63], "rocprim_temp_bytes": int(tb.value),
                "outputs_equal": True, "speedup_vs_rocprim": round(res["rocprim"]["ms"] / res["nxb"]["ms"], 3)})
    return res


def batch_section():
    """2048 meshes (log-uniform 64..65536 triangles): one graph launch vs per-mesh calls."""
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import native, scenes
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
    return {
        "workload": "batch_2048", "meshes": int(len(boff) - 1), "triangles": int(bn), "morton_bits": 32,
        "graph_ms": round(g_ms, 3), "graph_Mtri/s": round(bn / g_ms / 1e3, 1),
        "stage_ms": dict(zip(("bounds", "morton", "sort", "hploc", "bvh8", "total"), (round(x, 3) for x in st))),
        "sequential_ms": round(seq_ms, 3), "sequential_Mtri/s": round(bn / seq_ms / 1e3, 1),
        "speedup_vs_sequential": round(seq_ms / g_ms, 2),
        "autotuned": {"sort_digit_bits": tuned.config.sortDigitBits,
                      "hploc_blocks_per_cu": tuned.config.hplocBlocksPerCu or 12,
                      "search_radius": tuned.config.searchRadius,
                      "merge_mode": int(tuned.config.mergeMode), "bvh8_lanes": tuned.bvh8_lanes,
                      "graph_ms": round(tuned.ms, 3), "graph_Mtri/s": round(bn / tuned.ms / 1e3, 1),
                      "bvh8_sah_ratio_vs_reference_settings": round(tuned.sah / tuned.reference_sah, 5),
                      # the quality trials the choice was made from (radius, merge schedule)
                      "quality_trials": [{"search_radius": t["radius"], "merge_mode": t["merge_mode"],
                                          "graph_ms": round(t["ms"], 3),
                                          "bvh8_sah_ratio": round(t["sah"] / tuned.reference_sah, 5)}
                                         for t in tuned.trials if t["sah"] == t["sah"]]}}


# ---------------------------------------------------------------- counters
def _counter_rows(path_glob):
    """(kernel name, counter value) rows of rocprofv3 counter_collection CSVs."""
    rows = []
    for path in glob.glob(path_glob, recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
                val = r.get("Counter_Value") or r.get("Counter-Value")
                if val is not None:
                    rows.append((name, float(val)))
    return rows


def _per_dispatch(rows, kernels):
    vals = [v for name, v in rows if any(k in name for k in kernels)]
    return vals


def pmc_summary(src: str, workload: str, bits: int, n: int, m8: int, builds: int) -> dict:
    """profiles/pmc_traffic.json entry of one workload from rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch), scaled by the factor a
    plain pair copy of known size reads for each counter."""
    cf = _per_dispatch(_counter_rows(os.path.join(src, "copy_fetch", "**", "*counter_collection.csv")),
                       ("k_copy_pairs",))
    cw = _per_dispatch(_counter_rows(os.path.join(src, "copy_write", "**", "*counter_collection.csv")),
                       ("k_copy_pairs",))
    copy_bytes = 8 * (1 << 26)
    fetch_factor = copy_bytes / (1024.0 * float(np.median(cf)))
    write_factor = copy_bytes / (1024.0 * float(np.median(cw)))
    fr = _counter_rows(os.path.join(src, "fetch", "**", "*counter_collection.csv"))
    wr = _counter_rows(os.path.join(src, "write", "**", "*counter_collection.csv"))
    out = {"calibration": {"kernel": "k_copy_pairs (2^26 uint32 key/value pairs)", "read_bytes": copy_bytes,
                           "write_bytes": copy_bytes, "fetch_factor": round(fetch_factor, 4),
                           "write_factor": round(write_factor, 4)}}
    entry = {}
    for stage, kernels in STAGE_KERNELS.items():
        f = sum(_per_dispatch(fr, kernels)) * 1024.0 * fetch_factor / builds
        w = sum(_per_dispatch(wr, kernels)) * 1024.0 * write_factor / builds
        alg = stage_bytes(stage, n, bits // 8, m8)
        entry[stage] = {"fetch_bytes": int(f), "write_bytes": int(w), "traffic_bytes": int(f + w),
                        "algorithmic_bytes": int(alg), "traffic_over_algorithmic": round((f + w) / alg, 3)}
    out[f"{workload}/{bits}"] = entry
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="blobs_28m1")
    ap.add_argument("--bits", type=int, default=64)
    ap.add_argument("--mode", default="wave32", choices=["wave32", "wave64"])
    ap.add_argument("--radius", type=int, default=8)
    ap.add_argument("--only-headline", action="store_true",
                    help="time the headline workload only (for profiler runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the other workloads")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-trace", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a plain copy kernel of known size (2^26 key/value pairs: 512 MiB read, "
                         "512 MiB written) for scaling the FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="write profiles/pmc_traffic.json from rocprofv3 counter runs under DIR")
    ap.add_argument("--pmc-builds", type=int, default=0,
                    help="builds per counter run (default: --steps + --warmup + 1)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    from nexusbvh_amd import builder as B
    from nexusbvh_amd import native, scenes

    if args.pmc_summary:
        # needs n and the BVH8 node count of the workload: both come from the
        # quality build in the headline line of the same counter runs
        with open(os.path.join(args.pmc_summary, "headline.json")) as f:
            head = json.loads(f.read().strip().splitlines()[-1])
        builds = args.pmc_builds or (head["steps"] + head["warmup"] + 1)
        doc = pmc_summary(args.pmc_summary, head["config"]["workload"], head["config"]["morton_bits"],
                          head["config"]["triangles"], head["bvh8_nodes"], builds)
        try:
            with open(TRAFFIC_PATH) as f:
                old = json.load(f)
        except (OSError, ValueError):
            old = {}
        old.update(doc)
        with open(TRAFFIC_PATH, "w") as f:
            json.dump(old, f, indent=1)
        print(json.dumps(doc, indent=1))
        return 0

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

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    log(f"generating {args.workload}")
    tris = scenes.workload(args.workload)
    n = len(tris)
    d = B.DevicePrimitives.from_numpy(tris)
    mode = B.MergeMode.WAVE32 if args.mode == "wave32" else B.MergeMode.WAVE64
    cfg = B.BuildConfig(mortonBits=args.bits, mergeMode=mode, searchRadius=args.radius, metricsFlags=1)
    log(f"headline: {args.steps} builds of {n} triangles")
    ms, avg, m8 = measure(d, n, cfg, args.steps, args.warmup)
    kb = args.bits // 8
    stages = stage_table(avg, n, kb, m8)
    dom = max(STAGE_KEYS, key=lambda k: stages[k]["ms"])
    dom_ms = stages[dom]["ms"]
    dom_bytes = stage_bytes(dom, n, kb, m8)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    value = n / (ms * 1e-3) / 1e6
    q = quality(d, n, cfg)
    kernel_of = {"bounds": "k_scene_bounds", "morton": "k_morton_hist", "sort": "sort passes",
                 "hploc": "k_hploc", "bvh8": "k_bvh8_top + k_bvh8_level"}

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
        "roofline": {"bound": "hbm", "kernel": kernel_of[dom], "stage": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, args.bits, dom),
                     "algorithmic_bytes": int(dom_bytes), "avg_ms": dom_ms},
        "stages": stages,
        "stage_sum_ms": round(avg.totalTime, 4),
        "bvh8_nodes": m8,
        **q,
    }
    if args.only_headline:
        print(json.dumps(out), flush=True)
        return 0

    if not args.no_secondary:
        log("headline scene with 32-bit keys and as AABB input")
        out["blobs_28m1_k32"] = workload_section(args.workload, d, n, 32, args.mode, args.radius, steps=5)
        log("headline scene as BVH2-only builds (32- and 64-bit keys)")
        out["bvh2_only"] = {f"k{b}": bvh2_section(args.workload, d, n, b, cfg) for b in (32, 64)}
        boxes = scenes.boxes_of(tris)
        da = B.DevicePrimitives.from_numpy(boxes, B.PrimType.AABB)
        del boxes
        msa, avga, m8a = measure(da, n, cfg, 5, 2)
        out["aabb_input"] = {"workload": args.workload + " as AABBs", "morton_bits": args.bits,
                             "ms_per_build": round(msa, 4), "Mtri/s": round(n / (msa * 1e-3) / 1e6, 1),
                             "stages": stage_table(avga, n, kb, m8a, prim_bytes=24)}
        del da
        log("sort yardstick (rocPRIM radix_sort_pairs on the same keys)")
        out["sort_yardstick"] = sort_yardstick_section(d, n)
        log("learned radius")
        out["learned_radius"] = learned_section(d, (0.5, 0.0))
        lr = out["learned_radius"]["learned_min_confidence_0.5"]
        out["sah_ratio_learned_vs_r8"] = lr["sah_ratio_vs_r8"]
    del d

    if not args.no_secondary:
        log("icosphere_1m1, slivers_2m9, uniform_300k")
        ico = scenes.workload("icosphere_1m1")
        di = B.DevicePrimitives.from_numpy(ico)
        out["icosphere_1m1"] = {f"k{b}": workload_section("icosphere_1m1", di, len(ico), b, "wave32") for b in (32, 64)}
        if not args.no_trace:
            out["traversal"] = trace_section(di, len(ico), B.BuildConfig(mortonBits=32))
        del di
        sl = scenes.workload("slivers_2m9")
        ds = B.DevicePrimitives.from_numpy(sl)
        w32 = workload_section("slivers_2m9", ds, len(sl), 64, "wave32")
        w64 = workload_section("slivers_2m9", ds, len(sl), 64, "wave64")
        out["slivers_2m9"] = {"wave32": w32, "wave64": w64,
                              "sah_ratio_wave64_vs_wave32": {"bvh2": round(w64["bvh2_sah"] / w32["bvh2_sah"], 5),
                                                             "bvh8": round(w64["bvh8_sah"] / w32["bvh8_sah"], 5)}}
        del ds, sl
        un = scenes.workload("uniform_300k")
        du = B.DevicePrimitives.from_numpy(un)
        out["uniform_300k"] = workload_section("uniform_300k", du, len(un), 32, "wave32")
        del du

    if not args.no_batch:
        log("batched builds")
        out["batched"] = batch_section()

    if rank == 0 and not args.no_cpu_baseline:
        log("CPU baseline (oracle, one core, one build of the headline scene)")
        from oracle import oracle as O
        t0 = time.perf_counter()
        ref = O.build(tris, 1, args.bits, 32 if mode == B.MergeMode.WAVE32 else 64, args.radius)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / dt / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
                               "sample": f"one full single-threaded oracle build (bounds, Morton, std::sort, "
                                         f"sequential H-PLOC, BVH8) of the whole {args.workload} scene "
                                         f"({n} triangles, {args.bits}-bit keys), {dt:.1f} s",
                               "stage_ms": [round(float(x), 1) for x in ref["stage_ms"]],
                               "host_cpu": cpu_model(), "host_nproc": cpu_nproc()}
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
