"""Python mirror of the NexusBVH builder API over libnxb.so.

Same names, argument meaning and ownership as NexusBVH/include/NXB/BVHBuilder.h:18-55
and BVHBuildMetrics.h:7-107:

    BuildBVH2(primitives, primCount, buildConfig=BuildConfig(), buildMetrics=None) -> BVH2
    BuildBVH8(primitives, primCount, buildConfig=BuildConfig(), buildMetrics=None) -> BVH8
    ToHost(bvh2) -> host BVH2 (numpy nodes)        FreeHostBVH(bvh2)
    FreeDeviceBVH(bvh2 | bvh8)                     BenchmarkBuild(func, warmup, measured, *args)

`primitives` is a DevicePrimitives (device buffer + primitive type): the type
plays the role of the PrimT template argument (Triangle or AABB). Errors are
raised as NxbError carrying the nxb_status code (the reference exits).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from . import native
from .native import BVH2, BVH8, AABB, BuildConfigC, DeviceBuffer, MetricsC, check, lib

# numpy views of the device formats (include/nxb.h)
BVH2_NODE_DTYPE = np.dtype([("bmin", "<f4", (3,)), ("bmax", "<f4", (3,)), ("left", "<u4"), ("right", "<u4")])
BVH8_NODE_DTYPE = np.dtype([("p", "<f4", (3,)), ("e", "u1", (3,)), ("imask", "u1"), ("child_base", "<u4"),
                            ("prim_base", "<u4"), ("meta", "u1", (8,)), ("qlox", "u1", (8,)), ("qloy", "u1", (8,)),
                            ("qloz", "u1", (8,)), ("qhix", "u1", (8,)), ("qhiy", "u1", (8,)), ("qhiz", "u1", (8,))])
assert BVH2_NODE_DTYPE.itemsize == 32 and BVH8_NODE_DTYPE.itemsize == 80


class PrimType(IntEnum):
    """NXB::PrimType (NexusBVH/include/NXB/BVH.h:10-14)."""
    AABB = 0
    TRIANGLE = 1


class MergeMode(IntEnum):
    WAVE32 = 0   # the reference schedule (16 per child, 32-cluster window)
    WAVE64 = 1   # native 64-lane schedule (32 per child, 64-cluster window)


class SlotAssignment(IntEnum):
    GREEDY = 0    # the reference's assignment (WideConverter.cu:106-153)
    AUCTION = 1   # its compiled-out auction (WideConverter.cu:52-102)


class RadiusPolicy(IntEnum):
    FIXED = 0
    TABLE = 1
    LEARNED = 2   # per-segment radius proposed by the MLP (mlpParams)


_CONFIG_CACHE: dict = {}


@dataclass
class BuildConfig:
    """BuildConfig (NexusBVH/include/NXB/BuildConfig.h:6-12) plus the options of this build."""
    prioritizeSpeed: bool = False      # True: 32-bit Morton keys
    mortonBits: int = 0                # 0: follow prioritizeSpeed; else 32 or 64
    searchRadius: int = 8
    mergeMode: int = MergeMode.WAVE32
    radiusPolicy: int = RadiusPolicy.FIXED
    segmentLog2: int = 12
    radiusTable: DeviceBuffer | None = None   # uint8 per segment (device)
    radiusTableLen: int = 0
    sortDigitBits: int = 0
    spinLimit: int = 0
    metricsFlags: int = 0              # 1: stage times only (no SAH in the metrics)
    mlpParams: DeviceBuffer | None = None     # learned policy: NXB_MLP_PARAMS float32 (device)
    mlpMinConfidence: float = 0.5
    slotAssignment: int = SlotAssignment.GREEDY   # BVH8 child slots
    hplocBlocksPerCu: int = 0          # H-PLOC resident blocks per CU; 0: default (12)

    def to_c(self) -> BuildConfigC:
        """The C struct (cached per distinct field values; the library only reads it)."""
        key = (self.prioritizeSpeed, self.mortonBits, self.searchRadius, self.mergeMode, self.radiusPolicy,
               self.segmentLog2, self.radiusTable.address if self.radiusTable is not None else None,
               self.radiusTableLen, self.sortDigitBits, self.spinLimit, self.metricsFlags,
               self.mlpParams.address if self.mlpParams is not None else None, self.mlpMinConfidence,
               self.slotAssignment, self.hplocBlocksPerCu)
        c = _CONFIG_CACHE.get(key)
        if c is None:
            if len(_CONFIG_CACHE) > 256:
                _CONFIG_CACHE.clear()
            c = _CONFIG_CACHE[key] = self._make_c()
        return c

    def _make_c(self) -> BuildConfigC:
        c = BuildConfigC()
        lib().nxb_default_config(ctypes.byref(c))
        c.prioritize_speed = 1 if self.prioritizeSpeed else 0
        c.morton_bits = int(self.mortonBits)
        c.search_radius = int(self.searchRadius)
        c.merge_mode = int(self.mergeMode)
        c.radius_policy = int(self.radiusPolicy)
        c.segment_log2 = int(self.segmentLog2)
        c.d_radius_table = self.radiusTable.address if self.radiusTable is not None else None
        c.radius_table_len = int(self.radiusTableLen)
        c.sort_digit_bits = int(self.sortDigitBits)
        c.spin_limit = int(self.spinLimit)
        c.metrics_flags = int(self.metricsFlags)
        c.d_mlp_params = self.mlpParams.address if self.mlpParams is not None else None
        c.mlp_min_confidence = float(self.mlpMinConfidence)
        c.slot_assignment = int(self.slotAssignment)
        c.hploc_blocks_per_cu = int(self.hplocBlocksPerCu)
        return c


@dataclass
class BVHBuildMetrics:
    """BVHBuildMetrics (NexusBVH/include/NXB/BVHBuildMetrics.h:7-50)."""
    computeSceneBoundsTime: float = 0.0
    computeMortonCodesTime: float = 0.0
    radixSortTime: float = 0.0
    bvhBuildTime: float = 0.0
    bvh8ConversionTime: float = 0.0
    totalTime: float = 0.0
    bvh2Cost: float = 0.0
    bvh8Cost: float = 0.0
    averageChildPerNode: float = 0.0

    _FIELDS = ("computeSceneBoundsTime", "computeMortonCodesTime", "radixSortTime", "bvhBuildTime",
               "bvh8ConversionTime", "totalTime", "bvh2Cost", "bvh8Cost", "averageChildPerNode")

    def _load(self, m: MetricsC) -> None:
        vals = (m.compute_scene_bounds_time, m.compute_morton_codes_time, m.radix_sort_time, m.bvh_build_time,
                m.bvh8_conversion_time, m.total_time, m.bvh2_cost, m.bvh8_cost, m.average_child_per_node)
        for k, v in zip(self._FIELDS, vals):
            setattr(self, k, float(v))

    def __iadd__(self, o: "BVHBuildMetrics") -> "BVHBuildMetrics":
        for k in self._FIELDS:
            setattr(self, k, getattr(self, k) + getattr(o, k))
        return self

    def __truediv__(self, d: float) -> "BVHBuildMetrics":
        r = BVHBuildMetrics()
        for k in self._FIELDS:
            setattr(r, k, getattr(self, k) / d)
        return r


@dataclass
class DevicePrimitives:
    """Device-resident primitives: Triangle (36 B) or AABB (24 B) records."""
    buffer: DeviceBuffer
    count: int
    prim_type: PrimType

    @classmethod
    def from_numpy(cls, arr: np.ndarray, prim_type: PrimType | None = None) -> "DevicePrimitives":
        a = np.ascontiguousarray(arr, dtype=np.float32)
        per = a.shape[-1] if a.ndim == 2 else None
        if prim_type is None:
            if per == 9:
                prim_type = PrimType.TRIANGLE
            elif per == 6:
                prim_type = PrimType.AABB
            else:
                raise ValueError("expected an (n, 9) triangle or (n, 6) box array")
        floats = 9 if prim_type == PrimType.TRIANGLE else 6
        count = a.size // floats
        return cls(DeviceBuffer.from_array(a.reshape(-1)), count, PrimType(prim_type))

    @property
    def address(self) -> int:
        return self.buffer.address

    def view(self, first: int, count: int) -> "DevicePrimitives":
        """Primitives [first, first + count) without a copy."""
        per = 36 if self.prim_type == PrimType.TRIANGLE else 24
        return DevicePrimitives(native.DeviceView(self.address + first * per, count * per, self.buffer), count,
                                self.prim_type)


def _aabb_tuple(b: AABB):
    return ((b.bmin.x, b.bmin.y, b.bmin.z), (b.bmax.x, b.bmax.y, b.bmax.z))


def _prims(primitives) -> tuple[int, int]:
    if isinstance(primitives, DevicePrimitives):
        return primitives.address, int(primitives.prim_type)
    raise TypeError("primitives must be DevicePrimitives")


def _stream(stream) -> int | None:
    return int(stream) if stream else None


def BuildBVH2(primitives, primCount: int, buildConfig: BuildConfig | None = None,
              buildMetrics: BVHBuildMetrics | None = None, stream=None) -> BVH2:
    """NXB::BuildBVH2<PrimT> (BVHBuilder.h:18-19): returns a device BVH2 handle."""
    ptr, ptype = _prims(primitives)
    cfg = (buildConfig or BuildConfig()).to_c()
    out = BVH2()
    m = MetricsC()
    st = lib().nxb_build_bvh2(ptr, ptype, primCount, ctypes.byref(cfg), ctypes.byref(out),
                              ctypes.byref(m) if buildMetrics is not None else None, _stream(stream))
    check(st, "BuildBVH2")
    if buildMetrics is not None:
        buildMetrics._load(m)
    return out


def BuildBVH8(primitives, primCount: int, buildConfig: BuildConfig | None = None,
              buildMetrics: BVHBuildMetrics | None = None, stream=None) -> BVH8:
    """NXB::BuildBVH8<PrimT> (BVHBuilder.h:30-31): returns a device BVH8 handle."""
    ptr, ptype = _prims(primitives)
    cfg = (buildConfig or BuildConfig()).to_c()
    out = BVH8()
    m = MetricsC()
    st = lib().nxb_build_bvh8(ptr, ptype, primCount, ctypes.byref(cfg), ctypes.byref(out),
                              ctypes.byref(m) if buildMetrics is not None else None, _stream(stream))
    check(st, "BuildBVH8")
    if buildMetrics is not None:
        buildMetrics._load(m)
    return out


@dataclass
class DeviceIndexedMesh:
    """Device-resident indexed triangle mesh: vertices (3 floats each) and
    three uint32 vertex indices per triangle."""
    vertices: DeviceBuffer
    indices: DeviceBuffer
    vertexCount: int
    triCount: int

    @classmethod
    def from_numpy(cls, vertices: np.ndarray, indices: np.ndarray) -> "DeviceIndexedMesh":
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        return cls(DeviceBuffer.from_array(v.reshape(-1)), DeviceBuffer.from_array(i.reshape(-1)), len(v), len(i))


def BuildBVH8Indexed(mesh: DeviceIndexedMesh, buildConfig: BuildConfig | None = None,
                     buildMetrics: BVHBuildMetrics | None = None, stream=None) -> BVH8:
    """BVH8 of an indexed mesh (no counterpart in the reference); equal to
    BuildBVH8 on the expanded triangles."""
    cfg = (buildConfig or BuildConfig()).to_c()
    out = BVH8()
    m = MetricsC()
    check(lib().nxb_build_bvh8_indexed(mesh.vertices.address, mesh.vertexCount, mesh.indices.address, mesh.triCount,
                                       ctypes.byref(cfg), ctypes.byref(out),
                                       ctypes.byref(m) if buildMetrics is not None else None, _stream(stream)),
          "BuildBVH8Indexed")
    if buildMetrics is not None:
        buildMetrics._load(m)
    return out


def BuildBVH2Indexed(mesh: DeviceIndexedMesh, buildConfig: BuildConfig | None = None,
                     buildMetrics: BVHBuildMetrics | None = None, stream=None) -> BVH2:
    """BVH2 of an indexed mesh (no counterpart in the reference)."""
    cfg = (buildConfig or BuildConfig()).to_c()
    out = BVH2()
    m = MetricsC()
    check(lib().nxb_build_bvh2_indexed(mesh.vertices.address, mesh.vertexCount, mesh.indices.address, mesh.triCount,
                                       ctypes.byref(cfg), ctypes.byref(out),
                                       ctypes.byref(m) if buildMetrics is not None else None, _stream(stream)),
          "BuildBVH2Indexed")
    if buildMetrics is not None:
        buildMetrics._load(m)
    return out


@dataclass
class HostBVH2:
    nodes: np.ndarray            # BVH2_NODE_DTYPE
    nodeCount: int
    primCount: int
    bounds: tuple
    _c: BVH2 = field(default=None, repr=False)

    @property
    def root(self) -> int:
        return self.nodeCount - 1


@dataclass
class HostBVH8:
    nodes: np.ndarray            # BVH8_NODE_DTYPE, root at 0
    primIdx: np.ndarray          # uint32
    nodeCount: int
    primCount: int
    bounds: tuple


def ToHost(deviceBvh):
    """NXB::ToHost(BVH2) (BVHBuilder.h:40); for a BVH8 (no counterpart in the
    reference) the nodes and primIdx."""
    if isinstance(deviceBvh, BVH8):
        h8 = BVH8()
        check(lib().nxb_bvh8_to_host(ctypes.byref(deviceBvh), ctypes.byref(h8)), "ToHost")
        try:
            nb, pc = h8.node_count, h8.prim_count
            nodes = (np.frombuffer(bytes((ctypes.c_char * (nb * 80)).from_address(h8.nodes)), dtype=BVH8_NODE_DTYPE)
                     .copy() if nb else np.zeros(0, dtype=BVH8_NODE_DTYPE))
            prims = (np.frombuffer(bytes((ctypes.c_char * (pc * 4)).from_address(h8.prim_idx)), dtype=np.uint32)
                     .copy() if pc else np.zeros(0, dtype=np.uint32))
        finally:
            lib().nxb_free_host_bvh8(ctypes.byref(h8))
        return HostBVH8(nodes, prims, nb, pc, _aabb_tuple(h8.bounds))
    h = BVH2()
    check(lib().nxb_bvh2_to_host(ctypes.byref(deviceBvh), ctypes.byref(h)), "ToHost")
    n = h.node_count
    if n:
        buf = (ctypes.c_char * (n * 32)).from_address(h.nodes)
        nodes = np.frombuffer(bytes(buf), dtype=BVH2_NODE_DTYPE).copy()
    else:
        nodes = np.zeros(0, dtype=BVH2_NODE_DTYPE)
    return HostBVH2(nodes, n, h.prim_count, _aabb_tuple(h.bounds), h)


def FreeHostBVH(hostBvh: HostBVH2) -> None:
    """NXB::FreeHostBVH (BVHBuilder.h:45)."""
    if hostBvh._c is not None:
        lib().nxb_free_host_bvh2(ctypes.byref(hostBvh._c))
        hostBvh._c = None


def FreeDeviceBVH(deviceBvh) -> None:
    """NXB::FreeDeviceBVH (BVHBuilder.h:50,55) for a BVH2 or BVH8 handle."""
    if isinstance(deviceBvh, BVH2):
        check(lib().nxb_free_device_bvh2(ctypes.byref(deviceBvh)), "FreeDeviceBVH")
    elif isinstance(deviceBvh, BVH8):
        check(lib().nxb_free_device_bvh8(ctypes.byref(deviceBvh)), "FreeDeviceBVH")
    else:
        raise TypeError("expected BVH2 or BVH8")


def bvh8_to_host(deviceBvh: BVH8) -> tuple[np.ndarray, np.ndarray]:
    """Copies a device BVH8 (nodes, primIdx) to numpy arrays."""
    nodes = native.copy_to_host(deviceBvh.nodes or 0, deviceBvh.node_count * 80, np.uint8).view(BVH8_NODE_DTYPE)
    prims = native.copy_to_host(deviceBvh.prim_idx or 0, deviceBvh.prim_count * 4, np.uint32)
    return nodes, prims


def bvh_bounds(bvh) -> tuple:
    return _aabb_tuple(bvh.bounds)


def sah_bvh2(bvh: BVH2) -> float:
    c = ctypes.c_double(0)
    check(lib().nxb_sah_bvh2(ctypes.byref(bvh), ctypes.byref(c), None), "sah_bvh2")
    return c.value


def sah_bvh8(bvh: BVH8) -> float:
    c = ctypes.c_double(0)
    check(lib().nxb_sah_bvh8(ctypes.byref(bvh), ctypes.byref(c), None), "sah_bvh8")
    return c.value


# (label, BVHBuildMetrics field) of each stage, in build order
BUILD_STAGES = (("scene bounds", "computeSceneBoundsTime"), ("Morton codes", "computeMortonCodesTime"),
                ("radix sort", "radixSortTime"), ("BVH2 H-PLOC", "bvhBuildTime"),
                ("BVH8 collapse", "bvh8ConversionTime"))
HBM_PEAK_GBS = 8000.0


def stage_bytes_per_prim(stage: int, prim_bytes: int, key_bytes: int, bvh8_nodes_per_prim: float = 0.0,
                         bvh8: bool = False) -> float:
    """Algorithmic HBM bytes per primitive of build stage `stage` (order of
    BUILD_STAGES): what the stage must read and write at least.
    bounds: primitive in, 32-B leaf out; Morton: 24-B leaf box in, key and
    index out; sort: key + index read and written per 8-bit digit pass (4
    passes for 32-bit keys, 8 for 64-bit) plus one read of the keys for the
    digit counts; BVH2: sorted key and index, leaf box gather, the 32-B node
    (in a BVH8 build, `bvh8`, its 64-B wide record instead), cluster index,
    parent reset and two parent exchanges; BVH8: one 64-B wide record per
    primitive, the primitive index, and per wide node its 80 bytes and its
    work item written and read (16 B)."""
    k = float(key_bytes)
    return (prim_bytes + 32.0, 24.0 + k + 4.0, 2.0 * (k + 4.0) * (4.0 if key_bytes == 4 else 8.0) + k,
            k + (108.0 if bvh8 else 76.0), 68.0 + 96.0 * bvh8_nodes_per_prim)[stage]


def BenchmarkBuild(func, warmupIterations: int, measuredIterations: int, *args, verbose: bool = True):
    """Warm up, build measuredIterations times with metrics and return the
    mean metrics (the contract of BenchmarkBuild, BVHBuildMetrics.h:62-107).
    args are BuildBVH2/BuildBVH8's (primitives, primCount[, config]). The
    report gives, per stage, mean and fastest ms, million primitives per
    second and algorithmic GB/s (stage_bytes_per_prim) with its share of the
    8 TB/s HBM peak, and the median and fastest whole-build time."""
    for _ in range(warmupIterations):
        FreeDeviceBVH(func(*args, BVHBuildMetrics()))
    runs = []
    for _ in range(measuredIterations):
        m = BVHBuildMetrics()
        FreeDeviceBVH(func(*args, m))
        runs.append(m)
    mean = BVHBuildMetrics()
    for m in runs:
        mean += m
    if runs:
        mean = mean / float(len(runs))
    if not verbose:
        return mean
    wide = func is BuildBVH8 or func is BuildBVH8Indexed
    prims = args[0] if args else None
    if isinstance(prims, DeviceIndexedMesh):
        count = prims.triCount
    else:
        count = int(args[1]) if len(args) > 1 and isinstance(args[1], (int, np.integer)) else 0
    cfg = next((a for a in args if isinstance(a, BuildConfig)), BuildConfig())
    key_bytes = (cfg.mortonBits or (32 if cfg.prioritizeSpeed else 64)) // 8
    # an indexed triangle reads its three 4-B indices and three 12-B vertices
    prim_bytes = 24 if getattr(prims, "prim_type", PrimType.TRIANGLE) == PrimType.AABB else 36
    if isinstance(prims, DeviceIndexedMesh):
        prim_bytes = 48
    per8 = ((count - 1) / (mean.averageChildPerNode - 1.0) / count
            if wide and count > 1 and mean.averageChildPerNode > 1.0 else 0.0)
    print(f"--- NXB build benchmark: {'BVH8' if wide else 'BVH2'}, {count} primitives, {8 * key_bytes}-bit Morton, "
          f"{warmupIterations} warm-up + {measuredIterations} measured builds ---")
    print(f"{'stage':<16}{'mean ms':>10}{'min ms':>10}{'Mprim/s':>12}{'GB/s':>10}{'of HBM':>10}")
    rows = [(label, field, stage_bytes_per_prim(i, prim_bytes, key_bytes, per8, wide))
            for i, (label, field) in enumerate(BUILD_STAGES[:5 if wide else 4])]
    rows.append(("whole build", "totalTime", sum(r[2] for r in rows)))
    for label, field, per in rows:
        vals = [getattr(m, field) for m in runs]
        avg = sum(vals) / max(1, len(vals))
        line = f"{label:<16}{avg:>10.5f}{min(vals, default=0.0):>10.5f}"
        if avg > 0 and count > 0:
            gbs = per * count / (avg * 1e6)
            line += f"{count / (avg * 1e3):>12.1f}{gbs:>10.1f}{100 * gbs / HBM_PEAK_GBS:>9.1f}%"
        print(line)
    totals = sorted(m.totalTime for m in runs)
    if totals:
        print(f"build time median {totals[len(totals) // 2]:.4f} ms, fastest {totals[0]:.4f} ms")
    print(f"SAH: BVH2 {mean.bvh2Cost:.4f}" +
          (f", BVH8 {mean.bvh8Cost:.4f}, children per BVH8 node {mean.averageChildPerNode:.4f}" if wide else ""))
    return mean
