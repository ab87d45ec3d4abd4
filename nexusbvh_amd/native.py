"""ctypes bindings of libnxb.so (the C ABI in include/nxb.h) and of the HIP runtime.

The library is loaded from nexusbvh_amd/lib/. There is no fallback: if the
library is missing, importing a symbol raises NativeLibraryMissing.

GPU processes here do not import torch: torch-ROCm ships its own copy of the
HIP runtime, and libnxb.so links the system one, so device memory is managed
through the same libamdhip64 that libnxb.so uses.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_void_p

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
LIBNXB_PATH = os.path.join(LIB_DIR, "libnxb.so")
HIP_RUNTIME = os.environ.get("NXB_HIP_RUNTIME", "/opt/rocm/lib/libamdhip64.so")

INVALID_IDX = 0xFFFFFFFF
MAX_PRIMS = 1 << 30


class NativeLibraryMissing(RuntimeError):
    pass


# ---- C structs (layouts of include/nxb.h) -----------------------------------
class Float3(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class AABB(Structure):
    _fields_ = [("bmin", Float3), ("bmax", Float3)]


class BVH2(Structure):
    _fields_ = [("nodes", c_void_p), ("node_count", c_uint32), ("prim_count", c_uint32), ("bounds", AABB)]


class BVH8(Structure):
    _fields_ = [("nodes", c_void_p), ("node_count", c_uint32), ("prim_idx", c_void_p),
                ("prim_count", c_uint32), ("bounds", AABB)]


class BuildConfigC(Structure):
    _fields_ = [("prioritize_speed", c_uint32), ("morton_bits", c_uint32), ("search_radius", c_uint32),
                ("merge_mode", c_uint32), ("radius_policy", c_uint32), ("segment_log2", c_uint32),
                ("d_radius_table", c_void_p), ("radius_table_len", c_uint32),
                ("sort_digit_bits", c_uint32), ("spin_limit", c_uint32), ("metrics_flags", c_uint32),
                ("d_mlp_params", c_void_p), ("mlp_min_confidence", c_float), ("slot_assignment", c_uint32),
                ("hploc_blocks_per_cu", c_uint32)]


class MetricsC(Structure):
    _fields_ = [("compute_scene_bounds_time", c_float), ("compute_morton_codes_time", c_float),
                ("radix_sort_time", c_float), ("bvh_build_time", c_float), ("bvh8_conversion_time", c_float),
                ("total_time", c_float), ("bvh2_cost", c_float), ("bvh8_cost", c_float),
                ("average_child_per_node", c_float)]


assert ctypes.sizeof(AABB) == 24
assert ctypes.sizeof(BVH2) == 40
assert ctypes.sizeof(BVH8) == 56
assert ctypes.sizeof(MetricsC) == 36

STATUS = {0: "NXB_SUCCESS", 1: "NXB_ERROR_INVALID_VALUE", 2: "NXB_ERROR_OUT_OF_MEMORY",
          3: "NXB_ERROR_NONFINITE_INPUT", 4: "NXB_ERROR_TOO_LARGE", 5: "NXB_ERROR_TIMEOUT",
          6: "NXB_ERROR_HIP", 7: "NXB_ERROR_NOT_SUPPORTED"}

# Every symbol include/nxb.h declares: name -> (restype, argtypes)
_SIGNATURES = {
    "nxb_default_config": (None, [POINTER(BuildConfigC)]),
    "nxb_status_string": (c_char_p, [c_int]),
    "nxb_version": (c_char_p, []),
    "nxb_build_bvh2": (c_int, [c_void_p, c_int, c_uint32, POINTER(BuildConfigC), POINTER(BVH2),
                               POINTER(MetricsC), c_void_p]),
    "nxb_build_bvh8": (c_int, [c_void_p, c_int, c_uint32, POINTER(BuildConfigC), POINTER(BVH8),
                               POINTER(MetricsC), c_void_p]),
    "nxb_bvh2_to_host": (c_int, [POINTER(BVH2), POINTER(BVH2)]),
    "nxb_free_host_bvh2": (None, [POINTER(BVH2)]),
    "nxb_free_device_bvh2": (c_int, [POINTER(BVH2)]),
    "nxb_free_device_bvh8": (c_int, [POINTER(BVH8)]),
    "nxb_stage_workspace_bytes": (c_size_t, [c_uint32, c_uint32]),
    "nxb_compute_scene_bounds": (c_int, [c_void_p, c_int, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nxb_compute_morton": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    "nxb_radix_sort_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_uint32,
                                     c_uint32, c_uint32, c_void_p, c_size_t, c_void_p]),
    "nxb_build_hploc": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, POINTER(BuildConfigC), c_void_p,
                                c_void_p, c_size_t, c_void_p]),
    "nxb_collapse_bvh8": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p, POINTER(c_uint32), c_uint32,
                                  c_void_p, c_size_t, c_void_p]),
    "nxb_sah_bvh2": (c_int, [POINTER(BVH2), POINTER(c_double), c_void_p]),
    "nxb_sah_bvh8": (c_int, [POINTER(BVH8), POINTER(c_double), c_void_p]),
    "nxb_build_bvh2_noalloc": (c_int, [c_void_p, c_int, c_uint32, POINTER(BuildConfigC), c_void_p, POINTER(BVH2),
                                       c_void_p]),
    "nxb_build_bvh2_indexed": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, POINTER(BuildConfigC), POINTER(BVH2),
                                       c_void_p, c_void_p]),
    "nxb_build_bvh8_indexed": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, POINTER(BuildConfigC), POINTER(BVH8),
                                       c_void_p, c_void_p]),
    "nxb_bvh8_to_host": (c_int, [POINTER(BVH8), POINTER(BVH8)]),
    "nxb_free_host_bvh8": (None, [POINTER(BVH8)]),
    "nxb_trace_closest": (c_int, [POINTER(BVH8), c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    "nxb_trace_closest_indexed": (c_int, [POINTER(BVH8), c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, c_void_p,
                                          c_void_p, c_void_p]),
    "nxb_bvh8_capacity": (c_uint32, [c_uint32]),
    "nxb_segment_features": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_void_p,
                                     c_void_p]),
    "nxb_segment_sah": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_void_p]),
    "nxb_mlp_radius_table": (c_int, [c_void_p, c_void_p, c_uint32, c_float, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "nxb_mlp_gradients": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    "nxb_mlp_adam_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_float, c_float, c_float,
                                  c_float, c_void_p]),
    "nxb_last_learned_stats": (None, [POINTER(c_uint32), POINTER(c_uint32)]),
    "nxb_batch_plan_create": (c_int, [c_void_p, c_int, c_uint32, c_void_p, POINTER(BuildConfigC), c_uint32,
                                      POINTER(c_void_p)]),
    "nxb_batch_plan_run": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nxb_batch_plan_mesh": (c_int, [c_void_p, c_uint32, POINTER(BVH8)]),
    "nxb_batch_plan_node_total": (c_uint32, [c_void_p]),
    "nxb_batch_plan_destroy": (None, [c_void_p]),
}

_lib = None
_hip = None


def lib() -> ctypes.CDLL:
    """libnxb.so with typed entry points; raises NativeLibraryMissing if not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBNXB_PATH):
            raise NativeLibraryMissing(
                f"{LIBNXB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        l = ctypes.CDLL(LIBNXB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGNATURES)


class NxbError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        name = STATUS.get(status, str(status))
        msg = lib().nxb_status_string(status).decode()
        super().__init__(f"{where}: {name} ({msg})" if where else f"{name} ({msg})")


def check(status: int, where: str = "") -> None:
    if status != 0:
        raise NxbError(status, where)


# ---- HIP runtime ------------------------------------------------------------
class HipError(RuntimeError):
    pass


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        h = ctypes.CDLL(HIP_RUNTIME)
        h.hipMalloc.argtypes = [POINTER(c_void_p), c_size_t]
        h.hipFree.argtypes = [c_void_p]
        h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        h.hipMemset.argtypes = [c_void_p, c_int, c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [POINTER(c_int)]
        h.hipGetErrorString.restype = c_char_p
        h.hipGetErrorString.argtypes = [c_int]
        h.hipEventCreate.argtypes = [POINTER(c_void_p)]
        h.hipEventRecord.argtypes = [c_void_p, c_void_p]
        h.hipEventSynchronize.argtypes = [c_void_p]
        h.hipEventElapsedTime.argtypes = [POINTER(c_float), c_void_p, c_void_p]
        h.hipEventDestroy.argtypes = [c_void_p]
        h.hipStreamCreate.argtypes = [POINTER(c_void_p)]
        h.hipStreamDestroy.argtypes = [c_void_p]
        h.hipStreamSynchronize.argtypes = [c_void_p]
        _hip = h
    return _hip


def hip_check(err: int, what: str) -> None:
    if err != 0:
        raise HipError(f"{what}: {hip().hipGetErrorString(err).decode()} ({err})")


def device_count() -> int:
    try:
        n = c_int(0)
        if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
            return 0
        return n.value
    except OSError:
        return 0


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """Owning device allocation (hipMalloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = c_void_p(0)
        if self.nbytes > 0:
            hip_check(hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes), f"hipMalloc({self.nbytes})")

    @property
    def address(self) -> int:
        return self.ptr.value or 0

    @classmethod
    def from_array(cls, arr) -> "DeviceBuffer":
        import numpy as np
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if a.nbytes:
            hip_check(hip().hipMemcpy(b.ptr, a.ctypes.data_as(c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return b

    def to_array(self, dtype, count: int | None = None):
        import numpy as np
        dt = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dt.itemsize
        out = np.empty(count, dtype=dt)
        if out.nbytes:
            hip_check(hip().hipMemcpy(out.ctypes.data_as(c_void_p), self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """Non-owning window [address, address + nbytes) of device memory."""

    def __init__(self, address: int, nbytes: int, owner=None):
        self.ptr = c_void_p(address)
        self.nbytes = int(nbytes)
        self.owner = owner   # keeps the owning buffer alive

    @property
    def address(self) -> int:
        return self.ptr.value or 0


def copy_device(dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int) -> None:
    if nbytes:
        hip_check(hip().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def zero(buf: "DeviceBuffer", nbytes: int | None = None) -> None:
    n = buf.nbytes if nbytes is None else nbytes
    if n:
        hip_check(hip().hipMemset(buf.ptr, 0, n), "hipMemset")


def copy_to_host(ptr: int, nbytes: int, dtype):
    import numpy as np
    dt = np.dtype(dtype)
    out = np.empty(nbytes // dt.itemsize, dtype=dt)
    if nbytes:
        hip_check(hip().hipMemcpy(out.ctypes.data_as(c_void_p), c_void_p(ptr), nbytes, D2H), "hipMemcpy D2H")
    return out


def synchronize() -> None:
    hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream (default priority)."""

    def __init__(self):
        self.s = c_void_p(0)
        hip_check(hip().hipStreamCreate(ctypes.byref(self.s)), "hipStreamCreate")

    def __int__(self) -> int:
        return int(self.s.value)

    def synchronize(self) -> None:
        hip_check(hip().hipStreamSynchronize(self.s), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.s.value:
                hip().hipStreamSynchronize(self.s)
                hip().hipStreamDestroy(self.s)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.e = c_void_p(0)
        hip_check(hip().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream: int | None = None) -> None:
        hip_check(hip().hipEventRecord(self.e, c_void_p(stream or 0)), "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        hip_check(hip().hipEventSynchronize(end.e), "hipEventSynchronize")
        t = c_float(0)
        hip_check(hip().hipEventElapsedTime(ctypes.byref(t), self.e, end.e), "hipEventElapsedTime")
        return t.value

    def __del__(self):
        try:
            if self.e.value:
                hip().hipEventDestroy(self.e)
        except Exception:
            pass
