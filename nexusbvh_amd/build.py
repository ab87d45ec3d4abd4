"""Builds the native libraries in-tree.

  nexusbvh_amd/lib/libnxb.so         HIP kernels (gfx950) + host orchestration, C ABI of include/nxb.h
  nexusbvh_amd/lib/libnxb_scenes.so  synthetic scene generators (host C++)
  nexusbvh_amd/lib/libnxb_yardstick.so  rocPRIM's radix sort behind a C ABI (a yardstick bench.py
                                     times beside the pipeline's sort; not linked into libnxb.so)
  oracle/lib/libnxb_oracle.so        CPU oracle (test infrastructure, host C++)

Device code is compiled without FMA contraction and with correctly rounded
FP32 division, the same arithmetic the CPU oracle is compiled with.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
ORACLE = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE, "lib")
OBJ = os.path.join(PKG, "lib", "obj")

HIP_SOURCES = ["setup.hip", "radix_sort.hip", "hploc.hip", "bvh8.hip", "sah.hip", "learned.hip", "batch.hip", "trace.hip",
               "api.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NXB_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip() -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(HIP_SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    out = os.path.join(LIB, "libnxb.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_yardstick() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libnxb_yardstick.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "rocprim_yardstick.hip"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cpp_benchmark() -> str:
    """The C++ drop-in user (NXB/BVHBuilder.h), linked against libnxb.so."""
    out = os.path.join(LIB, "nxb_benchmark")
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "nxb_benchmark.cpp"),
          "-L", LIB, "-lnxb", "-lnxb_scenes", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_scenes() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libnxb_scenes.so")
    _run(["g++", *HOST_FLAGS, "-shared", "-pthread", os.path.join(CSRC, "scenes.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> str:
    os.makedirs(ORACLE_LIB, exist_ok=True)
    out = os.path.join(ORACLE_LIB, "libnxb_oracle.so")
    _run(["g++", *HOST_FLAGS, "-shared", os.path.join(ORACLE, "nxb_oracle.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose: bool = True) -> None:
    with ThreadPoolExecutor(max_workers=4) as ex:
        futs = [ex.submit(build_hip), ex.submit(build_scenes), ex.submit(build_oracle), ex.submit(build_yardstick)]
        outs = [f.result() for f in futs]
    outs.append(build_cpp_benchmark())
    if verbose:
        for o in outs:
            print("built", os.path.relpath(o, ROOT))


def clean() -> None:
    shutil.rmtree(LIB, ignore_errors=True)
    shutil.rmtree(ORACLE_LIB, ignore_errors=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "clean":
        clean()
    else:
        build_all()
