"""C ABI checks that need no GPU: the library loads, exports every entry
point include/nxb.h declares, and rejects invalid arguments with status codes
before touching the device."""
import ctypes
import os
import re

import pytest

from nexusbvh_amd import native
from nexusbvh_amd.native import BVH2, BVH8, BuildConfigC, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nxb.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nxb_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    names = header_functions()
    assert len(names) >= 18
    l = lib()
    for name in names:
        assert hasattr(l, name), name
    assert sorted(native.declared_symbols()) == names


def test_struct_layouts():
    assert ctypes.sizeof(native.AABB) == 24
    assert ctypes.sizeof(native.MetricsC) == 9 * 4
    assert ctypes.sizeof(BuildConfigC) == 72  # sizeof(nxb_build_config) with gcc x86-64
    from nexusbvh_amd.builder import BVH2_NODE_DTYPE, BVH8_NODE_DTYPE
    assert BVH2_NODE_DTYPE.itemsize == 32 and BVH8_NODE_DTYPE.itemsize == 80


def test_default_config_is_reference_default():
    c = BuildConfigC()
    lib().nxb_default_config(ctypes.byref(c))
    assert c.prioritize_speed == 0 and c.morton_bits == 0  # 64-bit Morton (BuildConfig.h:11)
    assert c.search_radius == 8 and c.merge_mode == 0


def test_status_strings():
    for code in range(8):
        assert lib().nxb_status_string(code)
    assert b"gfx950" in lib().nxb_version()


def test_bvh8_capacity():
    cap = lib().nxb_bvh8_capacity
    assert cap(0) == 0 and cap(1) == 1 and cap(2) == 1 and cap(7) == 4
    for n in (1, 2, 3, 100, 300000, 28100000):
        assert cap(n) == 1 + (4 * n - 2) // 7


def test_invalid_arguments_return_status_without_gpu():
    l = lib()
    out2, out8 = BVH2(), BVH8()
    c = BuildConfigC()
    l.nxb_default_config(ctypes.byref(c))
    # null output
    assert l.nxb_build_bvh2(ctypes.c_void_p(16), 1, 10, ctypes.byref(c), None, None, None) == 1
    # bad primitive type
    assert l.nxb_build_bvh2(ctypes.c_void_p(16), 7, 10, ctypes.byref(c), ctypes.byref(out2), None, None) == 1
    # null primitives with n > 0
    assert l.nxb_build_bvh8(None, 1, 10, ctypes.byref(c), ctypes.byref(out8), None, None) == 1
    # too many primitives
    assert l.nxb_build_bvh8(ctypes.c_void_p(16), 1, (1 << 30) + 1, ctypes.byref(c), ctypes.byref(out8), None,
                            None) == 4
    # bad radius / morton bits / merge mode
    for field, val in (("search_radius", 0), ("search_radius", 17), ("morton_bits", 48), ("merge_mode", 5)):
        bad = BuildConfigC()
        l.nxb_default_config(ctypes.byref(bad))
        setattr(bad, field, val)
        assert l.nxb_build_bvh2(ctypes.c_void_p(16), 1, 10, ctypes.byref(bad), ctypes.byref(out2), None, None) == 1
    # empty scene: success, empty BVH (no device work)
    assert l.nxb_build_bvh2(None, 1, 0, ctypes.byref(c), ctypes.byref(out2), None, None) == 0
    assert out2.node_count == 0 and not out2.nodes
    assert l.nxb_build_bvh8(None, 0, 0, None, ctypes.byref(out8), None, None) == 0
    assert out8.node_count == 0
    # sort argument checks
    assert l.nxb_radix_sort_pairs(None, None, None, None, 10, 48, 0, 32, 8, None, 0, None) == 1
    assert l.nxb_radix_sort_pairs(None, None, None, None, 10, 32, 4, 2, 8, None, 0, None) == 1
    assert l.nxb_radix_sort_pairs(None, None, None, None, 10, 32, 0, 32, 9, None, 0, None) == 1


def test_stage_workspace_bytes_grows_with_n():
    f = lib().nxb_stage_workspace_bytes
    assert f(1000, 32) < f(1000, 64) < f(100000, 64)


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "LIBNXB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(native.NativeLibraryMissing):
        native.lib()


def test_dropin_cpp_headers_compile_with_gcc(tmp_path):
    """NXB/*.h (the reference's C++ API over the C ABI) compile for host code,
    including their layout static_asserts against include/nxb.h."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text(
        "#include <NXB/BVHBuilder.h>\n"
        "static_assert(sizeof(NXB::BVH8::Node) == 80, \"node\");\n"
        "int main() {\n"
        "  NXB::BuildConfig c;\n"
        "  auto f = &NXB::BuildBVH8<NXB::Triangle>;\n"
        "  auto g = &NXB::BuildBVH2<NXB::AABB>;\n"
        "  (void)f; (void)g;\n"
        "  return c.prioritizeSpeed ? 1 : 0;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I", os.path.join(ROOT, "include"), "-fsyntax-only", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
