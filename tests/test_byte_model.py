"""The algorithmic byte model that roofline fractions divide by: the Python
copy (nexusbvh_amd.builder.stage_bytes_per_prim, used by bench.py and the
Python metrics report) and the C++ copy (NXB::StageBytesPerPrim in
include/NXB/BVHBuildMetrics.h, used by the drop-in C++ API) give the same
figures, and those figures are the per-stage byte counts the kernels' data
layout fixes (36-B Triangle, 24-B AABB, 32-B BVH2 node, 64-B wide record,
80-B BVH8 node)."""
import itertools
import os
import subprocess

import pytest

from nexusbvh_amd.builder import stage_bytes_per_prim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Per-stage bytes per primitive for 36-B triangles, counted field by field.
# bounds: 36-B triangle read, 32-B leaf node written.
# Morton: 24-B leaf box read; key (8 B or 4 B) and 4-B index written.
# sort: 8 (64-bit keys) or 4 (32-bit) digit passes, each reading and writing
#       key + index, plus one key read for the digit counts.
# BVH2: key and index read (k + 4), leaf box gathered (24), node written
#       (32, or the 64-B wide record in a BVH8 build), cluster index written
#       (4), parent word reset (4) and two parent exchanges (8).
# BVH8: 64-B wide record and 4-B primitive index per primitive; 80-B node
#       plus 16-B work item (written and read) per wide node.
EXPECTED_TRIANGLE = {
    (8, False): (68, 36, 200, 8 + 4 + 24 + 32 + 4 + 4 + 8),
    (8, True): (68, 36, 200, 8 + 4 + 24 + 64 + 4 + 4 + 8),
    (4, False): (68, 32, 4 * 2 * 8 + 4, 4 + 4 + 24 + 32 + 4 + 4 + 8),
    (4, True): (68, 32, 4 * 2 * 8 + 4, 4 + 4 + 24 + 64 + 4 + 4 + 8),
}


@pytest.mark.parametrize("key_bytes,bvh8", sorted(EXPECTED_TRIANGLE))
def test_python_model_matches_field_counts(key_bytes, bvh8):
    want = EXPECTED_TRIANGLE[(key_bytes, bvh8)]
    got = tuple(stage_bytes_per_prim(s, 36, key_bytes, 0.0, bvh8) for s in range(4))
    assert got == want
    # AABB primitives only change the bounds stage (24-B box in).
    assert stage_bytes_per_prim(0, 24, key_bytes, 0.0, bvh8) == 24 + 32
    for m in (0.0, 0.25, 1.0 / 7.0, 4.0 / 7.0):
        assert stage_bytes_per_prim(4, 36, key_bytes, m, bvh8) == pytest.approx(64 + 4 + (80 + 16) * m)


def test_cpp_model_matches_python(tmp_path):
    """Compile NXB::StageBytesPerPrim with g++ and compare every stage over
    both key widths, both primitive types, BVH2/BVH8 builds and several
    wide-node ratios."""
    src = tmp_path / "bytes.cpp"
    src.write_text(
        "#include <NXB/BVHBuildMetrics.h>\n"
        "#include <cstdio>\n"
        "int main() {\n"
        "  const unsigned prim[2] = {36, 24}, key[2] = {8, 4};\n"
        "  const double m8[3] = {0.0, 0.25, 0.5714285714285714};\n"
        "  for (unsigned p : prim) for (unsigned k : key) for (int w = 0; w < 2; ++w)\n"
        "    for (double m : m8) for (int st = 0; st < NXB::kBuildStages; ++st) {\n"
        "      NXB::BuildShape s; s.primBytes = p; s.keyBytes = k; s.bvh8 = w != 0;\n"
        "      std::printf(\"%u %u %d %.17g %d %.17g\\n\", p, k, w, m, st, NXB::StageBytesPerPrim(st, s, m));\n"
        "    }\n"
        "  return 0;\n"
        "}\n")
    exe = tmp_path / "bytes"
    r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [line.split() for line in out if line]
    assert len(rows) == 2 * 2 * 2 * 3 * 5
    seen = set()
    for p, k, w, m, st, v in rows:
        p, k, w, st, m, v = int(p), int(k), int(w), int(st), float(m), float(v)
        assert v == stage_bytes_per_prim(st, p, k, m, bool(w)), (p, k, w, m, st)
        seen.add((p, k, w, st))
    assert seen == set(itertools.product((36, 24), (8, 4), (0, 1), range(5)))
