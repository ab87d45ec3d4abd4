"""Writes the golden fixtures of tests/golden/.

morton_vectors.json: known-answer Morton codes computed here with plain
  Python integers (float32 quantization by truncation, bit-by-bit interleave,
  x -> bit 0, y -> bit 1, z -> bit 2), independent of the oracle code.
small_scene.json: a 48-triangle scene and the oracle's outputs for it
  (sorted keys, permutation, canonical BVH2, BVH8 node count, SAH), used as a
  regression fixture for the oracle and for the GPU path.

Run: python tests/golden/make_golden.py   (needs the oracle library built)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def morton_py(x, y, z, bits):
    per = 10 if bits == 32 else 21
    scale = np.float32(1023.0 if bits == 32 else 2097151.0)

    def q(v):
        f = np.float32(v) * scale
        if not (f > 0):
            return 0
        return int(min(f, scale))
    qs = [q(x), q(y), q(z)]
    code = 0
    for b in range(per):
        for axis in range(3):
            code |= ((qs[axis] >> b) & 1) << (3 * b + axis)
    return code


def main():
    pts = [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0.5, 0.25, 0.75), (0.1, 0.2, 0.3), (0.999, 0.001, 0.5),
           (1.0 / 3.0, 2.0 / 3.0, 0.125), (0.7071, 0.3333, 0.0001), (0.0625, 0.9375, 0.4375)]
    vec = [[x, y, z, f"{morton_py(x, y, z, 32):x}", f"{morton_py(x, y, z, 64):x}"] for x, y, z in pts]
    with open(os.path.join(HERE, "morton_vectors.json"), "w") as f:
        json.dump({"note": "x,y,z normalized; keys hex (32-bit, 64-bit)", "vectors": vec}, f, indent=1)

    from nexusbvh_amd import scenes
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_oracle import _oracle_canon, _oracle_pipeline
    tris = scenes.slivers(48, seed=2024) + np.float32(1.0)
    n = len(tris)
    cases = []
    for bits, width, radius in [(32, 32, 8), (64, 32, 8), (64, 64, 8), (64, 32, 2)]:
        full, sk, perm, scene = _oracle_pipeline(tris, bits, width, radius)
        b8, prim = O.collapse_bvh8(full, n)
        cases.append({"bits": bits, "width": width, "radius": radius,
                      "sorted_keys": [f"{int(k):x}" for k in sk], "perm": [int(p) for p in perm],
                      "bvh2_canon": _oracle_canon(full, n), "sah2": O.sah_bvh2(full, 2 * n - 1, scene),
                      "bvh8_nodes": int(len(b8)), "sah8": O.sah_bvh8(b8, len(b8), scene)})
    with open(os.path.join(HERE, "small_scene.json"), "w") as f:
        json.dump({"triangles": tris.tolist(), "cases": cases}, f)


if __name__ == "__main__":
    main()
