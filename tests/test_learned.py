"""Learned search radius: the numpy checker pinned on the CPU, and the gfx950
kernels (features, per-segment SAH, MLP inference/gradients, Adam, the
learned build policy) compared with it on the GPU.

Tolerances: the MLP kernels round operands to BF16 and accumulate in FP32
in MFMA order, the checker in numpy order, so logits and gradients agree to
~1e-2 relative (a value on a BF16 rounding boundary can round the other
way); features agree to 1e-4; per-segment SAH sums the same FP32 terms in
FP64, 1e-9 relative.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import learned as L
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(__file__))


def say(*a):
    print(*a, flush=True)


# ---- CPU: pin the checker -------------------------------------------------------

def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        one: one,
        one + 2.0 ** -8: one,                        # tie, even mantissa stays
        one + 3 * 2.0 ** -8: one + 2.0 ** -6,        # tie, rounds up to even
        one + 2.0 ** -8 + 2.0 ** -20: one + 2.0 ** -7,  # above the tie
        -2.5: -2.5,
        0.0: 0.0,
    }
    for x, want in cases.items():
        assert float(L.bf16(np.float32(x))) == want, x


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = L.init_params(3).astype(np.float64)
    p[L.FEATURES * L.HIDDEN: L.FEATURES * L.HIDDEN + L.HIDDEN] = rng.normal(0, 0.1, L.HIDDEN)  # b1 != 0
    X = rng.normal(0, 1, (37, L.FEATURES))
    y = rng.integers(0, L.CLASSES, 37)
    loss, g = L.mlp_loss_grad(p, X, y, emulate=False)
    idx = rng.choice(L.PARAMS, 300, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(L.PARAMS)
        e[i] = h
        lp, _ = L.mlp_loss_grad(p + e, X, y, emulate=False)
        lm, _ = L.mlp_loss_grad(p - e, X, y, emulate=False)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])
    assert loss > 0


def test_mlp_loss_closed_form():
    # zero weights: uniform softmax, loss = log 4, gradient only on b3
    p = np.zeros(L.PARAMS)
    X = np.ones((8, L.FEATURES))
    y = np.array([0, 1, 2, 3, 0, 0, 1, 3])
    loss, g = L.mlp_loss_grad(p, X, y, emulate=False)
    assert abs(loss - np.log(4)) < 1e-12
    gb3 = g[-L.CLASSES:]
    want = 0.25 - np.bincount(y, minlength=4) / 8
    assert np.allclose(gb3, want)
    assert np.all(g[:-L.CLASSES] == 0)


def test_labels_prefer_the_smallest_sufficient_radius():
    cost = np.array([[1.10, 1.00, 2.0], [1.05, 1.00, 1.5], [1.001, 1.00, 1.2], [1.000, 1.00, 1.0]])
    assert L.RADII == (2, 4, 8, 16)
    from nexusbvh_amd import learned as NL
    assert NL.labels_of(cost, 0.002).tolist() == [2, 0, 3]


def test_segment_features_of_a_hand_built_segment():
    # 4 unit-area-ish leaves in a 4 x 1 x 1 row, keys differing at chosen depths
    boxes = np.array([[i, 0, 0, i + 1, 1, 1] for i in range(4)], dtype=np.float32)
    keys = np.array([0b000, 0b001, 0b1000000, 0b1000000], dtype=np.uint32)  # depths: 9, 7, equal
    perm = np.arange(4, dtype=np.uint32)
    f = L.segment_features(keys, perm, boxes, 8, 32)[0]
    assert abs(f[0] - 2 / 16) < 1e-12                       # log2(4)/16
    assert abs(f[1] - np.log2(1 / 4) / 24) < 1e-12          # middle/largest extent
    assert abs(f[2] - np.log2(1 / 4) / 24) < 1e-12
    seg_a = 4 * 1 + 1 * 1 + 1 * 4
    assert abs(f[3] - np.log2(seg_a / 12) / 16) < 1e-12     # 4 leaves of half-area 3
    assert abs(f[4] - (np.log2(3) - np.log2(seg_a)) / 48) < 1e-8   # leaf log-areas in float32
    assert abs(f[5]) < 1e-8 and f[6] == 0 and abs(f[7]) < 1e-12
    # prefix depth d0 = depth(0 ^ 64) = (29 - 6) // 3 = 7; pairs at depth 9, 7, equal
    assert f[8:].tolist() == [1 / 3, 0, 1 / 3, 0, 0, 0, 0, 1 / 3]


def test_segment_sah_sums_to_the_internal_sah():
    rng = np.random.default_rng(5)
    n = 3000
    c = rng.random((n, 3), dtype=np.float32)
    tris = np.concatenate([c, c + 0.003, c + np.array([0.003, 0, 0.002], np.float32)], axis=1).astype(np.float32)
    r = O.build(tris, 1, 32, 32, 8, with_bvh8=False)
    nodes, scene = r["nodes"], r["scene"]
    boxes = np.concatenate([nodes["bmin"], nodes["bmax"]], axis=1).astype(np.float32)
    seg = L.segment_sah(boxes, n, 8, scene)
    assert seg.shape == (((n - 1) >> 8) + 1,)
    f32 = np.float32

    def area(b):
        d = (b[:, 3:] - b[:, :3]).astype(f32)
        return (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]).astype(f32)

    sa = area(scene.reshape(1, 6))[0]
    leaf = float(np.sum((f32(2.0) * (area(boxes[:n]) / sa)).astype(np.float64)))
    total = O.sah_bvh2(nodes, 2 * n - 1, scene)
    assert abs(seg.sum() + leaf - total) <= 1e-9 * total


# ---- GPU ----------------------------------------------------------------------

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def nx():
    from nexusbvh_amd import native
    if native.device_count() == 0:
        pytest.skip("no GPU")
    from nexusbvh_amd import learned as NL
    return NL


@gpu
def test_mlp_forward_matches_cpu(nx):
    from nexusbvh_amd.native import DeviceBuffer
    rng = np.random.default_rng(11)
    p = L.init_params(7)
    p[1024:1088] = rng.normal(0, 0.1, 64)
    for count in (1, 15, 16, 100, 1000):
        X = rng.normal(0, 1, (count, 16)).astype(np.float32)
        say("mlp forward", count)
        table, logits, fb = nx.mlp_radius_table(DeviceBuffer.from_array(p), DeviceBuffer.from_array(X), count, 0.0,
                                                want_logits=True)
        ref, _ = L.mlp_forward(p, X, True)
        assert np.allclose(logits, ref, rtol=1e-2, atol=2e-3 * np.abs(ref).max()), count
        margin = np.sort(ref, axis=1)
        clear = (margin[:, -1] - margin[:, -2]) > 1e-2
        want, _ = L.radius_table(p, X, 0.0)
        assert np.array_equal(table[clear], want[clear])
        assert fb == 0


@gpu
def test_radius_table_fallback(nx):
    from nexusbvh_amd.native import DeviceBuffer
    rng = np.random.default_rng(12)
    p = L.init_params(8) * 0.05          # nearly uniform softmax: every segment falls back
    X = rng.normal(0, 1, (300, 16)).astype(np.float32)
    table, _, fb = nx.mlp_radius_table(DeviceBuffer.from_array(p), DeviceBuffer.from_array(X), 300, 0.9)
    assert fb == 300 and np.all(table == 8)
    table, _, fb = nx.mlp_radius_table(DeviceBuffer.from_array(p), DeviceBuffer.from_array(X), 300, 0.0)
    assert fb == 0 and set(np.unique(table)) <= {2, 4, 8, 16}


@gpu
def test_mlp_gradients_match_cpu(nx):
    from nexusbvh_amd.native import DeviceBuffer
    rng = np.random.default_rng(13)
    p = L.init_params(9)
    p[1024:1088] = rng.normal(0, 0.1, 64)
    for count in (5, 64, 1000, 5000):
        X = rng.normal(0, 1, (count, 16)).astype(np.float32)
        y = rng.integers(0, 4, count).astype(np.uint8)
        say("mlp gradients", count)
        grads, loss = DeviceBuffer(L.PARAMS * 4), DeviceBuffer(4)
        nx.mlp_gradients(DeviceBuffer.from_array(p), DeviceBuffer.from_array(X), DeviceBuffer.from_array(y), count,
                         grads, loss)
        g = grads.to_array(np.float32, L.PARAMS)
        lg = float(loss.to_array(np.float32)[0])
        ref_loss, ref = L.mlp_loss_grad(p, X, y.astype(np.int64), True)
        assert abs(lg - ref_loss) <= 1e-3 * ref_loss
        scale = np.abs(ref).max()
        assert np.linalg.norm(g - ref) <= 2e-2 * np.linalg.norm(ref), count
        assert np.allclose(g, ref, rtol=5e-2, atol=2e-2 * scale), count


@gpu
def test_adam_step_matches_closed_form(nx):
    from nexusbvh_amd.native import DeviceBuffer, check, lib, synchronize
    rng = np.random.default_rng(14)
    p0 = rng.normal(0, 1, L.PARAMS).astype(np.float32)
    g = rng.normal(0, 1, L.PARAMS).astype(np.float32)
    m = np.zeros(L.PARAMS, np.float32)
    v = np.zeros(L.PARAMS, np.float32)
    P, G, M, V = (DeviceBuffer.from_array(a) for a in (p0, g, m, v))
    check(lib().nxb_mlp_adam_step(P.address, G.address, M.address, V.address, 1, ctypes.c_float(0.01),
                                  ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8), None))
    synchronize()
    # first step: m_hat = g, v_hat = g^2, update = lr * g / (|g| + eps)
    want = p0 - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(P.to_array(np.float32), want, rtol=1e-5, atol=1e-6)


@gpu
def test_training_learns_a_separable_problem(nx):
    rng = np.random.default_rng(15)
    A = rng.normal(0, 1, (16, 4))
    X = rng.normal(0, 1, (20000, 16)).astype(np.float32)
    y = np.argmax(X @ A, axis=1).astype(np.uint8)
    say("train separable")
    res = nx.train(X, y, steps=600, batch=4096, lr=3e-3, class_balance=False, log_every=200)
    acc = nx.accuracy(res.params, X, y)
    say("accuracy", acc, res.losses)
    assert res.losses[-1][1] < res.losses[0][1] * 0.3
    assert acc > 0.9


def _sorted_scene(tris, bits=64):
    from nexusbvh_amd import stages
    from nexusbvh_amd.builder import DevicePrimitives
    n = len(tris)
    prims = DevicePrimitives.from_numpy(tris)
    nodes, scene, _ = stages.scene_bounds(prims)
    keys, _ = stages.morton(nodes, n, scene, bits)
    b, e = O.sort_window(bits)
    ko, vo = stages.sort_pairs(keys, n, bits, b, e, 8)
    return prims, nodes, scene, ko, vo


@gpu
@pytest.mark.parametrize("bits", [32, 64])
def test_segment_features_match_cpu(nx, bits):
    from nexusbvh_amd import scenes
    tris = scenes.blobs_plane(70_001, 3, blobs=3, side=100, cell=0.2)
    n = len(tris)
    say("segment features", bits, n)
    _, nodes, scene, ko, vo = _sorted_scene(tris, bits)
    for seg_log2 in (8, 12):
        f = nx.segment_features(ko, vo, nodes, n, bits, seg_log2).to_array(np.float32).reshape(-1, 16)
        segs = nx.segment_count(n, seg_log2)
        f = f[:segs]
        dt = np.uint32 if bits == 32 else np.uint64
        keys = ko.to_array(dt, n)
        perm = vo.to_array(np.uint32, n)
        leaves = nodes.to_array(np.float32, n * 8).reshape(n, 8)[:, :6]
        ref = L.segment_features(keys, perm, leaves, seg_log2, bits)
        assert np.allclose(f[:, :8], ref[:, :8], rtol=1e-4, atol=1e-4), seg_log2
        assert np.allclose(f[:, 8:], ref[:, 8:], atol=1e-6), seg_log2


@gpu
def test_segment_sah_matches_cpu(nx):
    from nexusbvh_amd import scenes, stages
    from nexusbvh_amd.builder import BuildConfig
    from nexusbvh_amd.native import DeviceBuffer, copy_device
    tris = scenes.blobs_plane(50_000, 4, blobs=2, side=90, cell=0.2)
    n = len(tris)
    _, nodes, scene, ko, vo = _sorted_scene(tris)
    work = DeviceBuffer(n * 4)
    copy_device(work, vo, n * 4)
    stages.hploc(ko, work, n, 64, nodes, BuildConfig(mortonBits=64))
    for seg_log2 in (8, 10, 12):
        say("segment sah", seg_log2)
        got = nx.segment_sah(nodes, n, scene, seg_log2)
        nd = stages.nodes_to_host(nodes, 2 * n - 1)
        boxes = np.concatenate([nd["bmin"], nd["bmax"]], axis=1).astype(np.float32)
        ref = L.segment_sah(boxes, n, seg_log2, scene.to_array(np.float32, 6))
        assert np.allclose(got, ref, rtol=1e-9, atol=0), seg_log2


@gpu
def test_learned_policy_builds_a_valid_bvh(nx):
    import nxb_checks
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import scenes
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, RadiusPolicy
    from nexusbvh_amd.native import DeviceBuffer, lib
    tris = scenes.blobs_plane(120_000, 5, blobs=4, side=130, cell=0.15)
    n = len(tris)
    d = DevicePrimitives.from_numpy(tris)
    p = DeviceBuffer.from_array(L.init_params(21))
    for bits in (32, 64):
        cfg = BuildConfig(mortonBits=bits, radiusPolicy=RadiusPolicy.LEARNED, mlpParams=p, segmentLog2=10,
                          mlpMinConfidence=0.0)
        say("learned build", bits)
        dev2 = B.BuildBVH2(d, n, cfg)
        h2 = B.ToHost(dev2)
        B.FreeDeviceBVH(dev2)
        nxb_checks.bvh2_invariants(h2.nodes, n, scenes.boxes_of(tris), np.array(h2.bounds, dtype=np.float32).reshape(6))
        b8 = B.BuildBVH8(d, n, cfg)
        nodes8, prim = B.bvh8_to_host(b8)
        nxb_checks.bvh8_invariants(nodes8, prim, n)
        B.FreeDeviceBVH(b8)
        segs, fb = ctypes.c_uint32(0), ctypes.c_uint32(0)
        lib().nxb_last_learned_stats(ctypes.byref(segs), ctypes.byref(fb))
        assert segs.value == nx.segment_count(n, 10) and fb.value == 0


@gpu
def test_learned_policy_with_a_constant_table_equals_fixed_radius(nx):
    """Parameters whose output is always class 2 (radius 8) reproduce the
    fixed-radius-8 BVH2 byte for byte; always class 3 reproduces radius 16
    (merge tasks wider than a segment take search_radius, set to the same)."""
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import scenes
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, RadiusPolicy
    from nexusbvh_amd.native import DeviceBuffer
    tris = scenes.blobs_plane(60_000, 6, blobs=2, side=100, cell=0.2)
    n = len(tris)
    d = DevicePrimitives.from_numpy(tris)
    for cls, r in ((2, 8), (3, 16)):
        p = np.zeros(L.PARAMS, np.float32)
        p[-4 + cls] = 10.0                       # b3: constant logits
        cfg = BuildConfig(mortonBits=64, searchRadius=r, radiusPolicy=RadiusPolicy.LEARNED,
                          mlpParams=DeviceBuffer.from_array(p), mlpMinConfidence=0.5)
        got = []
        for c in (cfg, BuildConfig(mortonBits=64, searchRadius=r)):
            dev2 = B.BuildBVH2(d, n, c)
            got.append(B.ToHost(dev2).nodes)
            B.FreeDeviceBVH(dev2)
        assert got[0].tobytes() == got[1].tobytes(), r


def test_shipped_weights_file():
    from nexusbvh_amd import learned as NL
    p = NL.load()
    assert p.shape == (L.PARAMS,) and np.all(np.isfinite(p))
    import json
    meta = json.load(open(NL.META_PATH))
    assert meta["segment_log2"] == 12 and meta["scenes"] == 2000


@gpu
def test_trained_network_gains_over_fixed_radius(nx):
    """The shipped network against the method without it (fixed radius 8) on
    held-out scenes of the training family: lower BVH2 and BVH8 SAH."""
    params = nx.load()
    for idx in (100002, 100003):
        tris = nx.family_scene(idx)
        r = nx.compare_policies(tris, params, 0.0, top_radius=8, reps=1, names=["r8", "learned"])
        say("gain", idx, len(tris), r)
        assert r["learned"]["bvh2_sah"] < r["r8"]["bvh2_sah"]
        assert r["learned"]["bvh8_sah"] < r["r8"]["bvh8_sah"]


@gpu
def test_shipped_network_at_full_size(nx):
    """The 28.1M-triangle benchmark scene built with the shipped network at
    the library's default confidence threshold (0.5, so some segments fall
    back to radius 8): the BVH2 and BVH8 pass the structural invariants, the
    fallback is active on some but not all segments, and the BVH2 SAH is
    below that of fixed radius 8."""
    import nxb_checks
    from nexusbvh_amd import builder as B
    from nexusbvh_amd import scenes
    from nexusbvh_amd.builder import BuildConfig, DevicePrimitives, RadiusPolicy
    from nexusbvh_amd.native import DeviceBuffer, lib
    tris = scenes.workload("blobs_28m1")
    n = len(tris)
    d = DevicePrimitives.from_numpy(tris)
    p = DeviceBuffer.from_array(nx.load().astype(np.float32))
    cfg = BuildConfig(mortonBits=64, searchRadius=8, radiusPolicy=RadiusPolicy.LEARNED, mlpParams=p,
                      segmentLog2=12)
    assert cfg.mlpMinConfidence == 0.5
    say("learned full size: BVH2")
    m2 = B.BVHBuildMetrics()
    dev2 = B.BuildBVH2(d, n, cfg, m2)
    segs, fb = ctypes.c_uint32(0), ctypes.c_uint32(0)
    lib().nxb_last_learned_stats(ctypes.byref(segs), ctypes.byref(fb))
    assert segs.value == nx.segment_count(n, 12)
    assert 0 < fb.value < segs.value, (fb.value, segs.value)
    h2 = B.ToHost(dev2)
    B.FreeDeviceBVH(dev2)
    nxb_checks.bvh2_invariants(h2.nodes, n, scenes.boxes_of(tris), np.array(h2.bounds, dtype=np.float32).reshape(6))
    del h2
    say("learned full size: BVH8")
    m8 = B.BVHBuildMetrics()
    b8 = B.BuildBVH8(d, n, cfg, m8)
    nodes8, prim = B.bvh8_to_host(b8)
    B.FreeDeviceBVH(b8)
    nxb_checks.bvh8_invariants(nodes8, prim, n)
    r8 = B.BVHBuildMetrics()
    B.FreeDeviceBVH(B.BuildBVH8(d, n, BuildConfig(mortonBits=64), r8))
    say("learned full size SAH", m8.bvh2Cost, m8.bvh8Cost, "r8", r8.bvh2Cost, r8.bvh8Cost, "fallbacks", fb.value)
    assert m2.bvh2Cost == pytest.approx(m8.bvh2Cost, rel=1e-6)
    assert m8.bvh2Cost < r8.bvh2Cost
