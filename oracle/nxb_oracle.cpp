/*
 * nxb_oracle.cpp - CPU oracle of the NexusBVH construction pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This file is a sequential, single-threaded
 * restatement of the reference method. Only the test suite, the smoke check
 * and the CPU-baseline leg of the benchmark load it, always as the checker
 * (or as the timed CPU baseline), never as part of the GPU product path.
 *
 * Written from the reference sources (read as text; paths relative to the
 * NexusBVH tree):
 *   AABB/Triangle          NexusBVH/include/NXB/AABB.h:8-54, Triangle.h:10-48
 *   scene bounds           NexusBVH/src/Cuda/Setup.cu:12-39
 *   Morton codes           NexusBVH/src/Cuda/Setup.cu:42-59, BuilderUtils.h:108-253
 *   radix sort semantics   NexusBVH/src/Cuda/Setup.cu:62-112 (CUB stable LSD)
 *   H-PLOC                 NexusBVH/src/Cuda/BinaryBuilder.cu:16-273
 *   BVH8 collapse          NexusBVH/src/Cuda/WideConverter.cu:25-30,106-222,225-414
 *   auction slots          NexusBVH/src/Cuda/WideConverter.cu:52-102,336-361 (compiled out there)
 *   SAH                    NexusBVH/src/Cuda/Eval.cu:12-80
 *   conventions            NexusBVH/src/BVHBuilder.cpp:13-267
 * and, with no reference counterpart, a brute-force closest-hit ray query
 * (the check of the BVH8 traversal kernel).
 *
 * Arithmetic conventions (shared with the HIP kernels, built the same way):
 *   - no FMA contraction (compiled with -ffp-contract=off), IEEE division;
 *   - box min/max are IEEE 754-2019 minimum/maximum (-0 < +0), so a union of
 *     boxes is symmetric and order-independent, including signed zeros;
 *   - float -> integer quantization clamps NaN and negatives to 0.
 * These fix what the reference leaves to nvcc (fast-math division, fminf on
 * signed zeros, contraction inside AABB::Area).
 *
 * Parity pins (tests/test_oracle.py): Morton known answers and a bit-loop
 * interleave; sort = numpy stable argsort; bounds = numpy min/max; the
 * nearest-neighbour shuffle simulation = brute-force argmin with the
 * i+1,i-1,i+2,... tie order; N<=16 H-PLOC = textbook PLOC written in the
 * test; BVH2/BVH8 structural invariants; SAH closed forms; exhaustive check
 * that BVH8 quantization never exceeds 255; the auction's epsilon-
 * complementary slackness at its final prices, its 8-child epsilon bound
 * against brute-force optimal assignments, and an octant closed form.
 */
#include "../include/nxb.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <utility>
#include <vector>

namespace {

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float ffrom(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

/* IEEE 754-2019 minimum / maximum: NaN propagates, -0 < +0. */
inline float fmin19(float a, float b) {
    if (a != a || b != b) return a + b;
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
}
inline float fmax19(float a, float b) {
    if (a != a || b != b) return a + b;
    if (a > b) return a;
    if (b > a) return b;
    return std::signbit(a) ? b : a;
}

inline nxb_aabb box_union(const nxb_aabb& a, const nxb_aabb& b) {
    nxb_aabb r;
    r.bmin.x = fmin19(a.bmin.x, b.bmin.x);
    r.bmin.y = fmin19(a.bmin.y, b.bmin.y);
    r.bmin.z = fmin19(a.bmin.z, b.bmin.z);
    r.bmax.x = fmax19(a.bmax.x, b.bmax.x);
    r.bmax.y = fmax19(a.bmax.y, b.bmax.y);
    r.bmax.z = fmax19(a.bmax.z, b.bmax.z);
    return r;
}

/* AABB::Area (AABB.h:46-50): half the surface area, evaluated left to right. */
inline float box_area(const nxb_aabb& b) {
    float dx = b.bmax.x - b.bmin.x;
    float dy = b.bmax.y - b.bmin.y;
    float dz = b.bmax.z - b.bmin.z;
    return dx * dy + dy * dz + dz * dx;
}

inline bool finite3(const nxb_float3& v) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z);
}

/* AABB(v0, v1, v2) (AABB.h:15-19) via GetBounds (BuilderUtils.h:255-263). */
inline nxb_aabb prim_bounds(const void* prims, int type, uint32_t i, bool* bad) {
    nxb_aabb b;
    if (type == NXB_PRIM_TRIANGLE) {
        const nxb_triangle& t = static_cast<const nxb_triangle*>(prims)[i];
        b.bmin.x = fmin19(t.v0.x, fmin19(t.v1.x, t.v2.x));
        b.bmin.y = fmin19(t.v0.y, fmin19(t.v1.y, t.v2.y));
        b.bmin.z = fmin19(t.v0.z, fmin19(t.v1.z, t.v2.z));
        b.bmax.x = fmax19(t.v0.x, fmax19(t.v1.x, t.v2.x));
        b.bmax.y = fmax19(t.v0.y, fmax19(t.v1.y, t.v2.y));
        b.bmax.z = fmax19(t.v0.z, fmax19(t.v1.z, t.v2.z));
        if (!finite3(t.v0) || !finite3(t.v1) || !finite3(t.v2)) *bad = true;
    } else {
        b = static_cast<const nxb_aabb*>(prims)[i];
        if (!finite3(b.bmin) || !finite3(b.bmax)) *bad = true;
    }
    return b;
}

/* Quantize a normalized coordinate: (uint32)(n * scale), NaN/negative -> 0,
 * clamped to scale (BuilderUtils.h:235-253 truncating cast). */
inline uint32_t quantize(float n, float scale) {
    float q = n * scale;
    if (!(q > 0.0f)) return 0u;
    if (q > scale) q = scale;
    return (uint32_t)q;
}

/* Bit interleave written as a plain loop (x -> bit 0, y -> bit 1, z -> bit 2),
 * independent of the magic-mask formulation (BuilderUtils.h:108-226). */
inline uint64_t interleave3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint64_t code = 0;
    for (int b = 0; b < bits; ++b) {
        code |= (uint64_t)((x >> b) & 1u) << (3 * b);
        code |= (uint64_t)((y >> b) & 1u) << (3 * b + 1);
        code |= (uint64_t)((z >> b) & 1u) << (3 * b + 2);
    }
    return code;
}

template <typename K>
inline uint64_t morton_of(const nxb_aabb& b, const nxb_aabb& s);

/* Setup.cu:50-54 with BuilderUtils.h:234-241: 10 bits per axis. */
template <>
inline uint64_t morton_of<uint32_t>(const nxb_aabb& b, const nxb_aabb& s) {
    float cx = (b.bmin.x + b.bmax.x) * 0.5f;
    float cy = (b.bmin.y + b.bmax.y) * 0.5f;
    float cz = (b.bmin.z + b.bmax.z) * 0.5f;
    float nx = (cx - s.bmin.x) / (s.bmax.x - s.bmin.x);
    float ny = (cy - s.bmin.y) / (s.bmax.y - s.bmin.y);
    float nz = (cz - s.bmin.z) / (s.bmax.z - s.bmin.z);
    return interleave3(quantize(nx, 1023.0f), quantize(ny, 1023.0f), quantize(nz, 1023.0f), 10);
}
/* BuilderUtils.h:246-253: 21 bits per axis. */
template <>
inline uint64_t morton_of<uint64_t>(const nxb_aabb& b, const nxb_aabb& s) {
    float cx = (b.bmin.x + b.bmax.x) * 0.5f;
    float cy = (b.bmin.y + b.bmax.y) * 0.5f;
    float cz = (b.bmin.z + b.bmax.z) * 0.5f;
    float nx = (cx - s.bmin.x) / (s.bmax.x - s.bmin.x);
    float ny = (cy - s.bmin.y) / (s.bmax.y - s.bmin.y);
    float nz = (cz - s.bmin.z) / (s.bmax.z - s.bmin.z);
    return interleave3(quantize(nx, 2097151.0f), quantize(ny, 2097151.0f),
                       quantize(nz, 2097151.0f), 21);
}

/* ---- H-PLOC --------------------------------------------------------------- */

/* Delta (BinaryBuilder.cu:16-28). */
inline uint64_t delta_k(const uint32_t* m, uint32_t a, uint32_t b) {
    return (((uint64_t)m[a] << 32) | a) ^ (((uint64_t)m[b] << 32) | b);
}
inline uint64_t delta_k(const uint64_t* m, uint32_t a, uint32_t b) {
    uint64_t d = m[a] ^ m[b];
    return d ? d : (uint64_t)(a ^ b);
}

/* Node numbering. The reference takes ids from an atomic counter
 * (BinaryBuilder.cu:85-91), so its numbering depends on scheduling. Here a
 * cluster carries a representative, the smallest sorted position it covers;
 * lists stay ordered by representative, and merging clusters i < j retires
 * rep(j). Every position but 0 is retired exactly once, so
 * id = n - 1 + rep(j) numbers the n - 1 internal nodes bijectively onto
 * [n, 2n-2]. The root is placed at 2n-2 (BVHBuilder.cpp:195 relies on it);
 * the node that retired position n-1 (natural id 2n-2) then takes the
 * root's natural id. */
struct Hploc {
    nxb_bvh2_node* nodes;
    std::vector<uint32_t> cluster;   /* clusterIdx */
    std::vector<uint32_t> rep;       /* representative of each stored cluster */
    std::vector<uint32_t> parent;    /* parentIdx */
    uint32_t n;
    uint32_t next_id;                /* internal nodes created so far */
    uint32_t moved_parent;           /* parent of the node that retired n-1 */
    nxb_bvh2_node moved;             /* that node, until the root is known */
    uint32_t width;                  /* lanes in a merge group: 32 or 64 */
    uint32_t cap;                    /* per-child load cap = merge threshold */
    uint32_t radius;
    const uint8_t* table;
    uint32_t seg_log2, table_len;

    /* Per-segment policy: a merge task over sorted positions [l_start, r_end)
     * inside one segment's span takes the radius of the segment of l_start;
     * wider tasks (the upper levels) and table entries 0 take `radius`. */
    uint32_t radius_for(uint32_t l_start, uint32_t r_end) const {
        uint32_t r = radius;
        if (table && r_end - l_start <= (1u << seg_log2)) {
            uint32_t s = l_start >> seg_log2;
            if (s < table_len && table[s] >= 1) r = table[s];
        }
        return r > 16 ? 16 : r;
    }
};

/* FindNearestNeighbor (BinaryBuilder.cu:127-163) simulated lane by lane,
 * including the modulo-width wrap of the shuffles. */
void find_nn_shuffle(uint32_t num, uint32_t W, uint32_t R, const nxb_aabb* box, uint32_t* nn_out) {
    uint32_t ma[64], mi[64], area[64], ta[64], ti[64];
    for (uint32_t l = 0; l < W; ++l) { ma[l] = NXB_INVALID_IDX; mi[l] = NXB_INVALID_IDX; }
    for (uint32_t r = 1; r <= R; ++r) {
        for (uint32_t l = 0; l < W; ++l) {
            uint32_t nb = l + r;
            area[l] = NXB_INVALID_IDX;
            if (nb < num) {
                area[l] = fbits(box_area(box_union(box[nb % W], box[l])));
                if (area[l] < ma[l]) { ma[l] = area[l]; mi[l] = nb; }
            }
        }
        for (uint32_t l = 0; l < W; ++l) {
            uint32_t src = (l + r) % W;
            uint32_t a = ma[src], i = mi[src];
            if (area[l] < a) { a = area[l]; i = l; }
            ta[l] = a; ti[l] = i;
        }
        for (uint32_t l = 0; l < W; ++l) {
            uint32_t src = (uint32_t)(l - r) % W;
            ma[l] = ta[src]; mi[l] = ti[src];
        }
    }
    for (uint32_t l = 0; l < W; ++l) nn_out[l] = mi[l];
}

constexpr uint32_t kMovedId = 0xFFFFFFFDu;  /* temporary id of the node that retired n-1 */

void put_node(Hploc& h, uint32_t id, const nxb_bvh2_node& nd) {
    if (id == kMovedId) h.moved = nd;
    else h.nodes[id] = nd;
}

/* MergeClustersCreateBVH2Node (BinaryBuilder.cu:72-124) simulated lane by lane. */
uint32_t merge_step(Hploc& h, uint32_t num, const uint32_t* nn, uint32_t* cid, uint32_t* rep, nxb_aabb* box,
                    bool final) {
    const uint32_t W = h.width;
    bool merge[64], mutual[64];
    uint32_t merge_count = 0;
    for (uint32_t l = 0; l < W; ++l) {
        bool active = l < num;
        uint32_t nnnn = nn[nn[l] % W];
        mutual[l] = active && l == nnnn;
        merge[l] = mutual[l] && l < nn[l];
        merge_count += merge[l];
    }
    h.next_id += merge_count;
    const bool root_step = final && num - merge_count == 1;
    const uint32_t n = h.n;
    uint32_t ncid[64], nrep[64];
    nxb_aabb nbox[64];
    for (uint32_t l = 0; l < W; ++l) {
        ncid[l] = cid[l];
        nrep[l] = rep[l];
        nbox[l] = box[l];
        if (merge[l]) {
            uint32_t j = nn[l] % W;
            nxb_bvh2_node nd;
            nd.bounds = box_union(box[l], box[j]);
            nd.left = cid[l];
            nd.right = cid[j];
            uint32_t natural = n - 1 + rep[j];
            uint32_t id = root_step ? 2 * n - 2 : (rep[j] == n - 1 ? kMovedId : natural);
            if (root_step && natural != 2 * n - 2) {
                /* the moved node takes the root's natural id */
                h.nodes[natural] = h.moved;
                if (nd.left == kMovedId) nd.left = natural;
                else if (nd.right == kMovedId) nd.right = natural;
                else {
                    nxb_bvh2_node& p = h.nodes[h.moved_parent];
                    if (p.left == kMovedId) p.left = natural;
                    else p.right = natural;
                }
            } else if (nd.left == kMovedId || nd.right == kMovedId) {
                h.moved_parent = id;
            }
            put_node(h, id, nd);
            ncid[l] = id;
            nbox[l] = nd.bounds;
        }
    }
    /* order-preserving compaction of lanes with (merge || !mutual) */
    uint32_t k = 0;
    for (uint32_t l = 0; l < W; ++l) {
        if (merge[l] || !mutual[l]) { cid[k] = ncid[l]; rep[k] = nrep[l]; box[k] = nbox[l]; ++k; }
    }
    for (; k < W; ++k) cid[k] = NXB_INVALID_IDX;
    return num - merge_count;
}

/* PlocMerge with LoadIndices/StoreIndices (BinaryBuilder.cu:42-69, 165-197). */
void ploc_merge(Hploc& h, uint32_t left, uint32_t right, uint32_t split, bool final) {
    const uint32_t W = h.width, C = h.cap;
    uint32_t l_start = left, l_end = split, r_start = split, r_end = right + 1;
    uint32_t cid[64], rep[64];
    nxb_aabb box[64];
    for (uint32_t l = 0; l < W; ++l) { cid[l] = NXB_INVALID_IDX; rep[l] = 0; }
    uint32_t num_left = 0, num_right = 0;
    /* a child of at most C primitives was never merged: its list is its
     * leaves in sorted order, representative = position */
    const bool left_merged = l_end - l_start > C, right_merged = r_end - r_start > C;
    for (uint32_t l = 0; l < W; ++l) {
        uint32_t idx = l;
        if (idx < std::min(l_end - l_start, C)) {
            cid[l] = h.cluster[l_start + idx];
            rep[l] = left_merged ? h.rep[l_start + idx] : l_start + idx;
            if (cid[l] != NXB_INVALID_IDX) ++num_left;
        }
    }
    for (uint32_t l = 0; l < W; ++l) {
        uint32_t idx = l - num_left;
        if (idx < std::min(r_end - r_start, C)) {
            cid[l] = h.cluster[r_start + idx];
            rep[l] = right_merged ? h.rep[r_start + idx] : r_start + idx;
            if (cid[l] != NXB_INVALID_IDX) ++num_right;
        }
    }
    uint32_t num = num_left + num_right;
    for (uint32_t l = 0; l < num; ++l) box[l] = cid[l] == kMovedId ? h.moved.bounds : h.nodes[cid[l]].bounds;
    for (uint32_t l = num; l < W; ++l) box[l] = nxb_aabb{{0, 0, 0}, {0, 0, 0}};
    uint32_t threshold = final ? 1 : C;
    uint32_t R = h.radius_for(l_start, r_end);
    uint32_t nn[64];
    while (num > threshold) {
        find_nn_shuffle(num, W, R, box, nn);
        num = merge_step(h, num, nn, cid, rep, box, final);
    }
    for (uint32_t l = 0; l < num_left + num_right; ++l) {
        h.cluster[l_start + l] = cid[l];
        h.rep[l_start + l] = rep[l];
    }
}

/* BuildBVH2Kernel (BinaryBuilder.cu:200-273) run one "thread" at a time: each
 * thread climbs until it is the first to reach a parent. */
template <typename K>
void hploc_run(Hploc& h, const K* keys) {
    const uint32_t n = h.n;
    const uint32_t C = h.cap;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t left = i, right = i, split = 0;
        for (;;) {
            uint32_t prev;
            bool parent_right = (left == 0) ||
                (right != n - 1 && delta_k(keys, right, right + 1) < delta_k(keys, left - 1, left));
            if (parent_right) {
                prev = h.parent[right];
                h.parent[right] = left;
                if (prev != NXB_INVALID_IDX) { split = right + 1; right = prev; }
            } else {
                prev = h.parent[left - 1];
                h.parent[left - 1] = right;
                if (prev != NXB_INVALID_IDX) { split = left; left = prev; }
            }
            if (prev == NXB_INVALID_IDX) break;
            uint32_t size = right - left + 1;
            bool final = size == n;
            if (size > C || final) ploc_merge(h, left, right, split, final);
        }
    }
}

/* ---- BVH8 ----------------------------------------------------------------- */

inline uint32_t nib(uint32_t x, uint32_t i) { return (x >> (4 * i)) & 0xFu; }
inline void set_nib(uint32_t& x, uint32_t i, uint32_t v) {
    x &= ~(0xFu << (4 * i));
    x |= v << (4 * i);
}
inline uint32_t bits_below(uint32_t x, uint32_t i) {
    return (uint32_t)__builtin_popcount(x & ((1u << i) - 1u));
}
/* CeilLog2 / InvPow2 (BuilderUtils.h:335-347). */
inline uint32_t ceil_log2(float x) {
    uint32_t ix = fbits(x);
    uint32_t e = (ix >> 23) & 0xFFu;
    return e + ((ix & 0x7FFFFFu) != 0);
}
inline float inv_pow2(uint8_t e) { return ffrom((uint32_t)(254 - e) << 23); }

/* GetCost (WideConverter.cu:25-30). */
inline float slot_cost(uint32_t s, const nxb_float3& o) {
    float sx = ((s >> 2) & 1) ? -1.0f : 1.0f;
    float sy = ((s >> 1) & 1) ? -1.0f : 1.0f;
    float sz = (s & 1) ? -1.0f : 1.0f;
    return sx * o.x + sy * o.y + sz * o.z;
}

/* CreateBVH8Node (WideConverter.cu:156-206). Empty slots have all-zero
 * meta/quantization bytes (the reference leaves those bytes unset). */
nxb_bvh8_node make_bvh8_node(const nxb_bvh2_node* b2, const nxb_aabb& bounds,
                             const uint32_t* child, uint32_t child_base, uint32_t prim_base,
                             uint32_t assign, uint32_t inner_mask, uint32_t leaf_mask) {
    nxb_bvh8_node n;
    std::memset(&n, 0, sizeof(n));
    const float qstep = 1.0f / 255.0f;
    float dx = bounds.bmax.x - bounds.bmin.x;
    float dy = bounds.bmax.y - bounds.bmin.y;
    float dz = bounds.bmax.z - bounds.bmin.z;
    n.p = bounds.bmin;
    n.e[0] = (uint8_t)ceil_log2(dx * qstep);
    n.e[1] = (uint8_t)ceil_log2(dy * qstep);
    n.e[2] = (uint8_t)ceil_log2(dz * qstep);
    n.imask = 0;
    n.child_base = child_base;
    n.prim_base = prim_base;
    float ix = inv_pow2(n.e[0]), iy = inv_pow2(n.e[1]), iz = inv_pow2(n.e[2]);
    for (uint32_t i = 0; i < 8; ++i) {
        uint32_t a = nib(assign, i);
        if (inner_mask & (1u << i)) {
            n.imask |= (uint8_t)(1u << i);
            n.meta[i] = (uint8_t)((1u << 5) | (24 + i));
        } else if (a != 0xFu) {
            n.meta[i] = (uint8_t)((1u << 5) | bits_below(leaf_mask, i));
        } else {
            continue;
        }
        const nxb_aabb& c = b2[child[a]].bounds;
        n.qlox[i] = (uint8_t)(uint32_t)std::floor((c.bmin.x - bounds.bmin.x) * ix);
        n.qloy[i] = (uint8_t)(uint32_t)std::floor((c.bmin.y - bounds.bmin.y) * iy);
        n.qloz[i] = (uint8_t)(uint32_t)std::floor((c.bmin.z - bounds.bmin.z) * iz);
        n.qhix[i] = (uint8_t)(uint32_t)std::ceil((c.bmax.x - bounds.bmin.x) * ix);
        n.qhiy[i] = (uint8_t)(uint32_t)std::ceil((c.bmax.y - bounds.bmin.y) * iy);
        n.qhiz[i] = (uint8_t)(uint32_t)std::ceil((c.bmax.z - bounds.bmin.z) * iz);
    }
    return n;
}

/* AuctionAssignment (WideConverter.cu:52-102, compiled out in the reference
 * by its USE_AUCTION switch; cost table at WideConverter.cu:336-361):
 * Bertsekas' auction with epsilon scaling (theta = 8). offsets: parent
 * centroid - child centroid (both as min + max), unscaled. Bounded here: at
 * most 16 epsilon phases of 4096 bids; children left without a slot (bid
 * limit, non-finite costs) take the first free slots in child order. Returns
 * the slot nibbles (child index per slot, 0xF = empty); prices_out (8,
 * nullable) receives the final slot prices. */
constexpr uint32_t kAuctionMaxBids = 4096;
constexpr uint32_t kAuctionMaxPhases = 16;

uint32_t auction_assign(const nxb_aabb& pb, const nxb_float3* offsets, uint32_t n, float* prices_out = nullptr) {
    float max_dim = std::fmax(std::fmax(pb.bmax.x - pb.bmin.x, pb.bmax.y - pb.bmin.y), pb.bmax.z - pb.bmin.z);
    float scale = (max_dim == 0.0f ? 0.0f : 10.0f / max_dim) * 0.5f;
    nxb_float3 off[8];
    float max_cost = -3.402823466e+38f;
    for (uint32_t c = 0; c < n; ++c) {
        off[c] = {offsets[c].x * scale, offsets[c].y * scale, offsets[c].z * scale};
        float cost = std::fabs(off[c].x) + std::fabs(off[c].y) + std::fabs(off[c].z);
        if (cost > max_cost) max_cost = cost;
    }
    float prices[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float threshold = 1.0f / (float)n;
    float epsilon = std::fmax(max_cost, threshold);
    uint32_t assign = 0xFFFFFFFFu;
    bool stop = !(epsilon < INFINITY);
    for (uint32_t phase = 0; !stop && epsilon >= threshold && phase < kAuctionMaxPhases; ++phase) {
        assign = 0xFFFFFFFFu;
        uint32_t bidders = 0x76543210u, left = n, bids = 0;
        while (left > 0) {
            if (bids++ == kAuctionMaxBids) { stop = true; break; }
            uint32_t c = nib(bidders, --left);
            float win = -3.402823466e+38f, second = -3.402823466e+38f;
            uint32_t ws = 0xFu;
            for (uint32_t s = 0; s < 8; ++s) {
                float reward = slot_cost(s, off[c]) - prices[s];
                if (reward > win) { second = win; win = reward; ws = s; }
                else if (reward > second) second = reward;
            }
            if (ws == 0xFu) { stop = true; break; }
            prices[ws] += (win - second) + epsilon;
            uint32_t prev = nib(assign, ws);
            set_nib(assign, ws, c);
            if (prev != 0xFu) set_nib(bidders, left++, prev);
        }
        epsilon *= 0.125f;
    }
    for (uint32_t c = 0; c < n; ++c) {
        bool placed = false;
        for (uint32_t s = 0; s < 8; ++s) placed |= nib(assign, s) == c;
        if (placed) continue;
        for (uint32_t s = 0; s < 8; ++s)
            if (nib(assign, s) == 0xFu) { set_nib(assign, s, c); break; }
    }
    if (prices_out) std::memcpy(prices_out, prices, sizeof(prices));
    return assign;
}

/* Collapse + slot assignment (greedy or auction) of one inner BVH2 node
 * (WideConverter.cu:284-379). */
struct Collapsed {
    uint32_t child[8];
    uint32_t count;
    uint32_t assign;      /* nibble s = child index in slot s, 0xF = empty */
    uint32_t inner_mask;  /* slot space */
    uint32_t leaf_mask;   /* slot space */
};

Collapsed collapse_node(const nxb_bvh2_node* b2, uint32_t x, uint32_t slot_mode) {
    Collapsed c;
    uint32_t inner = 0, count = 0;
    int msb = 0;
    uint32_t lr[2] = {b2[x].left, b2[x].right};
    for (;;) {
        float a0 = box_area(b2[lr[0]].bounds);
        float a1 = box_area(b2[lr[1]].bounds);
        int first = a0 < a1 ? 0 : 1;
        for (int i = 0; i < 2; ++i) {
            uint32_t idx = i == 0 ? (uint32_t)msb : count;
            if (b2[lr[first]].left != NXB_INVALID_IDX) inner |= 1u << idx;
            c.child[idx] = lr[first];
            ++count;
            first = !first;
        }
        msb = inner ? 31 - __builtin_clz(inner) : -1;
        if (msb < 0 || count == 8) break;
        inner &= ~(1u << msb);
        --count;
        uint32_t nx = c.child[msb];
        lr[0] = b2[nx].left;
        lr[1] = b2[nx].right;
    }
    c.count = count;

    /* GreedyAssignment (WideConverter.cu:106-153). If every cost compares
     * false (non-finite geometry) the first free slot is taken. */
    const nxb_aabb& pb = b2[x].bounds;
    nxb_float3 pc = {pb.bmin.x + pb.bmax.x, pb.bmin.y + pb.bmax.y, pb.bmin.z + pb.bmax.z};
    uint32_t assign = 0xFFFFFFFFu;
    if (slot_mode == NXB_SLOT_AUCTION) {
        nxb_float3 offs[8];
        for (uint32_t k = 0; k < count; ++k) {
            const nxb_aabb& cb = b2[c.child[k]].bounds;
            offs[k] = {pc.x - (cb.bmin.x + cb.bmax.x), pc.y - (cb.bmin.y + cb.bmax.y),
                       pc.z - (cb.bmin.z + cb.bmax.z)};
        }
        assign = auction_assign(pb, offs, count);
    }
    for (uint32_t k = 0; slot_mode != NXB_SLOT_AUCTION && k < count; ++k) {
        const nxb_aabb& cb = b2[c.child[k]].bounds;
        nxb_float3 cc = {cb.bmax.x + cb.bmin.x, cb.bmax.y + cb.bmin.y, cb.bmax.z + cb.bmin.z};
        nxb_float3 off = {pc.x - cc.x, pc.y - cc.y, pc.z - cc.z};
        float best_cost = -3.402823466e+38f;
        uint32_t best = 0xFu, first_free = 0xFu;
        for (uint32_t s = 0; s < 8; ++s) {
            if (nib(assign, s) != 0xFu) continue;
            if (first_free == 0xFu) first_free = s;
            float cost = slot_cost(s, off);
            if (cost > best_cost) { best_cost = cost; best = s; }
        }
        if (best == 0xFu) best = first_free;
        set_nib(assign, best, k);
    }
    uint32_t im = 0, lm = 0;
    for (uint32_t s = 0; s < 8; ++s) {
        uint32_t a = nib(assign, s);
        if (a == 0xFu) continue;
        uint32_t bit = (inner >> a) & 1u;
        im |= bit << s;
        lm |= (bit ^ 1u) << s;
    }
    c.assign = assign;
    c.inner_mask = im;
    c.leaf_mask = lm;
    return c;
}

uint32_t collapse_all(const nxb_bvh2_node* b2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                      uint32_t slot_mode) {
    if (n == 0) return 0;
    if (n == 1) {
        /* CreateBVH8SingleLeaf (WideConverter.cu:209-222). */
        uint32_t child[8] = {0};
        prim_idx[0] = 0;
        out[0] = make_bvh8_node(b2, b2[0].bounds, child, 0, 0, 0xFFFFFFF0u, 0u, 1u);
        return 1;
    }
    uint32_t node_counter = 1, leaf_counter = 0;
    std::deque<std::pair<uint32_t, uint32_t>> q;
    q.emplace_back(2 * n - 2, 0u);
    while (!q.empty()) {
        uint32_t x = q.front().first, b = q.front().second;
        q.pop_front();
        Collapsed c = collapse_node(b2, x, slot_mode);
        uint32_t inner_count = (uint32_t)__builtin_popcount(c.inner_mask);
        uint32_t leaf_count = c.count - inner_count;
        uint32_t child_base = node_counter;
        node_counter += inner_count;
        uint32_t prim_base = 0;
        if (leaf_count > 0) { prim_base = leaf_counter; leaf_counter += leaf_count; }
        for (uint32_t s = 0; s < 8; ++s) {
            uint32_t a = nib(c.assign, s);
            if (a == 0xFu) continue;
            uint32_t ch = c.child[a];
            if (c.inner_mask & (1u << s))
                q.emplace_back(ch, child_base + bits_below(c.inner_mask, s));
            else
                prim_idx[prim_base + bits_below(c.leaf_mask, s)] = b2[ch].right;
        }
        out[b] = make_bvh8_node(b2, b2[x].bounds, c.child, child_base, prim_base, c.assign,
                                c.inner_mask, c.leaf_mask);
    }
    return node_counter;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

/* ComputeSceneBoundsKernel (Setup.cu:12-39). Returns 1 if a coordinate is
 * non-finite, else 0. */
int nxbo_scene_bounds(const void* prims, int type, uint32_t n, nxb_bvh2_node* nodes, nxb_aabb* scene) {
    bool bad = false;
    nxb_aabb s;
    s.bmin = nxb_float3{INFINITY, INFINITY, INFINITY};
    s.bmax = nxb_float3{-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        nxb_aabb b = prim_bounds(prims, type, i, &bad);
        nodes[i].bounds = b;
        nodes[i].left = NXB_INVALID_IDX;
        nodes[i].right = i;
        s = box_union(s, b);
    }
    *scene = s;
    return bad ? 1 : 0;
}

/* ComputeMortonCodesKernel (Setup.cu:42-59); keys have key_bits 32 or 64. */
void nxbo_morton(const nxb_bvh2_node* nodes, uint32_t n, const nxb_aabb* scene, uint32_t key_bits, void* keys) {
    if (key_bits == 32) {
        uint32_t* k = static_cast<uint32_t*>(keys);
        for (uint32_t i = 0; i < n; ++i) k[i] = (uint32_t)morton_of<uint32_t>(nodes[i].bounds, *scene);
    } else {
        uint64_t* k = static_cast<uint64_t*>(keys);
        for (uint32_t i = 0; i < n; ++i) k[i] = morton_of<uint64_t>(nodes[i].bounds, *scene);
    }
}

/* Morton of explicit normalized coordinates (known-answer tests). */
uint64_t nxbo_morton_point(float x, float y, float z, uint32_t key_bits) {
    if (key_bits == 32)
        return interleave3(quantize(x, 1023.0f), quantize(y, 1023.0f), quantize(z, 1023.0f), 10);
    return interleave3(quantize(x, 2097151.0f), quantize(y, 2097151.0f), quantize(z, 2097151.0f), 21);
}

/* RadixSort (Setup.cu:62-112): stable sort of (key, iota) on bits
 * [begin_bit, end_bit). Equivalent to std::sort on (key window, index). */
void nxbo_sort_pairs(const void* keys, uint32_t n, uint32_t key_bits, uint32_t begin_bit,
                     uint32_t end_bit, void* keys_out, uint32_t* vals_out) {
    uint32_t width = end_bit - begin_bit;
    uint64_t mask = width >= 64 ? ~0ull : ((1ull << width) - 1ull);
    std::vector<std::pair<uint64_t, uint32_t>> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t k = key_bits == 32 ? (uint64_t)static_cast<const uint32_t*>(keys)[i]
                                    : static_cast<const uint64_t*>(keys)[i];
        v[i] = {(k >> begin_bit) & mask, i};
    }
    std::sort(v.begin(), v.end());
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t j = v[i].second;
        vals_out[i] = j;
        if (key_bits == 32)
            static_cast<uint32_t*>(keys_out)[i] = static_cast<const uint32_t*>(keys)[j];
        else
            static_cast<uint64_t*>(keys_out)[i] = static_cast<const uint64_t*>(keys)[j];
    }
}

/* FindNearestNeighbor on one cluster list (tests). box6: num*6 floats. */
void nxbo_find_nn(uint32_t num, uint32_t width, uint32_t radius, const float* box6, uint32_t* nn_out) {
    nxb_aabb box[64];
    for (uint32_t l = 0; l < 64; ++l) box[l] = nxb_aabb{{0, 0, 0}, {0, 0, 0}};
    for (uint32_t l = 0; l < num && l < width; ++l) std::memcpy(&box[l], box6 + 6 * l, 24);
    find_nn_shuffle(num, width, radius, box, nn_out);
}

/* BuildBVH2Kernel (BinaryBuilder.cu:200-273). nodes[0..n) must hold leaves.
 * width 32 = reference schedule (cap 16), 64 = native wave64 schedule (cap 32).
 * Returns the number of internal nodes created. */
uint32_t nxbo_hploc(const void* sorted_keys, uint32_t key_bits, const uint32_t* sorted_vals, uint32_t n,
                    nxb_bvh2_node* nodes, uint32_t width, uint32_t radius, const uint8_t* table,
                    uint32_t seg_log2, uint32_t table_len) {
    if (n == 0) return 0;
    Hploc h;
    h.nodes = nodes;
    h.cluster.assign(sorted_vals, sorted_vals + n);
    h.rep.assign(n, 0);
    h.parent.assign(n, NXB_INVALID_IDX);
    h.moved_parent = NXB_INVALID_IDX;
    std::memset(&h.moved, 0, sizeof(h.moved));
    h.n = n;
    h.next_id = n;
    h.width = width == 64 ? 64 : 32;
    h.cap = h.width / 2;
    h.radius = radius;
    h.table = table;
    h.seg_log2 = seg_log2;
    h.table_len = table_len;
    if (key_bits == 32) hploc_run(h, static_cast<const uint32_t*>(sorted_keys));
    else hploc_run(h, static_cast<const uint64_t*>(sorted_keys));
    return h.next_id - n;
}

/* BuildBVH8Kernel (WideConverter.cu:225-414) in FIFO order. Returns the node count. */
uint32_t nxbo_collapse_bvh8(const nxb_bvh2_node* b2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx) {
    return collapse_all(b2, n, out, prim_idx, NXB_SLOT_GREEDY);
}

/* The same with the child slots chosen by slot_mode (nxb_slot_assignment). */
uint32_t nxbo_collapse_bvh8_slots(const nxb_bvh2_node* b2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                                  uint32_t slot_mode) {
    return collapse_all(b2, n, out, prim_idx, slot_mode);
}

/* Slot nibbles of the auction for one parent box and n <= 8 child boxes
 * (6 floats each: min xyz, max xyz), with the offsets formed as in the
 * collapse; prices_out (8 floats, nullable): the final slot prices. */
uint32_t nxbo_auction_assign(const float* parent6, const float* child6, uint32_t n, float* prices_out) {
    if (n == 0 || n > 8) return 0xFFFFFFFFu;
    nxb_aabb pb;
    std::memcpy(&pb, parent6, sizeof(pb));
    nxb_float3 pc = {pb.bmin.x + pb.bmax.x, pb.bmin.y + pb.bmax.y, pb.bmin.z + pb.bmax.z};
    nxb_float3 offs[8];
    for (uint32_t k = 0; k < n; ++k) {
        const float* c = child6 + 6 * k;
        offs[k] = {pc.x - (c[0] + c[3]), pc.y - (c[1] + c[4]), pc.z - (c[2] + c[5])};
    }
    return auction_assign(pb, offs, n, prices_out);
}

/* ComputeBVH2CostKernel (Eval.cu:12-33), FP32 terms summed in FP64. */
double nxbo_sah_bvh2(const nxb_bvh2_node* nodes, uint32_t count, const nxb_aabb* scene) {
    float sa = box_area(*scene);
    double sum = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        float m = box_area(nodes[i].bounds) / sa;
        float t = nodes[i].left != NXB_INVALID_IDX ? 3.0f * m : 2.0f * m;
        sum += (double)t;
    }
    return sum;
}

/* ComputeBVH8CostKernel (Eval.cu:36-80): inner children weigh 2, leaves 3. */
double nxbo_sah_bvh8(const nxb_bvh8_node* nodes, uint32_t count, const nxb_aabb* scene) {
    float sa = box_area(*scene);
    double sum = 0.0;
    for (uint32_t k = 0; k < count; ++k) {
        const nxb_bvh8_node& n = nodes[k];
        float sx = ffrom((uint32_t)n.e[0] << 23), sy = ffrom((uint32_t)n.e[1] << 23),
              sz = ffrom((uint32_t)n.e[2] << 23);
        for (uint32_t i = 0; i < 8; ++i) {
            if (!n.meta[i]) continue;
            bool internal = (n.meta[i] & 31) >= 24;
            nxb_aabb c;
            c.bmin.x = n.p.x + sx * (float)n.qlox[i];
            c.bmin.y = n.p.y + sy * (float)n.qloy[i];
            c.bmin.z = n.p.z + sz * (float)n.qloz[i];
            c.bmax.x = n.p.x + sx * (float)n.qhix[i];
            c.bmax.y = n.p.y + sy * (float)n.qhiy[i];
            c.bmax.z = n.p.z + sz * (float)n.qhiz[i];
            float m = box_area(c) / sa;
            sum += (double)(internal ? 2.0f * m : 3.0f * m);
        }
    }
    return sum;
}

/* Ordered-tree comparison of two BVH2s from their roots: same shape, same
 * left/right order, bit-identical boxes, same leaf primitives.
 * Returns 0 when equal; else a nonzero code, and *where = visit index. */
int nxbo_compare_bvh2(const nxb_bvh2_node* a, uint32_t root_a, const nxb_bvh2_node* b, uint32_t root_b,
                      uint32_t node_count, uint64_t* where) {
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.reserve(128);
    st.emplace_back(root_a, root_b);
    uint64_t visits = 0;
    while (!st.empty()) {
        auto [x, y] = st.back();
        st.pop_back();
        if (x >= node_count || y >= node_count) { *where = visits; return 1; }
        const nxb_bvh2_node& p = a[x];
        const nxb_bvh2_node& q = b[y];
        if (std::memcmp(&p.bounds, &q.bounds, sizeof(nxb_aabb)) != 0) { *where = visits; return 2; }
        bool la = p.left == NXB_INVALID_IDX, lb = q.left == NXB_INVALID_IDX;
        if (la != lb) { *where = visits; return 3; }
        if (la) {
            if (p.right != q.right) { *where = visits; return 4; }
        } else {
            st.emplace_back(p.right, q.right);
            st.emplace_back(p.left, q.left);
        }
        if (++visits > node_count) { *where = visits; return 5; }
    }
    *where = visits;
    return visits == node_count ? 0 : 6;
}

/* Slot-wise BVH8 comparison from node 0: every byte except child_base and
 * prim_base must match; inner children are followed through child_base,
 * leaves compared through prim_idx. Returns 0 when equal. */
int nxbo_compare_bvh8(const nxb_bvh8_node* a, const uint32_t* pa, uint32_t ca,
                      const nxb_bvh8_node* b, const uint32_t* pb, uint32_t cb, uint32_t n,
                      uint64_t* where) {
    if (ca != cb) { *where = 0; return 1; }
    if (ca == 0) { *where = 0; return 0; }
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.emplace_back(0u, 0u);
    uint64_t visits = 0, leaves = 0;
    while (!st.empty()) {
        auto [x, y] = st.back();
        st.pop_back();
        if (x >= ca || y >= cb) { *where = visits; return 2; }
        nxb_bvh8_node p = a[x], q = b[y];
        uint32_t cbp = p.child_base, cbq = q.child_base, pbp = p.prim_base, pbq = q.prim_base;
        p.child_base = q.child_base = 0;
        p.prim_base = q.prim_base = 0;
        if (std::memcmp(&p, &q, sizeof(p)) != 0) { *where = visits; return 3; }
        uint32_t inner = p.imask, ki = 0;
        for (uint32_t i = 0; i < 8; ++i) {
            if (!p.meta[i]) continue;
            if (inner & (1u << i)) {
                st.emplace_back(cbp + ki, cbq + ki);
                ++ki;
            } else {
                uint32_t r = p.meta[i] & 31u;
                if (pbp + r >= n || pbq + r >= n) { *where = visits; return 4; }
                if (pa[pbp + r] != pb[pbq + r]) { *where = visits; return 5; }
                ++leaves;
            }
        }
        if (++visits > ca) { *where = visits; return 6; }
    }
    *where = visits;
    if (visits != ca) return 7;
    if (leaves != n) return 8;
    return 0;
}

/* Full CPU pipeline (the CPU baseline): bounds, Morton, sort, H-PLOC, BVH8.
 * stage_ms[5] receives per-stage wall times. Outputs: nodes (2n-1),
 * bvh8 (capacity nxb_bvh8_capacity(n)), prim_idx (n). Returns BVH8 node count. */
uint32_t nxbo_build(const void* prims, int type, uint32_t n, uint32_t key_bits, uint32_t width,
                    uint32_t radius, nxb_bvh2_node* nodes, nxb_aabb* scene, nxb_bvh8_node* bvh8,
                    uint32_t* prim_idx, double* stage_ms) {
    double t0 = now_ms();
    nxbo_scene_bounds(prims, type, n, nodes, scene);
    double t1 = now_ms();
    std::vector<uint64_t> keys(n), sorted(n);
    std::vector<uint32_t> vals(n);
    nxbo_morton(nodes, n, scene, key_bits, keys.data());
    double t2 = now_ms();
    if (key_bits == 32) nxbo_sort_pairs(keys.data(), n, 32, 2, 32, sorted.data(), vals.data());
    else nxbo_sort_pairs(keys.data(), n, 64, 1, 64, sorted.data(), vals.data());
    double t3 = now_ms();
    nxbo_hploc(sorted.data(), key_bits, vals.data(), n, nodes, width, radius, nullptr, 0, 0);
    double t4 = now_ms();
    uint32_t c = bvh8 ? nxbo_collapse_bvh8(nodes, n, bvh8, prim_idx) : 0;
    double t5 = now_ms();
    if (stage_ms) {
        stage_ms[0] = t1 - t0; stage_ms[1] = t2 - t1; stage_ms[2] = t3 - t2;
        stage_ms[3] = t4 - t3; stage_ms[4] = t5 - t4;
    }
    return c;
}

// Ray traversal check (no counterpart in the reference, which does not
// trace): the closest hit of each ray by brute force over all triangles in
// index order, with the Moller-Trumbore test ("Fast, Minimum Storage
// Ray/Triangle Intersection", Moller and Trumbore 1997) in FP32 in the
// operation order of the GPU kernel; a hit counts for tmin <= t <= tmax and
// equal distances go to the smaller index. rays: 8 floats each (origin,
// tmin, dir, tmax); hits: t, prim, u, v.
void nxbo_trace_bruteforce(const float* tris, uint32_t n, const float* rays, uint32_t ray_count, void* hits_out) {
    struct V { float x, y, z; };
    auto sub = [](V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto cross = [](V a, V b) { return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
    auto dot = [](V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    struct Hit { float t; uint32_t prim; float u, v; };
    Hit* hits = static_cast<Hit*>(hits_out);
    for (uint32_t r = 0; r < ray_count; ++r) {
        const float* ray = rays + (size_t)r * 8;
        const V o{ray[0], ray[1], ray[2]}, d{ray[4], ray[5], ray[6]};
        const float tmin = ray[3], tmax = ray[7];
        float best = tmax, bu = 0.0f, bv = 0.0f;
        uint32_t bp = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < n; ++i) {
            const float* t9 = tris + (size_t)i * 9;
            const V v0{t9[0], t9[1], t9[2]}, v1{t9[3], t9[4], t9[5]}, v2{t9[6], t9[7], t9[8]};
            const V e1 = sub(v1, v0), e2 = sub(v2, v0);
            const V pv = cross(d, e2);
            const float det = dot(e1, pv);
            if (det == 0.0f) continue;
            const float inv = 1.0f / det;
            const V tv = sub(o, v0);
            const float u = dot(tv, pv) * inv;
            if (!(u >= 0.0f && u <= 1.0f)) continue;
            const V qv = cross(tv, e1);
            const float w = dot(d, qv) * inv;
            if (!(w >= 0.0f && u + w <= 1.0f)) continue;
            const float t = dot(e2, qv) * inv;
            if (!(t >= tmin && t <= best)) continue;
            if (t < best || i < bp) {
                best = t;
                bp = i;
                bu = u;
                bv = w;
            }
        }
        hits[r] = Hit{best, bp, bu, bv};
    }
}

}  // extern "C"
