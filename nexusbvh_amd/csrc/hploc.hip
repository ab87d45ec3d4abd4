// hploc.hip - H-PLOC BVH2 construction (Benthin et al. 2024).
//
// Replaces BuildBVH2Kernel and its helpers (NexusBVH/src/Cuda/BinaryBuilder.cu:
// FindParentId/Delta 16-40, LoadIndices/StoreIndices 42-69,
// MergeClustersCreateBVH2Node 72-124, FindNearestNeighbor 127-163,
// PlocMerge 165-197, kernel 200-273).
//
// One lane per sorted primitive climbs the LBVH bottom-up (Apetrei 2014):
// the first child to reach a parent (atomic exchange on parent[]) retires,
// the second continues with the union range. Whenever a lane's range holds
// more than C primitives (or is the whole scene) its clusters are merged
// with PLOC++ by a group of lanes:
//   * NXB_MERGE_WAVE32: groups of 32 lanes, C = 16 -- the reference schedule.
//     A wave64 runs two independent groups, i.e. two merge tasks at once.
//   * NXB_MERGE_WAVE64: one group of 64 lanes, C = 32.
// The structure built depends only on the LBVH and the geometry, not on the
// order in which tasks run.
//
// Node ids without atomics: the reference takes ids from one global counter
// (BinaryBuilder.cu:85-91); on MI355X one hot address caps at ~10^8
// atomics/s, which would dominate the build. Here every cluster carries a
// representative (the smallest sorted position it covers; lists stay ordered
// by it). Merging clusters i < j retires rep(j); every position but 0 is
// retired exactly once, so id = n - 1 + rep(j) numbers the internal nodes
// bijectively onto [n, 2n-2], deterministically. The root goes to 2n-2; the
// node that retired position n-1 is parked in a spare slot (kMoved) and
// moved to the root's natural id when the root is made.
//
// Nearest neighbours: the reference's shuffle scheme yields, for cluster i,
// the first minimum of union areas (compared as uint32 bits) in the order
// i+1, i-1, i+2, i-2, ..., i+R, i-R. Here every lane writes its box to LDS,
// computes the R forward areas a(i, i+r) and stores them in LDS; lane i then
// reads its backward areas a(i-r, i) from lane i-r's row and scans the
// candidates in that order. Unions use IEEE minimum/maximum, so
// a(i-r, i) == a(i, i-r) bit for bit.
//
// Batched builds (many independent meshes in one launch): the sorted keys
// are 64-bit (mesh << 30 | 30-bit Morton code) and each mesh's positions
// form a contiguous range [lo, lo + cnt). A merge task never leaves its mesh
// (the climb treats the mesh range as the whole array) and numbers its nodes
// with the mesh's own representatives: local id L = cnt - 1 + (rep - lo),
// stored at global id base + L. Leaves come first (leaf g = primitive g),
// then the internal nodes of mesh 0, 1, ...: base = N + lo - mesh - cnt. A
// single mesh is the case lo = 0, cnt = n, base = 0.
//
// Memory ordering across waves: everything one wave hands to another through
// a parent exchange (the cluster records and sizes a merge leaves) is
// written with agent-scope atomic stores and read with agent-scope atomic
// loads, which are coherent across XCDs without cache maintenance. A wave
// waits for its stores to complete before its next exchange; the reader's
// loads are issued only after its exchange returned. Leaves (written by an
// earlier kernel) use plain loads, and so do the internal nodes' stores,
// which only later kernels read (see create_node for the one exception).
//
// BVH8 builds of one mesh write each internal node as a 64-byte wide record
// (both children's boxes, ids and primitive counts) for the collapse, in
// place of its BVH2 node and size: at 28.1M triangles the records cost
// H-PLOC 0.21 ms and save the collapse 0.64 ms.
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

// Two waves per block; each wave is persistent and works through its own
// range of sorted positions.
constexpr int kHplocThreads = 128;
constexpr int kHplocWaves = kHplocThreads / 64;
constexpr int kMaxRadius = 16;

constexpr uint32_t kMoved = 0xFFFFFFFDu;  // temporary id of the node that retired n-1
constexpr uint32_t kRefill = 32;           // free lanes that make a wave start new positions
// Resident blocks per CU (6 waves per SIMD): at 28.1M primitives H-PLOC took
// 4.5 ms with 12 to 14, 4.9 ms with 8 and 5.1 ms with 16.
constexpr uint32_t kHplocBlocksPerCu = kHplocMaxBlocksPerCu;
constexpr uint32_t kHplocMaxChunk = 1024;
constexpr uint32_t kHplocRegions = 8;         // XCDs of the chip
constexpr uint32_t kHplocCounterStride = 32;  // words between region counters (own 128-byte line)

struct HplocLds {
    // cluster boxes, 32 bytes a slot (lx ly lz hx | hy hz), so that one
    // address serves both loads of a slot; kMaxRadius spare slots past the
    // last, so that the neighbour search reads slot + r unconditionally
    struct alignas(32) BoxSlot {
        float4 a;
        float2 b;
    } box[64 + kMaxRadius];
    uint32_t cid[64];
    uint32_t rep[64];
    uint32_t sz[64];                  // primitives in the cluster
    uint32_t nn[64];
    // in-wave hand-over of the parent exchange
    uint32_t mlane[64];               // slot & 63 -> a lane that posted it
    uint32_t mslot[64];               // lane -> slot it exchanges at rightwards this step, or kInvalid
    uint32_t mclaim[64];              // lane -> range end handed to it, or kInvalid
};

__device__ __forceinline__ uint64_t delta_k(const uint32_t* m, uint32_t a, uint32_t b) {
    return (((uint64_t)m[a] << 32) | a) ^ (((uint64_t)m[b] << 32) | b);
}
__device__ __forceinline__ uint64_t delta_k(const uint64_t* m, uint32_t a, uint32_t b) {
    uint64_t d = m[a] ^ m[b];
    return d ? d : (uint64_t)(a ^ b);
}
// Batched keys (mesh << 30 | 30-bit code): the 32-bit Delta of the mesh alone,
// with positions local to the mesh (first position lo). Both keys lie in one
// mesh, so the mesh bits in their low words cancel in the XOR.
__device__ __forceinline__ uint64_t delta_batch(const uint64_t* m, uint32_t a, uint32_t b, uint32_t lo) {
    return (((uint64_t)(uint32_t)m[a] << 32) | (a - lo)) ^ (((uint64_t)(uint32_t)m[b] << 32) | (b - lo));
}

__device__ __forceinline__ void lds_put_box(HplocLds& s, uint32_t i, const Box& b) {
    s.box[i].a = make_float4(b.lx, b.ly, b.lz, b.hx);
    s.box[i].b = make_float2(b.hy, b.hz);
}
__device__ __forceinline__ Box lds_get_box(const HplocLds& s, uint32_t i) {
    float4 a = s.box[i].a;
    float2 c = s.box[i].b;
    Box b;
    b.lx = a.x; b.ly = a.y; b.lz = a.z; b.hx = a.w; b.hy = c.x; b.hz = c.y;
    return b;
}

// FindNearestNeighbor (BinaryBuilder.cu:127-163) without branching on
// lanes: the same candidates in the same order as the reference (forward
// a(i, i+r) then backward a(i-r, i), r = 1..R, first minimum wins). Every
// forward area is computed (candidates past the group's clusters read spare
// or stale slots); those that are not clusters of the group, or lie beyond
// the lane's own radius R (the task's, uniform over its group), become
// kInvalid. The loop runs to Rw, the largest radius of the wave's groups,
// on scalar control.
__device__ __forceinline__ uint32_t nearest_uniform(const HplocLds& s, const Box& own, uint32_t slot, uint32_t gl,
                                                    uint32_t num, uint32_t Rw, uint32_t R) {
    // candidates r < lim are the lane's; the winner is kept as the signed
    // offset r or -r (0: none), so the loop does no per-lane index arithmetic.
    // A backward candidate needs no radius test: lane i-r of the same group
    // masked its forward area at distance r by the same radius.
    const int lim = min((int)num - (int)gl, (int)R + 1);
    const uint32_t lane4 = lane_id() * 4u;
    uint32_t best = kInvalid;
    int off = 0;
#pragma unroll 2
    for (uint32_t r = 1; r <= Rw; ++r) {
        const uint32_t a = __float_as_uint(box_area(box_union(own, lds_get_box(s, slot + r))));
        const uint32_t af = (int)r < lim ? a : kInvalid;
        // lane - r (wrapping for lanes below r, whose value is not used)
        uint32_t ab = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(lane4 - 4u * r), (int)af);
        ab = gl >= r ? ab : kInvalid;
        if (af < best) off = (int)r;
        best = min(best, af);
        if (ab < best) off = -(int)r;
        best = min(best, ab);
    }
    return off ? (uint32_t)((int)gl + off) : kInvalid;
}

// The same search unrolled for a radius known at compile time: constant LDS
// and lane-shift offsets, no loop control. A scheduling barrier after each
// candidate keeps the candidates' loads from being hoisted together (which
// would raise the kernel's register count and cost occupancy).
template <uint32_t R>
__device__ __forceinline__ uint32_t nearest_fixed(const HplocLds& s, const Box& own, uint32_t slot, uint32_t gl,
                                                  uint32_t num) {
    const int lim = (int)num - (int)gl;
    const uint32_t lane4 = lane_id() * 4u;
    uint32_t best = kInvalid;
    int off = 0;
#pragma unroll
    for (uint32_t r = 1; r <= R; ++r) {
        const uint32_t a = __float_as_uint(box_area(box_union(own, lds_get_box(s, slot + r))));
        const uint32_t af = (int)r < lim ? a : kInvalid;
        uint32_t ab = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(lane4 - 4u * r), (int)af);
        ab = gl >= r ? ab : kInvalid;
        if (af < best) off = (int)r;
        best = min(best, af);
        if (ab < best) off = -(int)r;
        best = min(best, ab);
        __builtin_amdgcn_sched_barrier(0);
    }
    return off ? (uint32_t)((int)gl + off) : kInvalid;
}

struct NodeStore {
    nxb_bvh2_node* nodes;
    bool wide_ok;              // node_count * 32 bytes fits a 32-bit buffer offset
    nxb_bvh2_node* moved;      // spare slot per mesh for kMoved
    uint32_t* moved_parent;    // per mesh: id of the node whose child is kMoved
    uint32_t* moved_size;      // per mesh: primitive count of the kMoved node
    uint32_t* nsize;           // primitive count of internal node id (index id - n); null: not kept
    uint32_t n;                // leaves (all meshes)
    uint32_t node_count;       // nodes (all meshes): 2n - meshes
    uint4* rec8 = nullptr;     // one mesh, BVH8 build: wide record of internal node id at id - n
    uint4* moved_rec = nullptr;  // spare wide record for kMoved
};

// The mesh of a merge task: sorted positions [lo, lo + cnt), local node id L
// stored at base + L.
struct MeshCtx {
    uint32_t lo, cnt, mesh, base;
};

// 16-byte agent-coherent accesses. The relaxed agent-scope atomics of 64-bit
// words compile to plain loads/stores with the sc1 cache policy; the same
// policy on 16-byte buffer loads/stores moves 32 bytes in two memory
// transactions instead of four. A buffer covers its whole array: offsets past
// it read zeros and drop stores.
constexpr int kCachePolicySc1 = 16;
constexpr int kRawBufferWord3 = 0x00020000;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Cluster records. After a merge over a range, position p of the range
// holds one of the clusters the merge left (the first `count` positions):
// its box, id and representative. The parent merge reads them from their
// positions in one round of loads, without a second, dependent load of the
// boxes by id. A merged range's first cluster always holds the range's first
// leaf, so its representative is the range start; that record's
// representative word carries the cluster count instead.
// Records are two planes of 16 bytes per position (plane A: lx ly lz hx;
// plane B: hy hz id rep), so that the 16-byte access of consecutive lanes to
// consecutive positions covers whole 64-byte segments: one store instruction
// writes 16 lanes x 16 bytes contiguously instead of every other 16 bytes.
// Positions never merged have no record: their cluster is the leaf, read by
// id from the sorted permutation and the leaf nodes.
struct ClusterRec {
    Box b;
    uint32_t cid, rep;
};

__device__ __forceinline__ ClusterRec load_rec(const NodeStore& ns, const uint32_t* recs, uint32_t p) {
    ClusterRec r;
    if (ns.wide_ok) {
        const __amdgpu_buffer_rsrc_t ra =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(recs), (short)0, (int)(ns.n * 16u), kRawBufferWord3);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t*>(recs + (size_t)ns.n * 4), (short)0, (int)(ns.n * 16u), kRawBufferWord3);
        u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(ra, (int)(p * 16u), 0, kCachePolicySc1);
        u32x4 c = __builtin_amdgcn_raw_buffer_load_b128(rb, (int)(p * 16u), 0, kCachePolicySc1);
        r.b.lx = __uint_as_float(a.x); r.b.ly = __uint_as_float(a.y); r.b.lz = __uint_as_float(a.z);
        r.b.hx = __uint_as_float(a.w); r.b.hy = __uint_as_float(c.x); r.b.hz = __uint_as_float(c.y);
        r.cid = c.z;
        r.rep = c.w;
    } else {
        const uint64_t* qa = reinterpret_cast<const uint64_t*>(recs + (size_t)p * 4);
        const uint64_t* qb = reinterpret_cast<const uint64_t*>(recs + (size_t)ns.n * 4 + (size_t)p * 4);
        uint64_t w0 = load_agent64(qa), w1 = load_agent64(qa + 1), w2 = load_agent64(qb), w3 = load_agent64(qb + 1);
        r.b.lx = __uint_as_float((uint32_t)w0); r.b.ly = __uint_as_float((uint32_t)(w0 >> 32));
        r.b.lz = __uint_as_float((uint32_t)w1); r.b.hx = __uint_as_float((uint32_t)(w1 >> 32));
        r.b.hy = __uint_as_float((uint32_t)w2); r.b.hz = __uint_as_float((uint32_t)(w2 >> 32));
        r.cid = (uint32_t)w3;
        r.rep = (uint32_t)(w3 >> 32);
    }
    return r;
}
__device__ __forceinline__ void store_rec(const NodeStore& ns, uint32_t* recs, uint32_t p, const Box& b, uint32_t cid,
                                          uint32_t rep) {
    if (ns.wide_ok) {
        const __amdgpu_buffer_rsrc_t ra =
            __builtin_amdgcn_make_buffer_rsrc(recs, (short)0, (int)(ns.n * 16u), kRawBufferWord3);
        const __amdgpu_buffer_rsrc_t rb =
            __builtin_amdgcn_make_buffer_rsrc(recs + (size_t)ns.n * 4, (short)0, (int)(ns.n * 16u), kRawBufferWord3);
        u32x4 a = {__float_as_uint(b.lx), __float_as_uint(b.ly), __float_as_uint(b.lz), __float_as_uint(b.hx)};
        u32x4 c = {__float_as_uint(b.hy), __float_as_uint(b.hz), cid, rep};
        __builtin_amdgcn_raw_buffer_store_b128(a, ra, (int)(p * 16u), 0, kCachePolicySc1);
        __builtin_amdgcn_raw_buffer_store_b128(c, rb, (int)(p * 16u), 0, kCachePolicySc1);
    } else {
        uint64_t* qa = reinterpret_cast<uint64_t*>(recs + (size_t)p * 4);
        uint64_t* qb = reinterpret_cast<uint64_t*>(recs + (size_t)ns.n * 4 + (size_t)p * 4);
        store_agent64(qa, (uint64_t)__float_as_uint(b.lx) | ((uint64_t)__float_as_uint(b.ly) << 32));
        store_agent64(qa + 1, (uint64_t)__float_as_uint(b.lz) | ((uint64_t)__float_as_uint(b.hx) << 32));
        store_agent64(qb, (uint64_t)__float_as_uint(b.hy) | ((uint64_t)__float_as_uint(b.hz) << 32));
        store_agent64(qb + 1, (uint64_t)cid | ((uint64_t)rep << 32));
    }
}

__device__ __forceinline__ void load_node_agent(const nxb_bvh2_node* p, uint64_t w[4]) {
    const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
    w[0] = load_agent64(q); w[1] = load_agent64(q + 1); w[2] = load_agent64(q + 2); w[3] = load_agent64(q + 3);
}
__device__ __forceinline__ void store_words_agent(nxb_bvh2_node* p, const uint64_t w[4]) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    store_agent64(q, w[0]); store_agent64(q + 1, w[1]); store_agent64(q + 2, w[2]); store_agent64(q + 3, w[3]);
}

// Wide record of an internal node, read by the BVH8 collapse of the same
// build: words 0-5 the left child's box, 6-11 the right child's, then the
// left and right ids and their primitive counts (64 bytes, one segment).
__device__ __forceinline__ void rec8_words(const Box& lb, const Box& rb, uint32_t left, uint32_t right,
                                           uint32_t lsz, uint32_t rsz, uint4 q[4]) {
    q[0] = make_uint4(__float_as_uint(lb.lx), __float_as_uint(lb.ly), __float_as_uint(lb.lz), __float_as_uint(lb.hx));
    q[1] = make_uint4(__float_as_uint(lb.hy), __float_as_uint(lb.hz), __float_as_uint(rb.lx), __float_as_uint(rb.ly));
    q[2] = make_uint4(__float_as_uint(rb.lz), __float_as_uint(rb.hx), __float_as_uint(rb.hy), __float_as_uint(rb.hz));
    q[3] = make_uint4(left, right, lsz, rsz);
}
__device__ __forceinline__ void store_rec8_agent(uint4* p, const uint4 q[4]) {
    uint64_t* w = reinterpret_cast<uint64_t*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        store_agent64(w + 2 * i, (uint64_t)q[i].x | ((uint64_t)q[i].y << 32));
        store_agent64(w + 2 * i + 1, (uint64_t)q[i].z | ((uint64_t)q[i].w << 32));
    }
}

// Creates the node merging clusters (left, right) with box b and size
// primitives (lb, rb, lsz, rsz: the two clusters' boxes and primitive
// counts, for the wide record). Returns its id.
template <bool REC>
__device__ __forceinline__ uint32_t create_node(const NodeStore& ns, const MeshCtx& mc, const Box& b, uint32_t left,
                                                uint32_t right, uint32_t right_rep, uint32_t size, bool root_step,
                                                uint32_t* flags, const Box& lb, const Box& rb, uint32_t lsz,
                                                uint32_t rsz) {
    const uint32_t n = ns.n;
    const uint32_t local = mc.cnt - 1 + (right_rep - mc.lo);
    const uint32_t root_local = 2 * mc.cnt - 2;
    const uint32_t natural = mc.base + local;
    uint32_t id = root_step ? mc.base + root_local : (local == root_local ? kMoved : natural);
    if (local > root_local || natural >= ns.node_count) { atomic_or_agent(flags, kFlagBadIndex); return 0; }
    uint32_t* moved_parent = ns.moved_parent + mc.mesh;
    if (REC && ns.rec8) {
        // BVH8 build: the wide record stands for the internal node (its box
        // is the union of the children's, its size their sum); the same
        // parking of kMoved, on records.
        if (root_step && local != root_local) {
            const uint64_t* mr = reinterpret_cast<const uint64_t*>(ns.moved_rec);
            uint64_t* dst = reinterpret_cast<uint64_t*>(ns.rec8 + 4ull * (natural - n));
#pragma unroll
            for (int i = 0; i < 8; ++i) dst[i] = load_agent64(mr + i);
            if (left == kMoved) {
                left = natural;
            } else if (right == kMoved) {
                right = natural;
            } else {
                const uint32_t p = load_agent(moved_parent);
                if (p < ns.node_count && p >= n) {
                    // the ids are the record's words 12 and 13
                    uint64_t* q = reinterpret_cast<uint64_t*>(ns.rec8 + 4ull * (p - n) + 3);
                    const uint64_t lr = load_agent64(q);
                    uint32_t pl = (uint32_t)lr, pr = (uint32_t)(lr >> 32);
                    if (pl == kMoved) pl = natural;
                    else if (pr == kMoved) pr = natural;
                    else atomic_or_agent(flags, kFlagBadIndex);
                    store_agent64(q, (uint64_t)pl | ((uint64_t)pr << 32));
                } else {
                    atomic_or_agent(flags, kFlagBadIndex);
                }
            }
        } else if (left == kMoved || right == kMoved) {
            store_agent(moved_parent, id);
        }
        uint4 q[4];
        rec8_words(lb, rb, left, right, lsz, rsz, q);
        if (id == kMoved) {
            store_rec8_agent(ns.moved_rec, q);
        } else if (left == kMoved || right == kMoved) {
            store_rec8_agent(ns.rec8 + 4ull * (id - n), q);
        } else {
            uint4* p = ns.rec8 + 4ull * (id - n);
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
        }
        return id;
    }
    nxb_bvh2_node* moved = ns.moved + mc.mesh;
    uint32_t* moved_size = ns.moved_size + mc.mesh;
    if (root_step && local != root_local) {
        // the parked node takes the root's natural id
        uint64_t w[4];
        load_node_agent(moved, w);
        store_words_agent(ns.nodes + natural, w);
        if (ns.nsize) ns.nsize[natural - n] = load_agent(moved_size);
        if (left == kMoved) {
            left = natural;
        } else if (right == kMoved) {
            right = natural;
        } else {
            uint32_t p = load_agent(moved_parent);
            if (p < ns.node_count) {
                uint64_t* q = reinterpret_cast<uint64_t*>(ns.nodes + p) + 3;
                uint64_t lr = load_agent64(q);
                uint32_t pl = (uint32_t)lr, pr = (uint32_t)(lr >> 32);
                if (pl == kMoved) pl = natural;
                else if (pr == kMoved) pr = natural;
                else atomic_or_agent(flags, kFlagBadIndex);
                store_agent64(q, (uint64_t)pl | ((uint64_t)pr << 32));
            } else {
                atomic_or_agent(flags, kFlagBadIndex);
            }
        }
    } else if (left == kMoved || right == kMoved) {
        store_agent(moved_parent, id);
    }
    // No wave of this kernel reads a node back, except the root step, which
    // rewrites the child ids of the node that points at the parked one: that
    // node is stored agent-coherently; every other node is a plain store
    // (whole lines in L2, written back when the kernel ends).
    if (id == kMoved) store_node_agent(moved, 0, b, left, right);
    else if (left == kMoved || right == kMoved) store_node_agent(ns.nodes, id, b, left, right);
    else store_node(ns.nodes, id, b, left, right);
    if (id == kMoved) store_agent(moved_size, size);
    else if (ns.nsize) ns.nsize[id - n] = size;
    return id;
}

// Per-segment radius policy: a merge task over sorted positions
// [l_start, r_end) within one segment's span takes the radius of the segment
// of l_start; wider tasks (the upper levels of the tree) and table entries 0
// take the fixed radius.
__device__ __forceinline__ uint32_t radius_for(const RadiusParams& rp, uint32_t l_start, uint32_t r_end) {
    uint32_t r = rp.radius;
    if (rp.table && r_end - l_start <= (1u << rp.seg_log2)) {
        uint32_t seg = l_start >> rp.seg_log2;
        if (seg < rp.table_len) {
            uint32_t t = rp.table[seg];
            if (t >= 1) r = t;
        }
    }
    return r > (uint32_t)kMaxRadius ? (uint32_t)kMaxRadius : r;
}

// PlocMerge for the group this lane belongs to. All lanes of the wave call
// it together; a group with has_task == false does nothing.
//   GW  : lanes per group (32 or 64), C = GW / 2
//   gl  : lane index within the group; gbase: first LDS slot of the group
//   gmask: 64-bit mask of the group's lanes
// REC: the build may write wide records (ns.rec8); batched builds never do.
template <int GW, bool REC>
__device__ __forceinline__ void ploc_merge(HplocLds& s, bool has_task, uint32_t l_start, uint32_t split,
                                           uint32_t r_end, bool final, uint32_t gl, uint32_t gbase,
                                           uint64_t gmask, const uint32_t* cluster, uint32_t* recs, uint32_t* csizes,
                                           const NodeStore& ns, const MeshCtx& mc, const RadiusParams& rp,
                                           uint32_t* flags) {
    constexpr uint32_t C = GW / 2;
    const uint32_t slot = gbase + gl;
    uint32_t cid = kInvalid, rep = 0, sz = 0, num_loaded = 0;

    // LoadIndices (BinaryBuilder.cu:42-57): the left child's clusters, then
    // the right child's. A child of at most C primitives was never merged:
    // its clusters are its leaves in sorted order (one primitive each), read
    // through the sorted permutation. A merged child left its clusters'
    // records at the front of its range, their count in the first record.
    // Lanes [0, C) read the left child's clusters and lanes [C, 2C) the
    // right child's, all at once; the valid entries are then compacted in
    // that order through LDS.
    {
        const bool lhalf = gl < C;
        const uint32_t k = lhalf ? gl : gl - C;
        const uint32_t first = lhalf ? l_start : split;
        const uint32_t span = lhalf ? split - l_start : r_end - split;
        const bool merged = span > C;
        const uint32_t pos = first + k;
        bool v = has_task && k < min(span, C);
        uint32_t lcid = kInvalid, lrep = 0, lsz = 0;
        Box b = {0, 0, 0, 0, 0, 0};
        if (v && merged) {
            const ClusterRec r = load_rec(ns, recs, pos);
            lsz = load_agent(&csizes[pos]);
            lcid = r.cid;
            lrep = r.rep;
            b = r.b;
        } else if (v) {
            // a leaf: written by an earlier kernel, plain loads
            lcid = cluster[pos];
            if (lcid >= ns.n) { atomic_or_agent(flags, kFlagBadIndex); lcid = 0; }
            b = load_box(ns.nodes, lcid);
            lrep = pos;
            lsz = 1u;
        }
        // the half's first lane holds the child's cluster count (merged child)
        const uint32_t count = (uint32_t)__shfl((int)lrep, (int)(gbase + (lhalf ? 0u : C)));
        if (v) {
            if (merged) {
                v = k < count;
                if (k == 0) lrep = first;
            }
            v = v && lcid != kInvalid;
            if (v && lcid != kMoved && lcid >= ns.node_count) atomic_or_agent(flags, kFlagBadIndex);
        }
        const uint64_t vm = ballot(v) & gmask;
        if (v) {
            const uint32_t dst = gbase + mask_rank(vm);
            lds_put_box(s, dst, b);
            s.cid[dst] = lcid;
            s.rep[dst] = lrep;
            s.sz[dst] = lsz;
        }
        wave_lds_sync();
        const uint32_t loaded_now = (uint32_t)__popcll(vm);
        if (gl < loaded_now) {
            cid = s.cid[slot];
            rep = s.rep[slot];
            sz = s.sz[slot];
        }
        wave_lds_sync();
        num_loaded = loaded_now;
    }
    const uint32_t loaded = num_loaded;
    uint32_t num = loaded;

    Box own = {0, 0, 0, 0, 0, 0};
    if (gl < num) own = lds_get_box(s, slot);
    const uint32_t threshold = final ? 1u : C;
    const uint32_t R = has_task ? radius_for(rp, l_start, r_end) : 1u;
    // one radius for every task (no per-segment table): the unrolled search
    // for radius 8; otherwise the loop runs to the largest radius of the
    // wave's groups (each group's radius is uniform over it)
    const uint32_t fixed_r = rp.table ? 0u : min(rp.radius, (uint32_t)kMaxRadius);
    const uint32_t r_wave = GW == 64 ? __builtin_amdgcn_readfirstlane(R)
                                     : max(__builtin_amdgcn_readlane(R, 0), __builtin_amdgcn_readlane(R, 32));

    // LDS slots [gbase, gbase + num) hold own/cid/rep/sz of the group's
    // clusters here and after every compaction below.
    while (num > threshold) {
        // ---- FindNearestNeighbor (BinaryBuilder.cu:127-163) ----
        // lane i computes the forward areas a(i, i+r); a(i-r, i) is lane
        // i-r's forward area at distance r, fetched with a lane shuffle
        uint32_t nn = kInvalid;
        if (fixed_r == 8) nn = nearest_fixed<8>(s, own, slot, gl, num);
        else nn = nearest_uniform(s, own, slot, gl, num, r_wave, R);
        if (gl >= num) nn = kInvalid;

        // ---- MergeClustersCreateBVH2Node (BinaryBuilder.cu:72-124) ----
        s.nn[slot] = nn;
        wave_lds_sync();
        uint32_t nnnn = nn < (uint32_t)GW ? s.nn[gbase + nn] : kInvalid;
        bool mutual = gl < num && nnnn == gl;
        bool merge = mutual && gl < nn;
        uint64_t mm = ballot(merge) & gmask;
        uint32_t mcount = (uint32_t)__popcll(mm);
        const bool root_step = final && num - mcount == 1;
        if (merge) {
            const Box nb = lds_get_box(s, gbase + nn);
            const uint32_t ncid = s.cid[gbase + nn];
            const uint32_t nrep = s.rep[gbase + nn];
            const uint32_t nsz = s.sz[gbase + nn];
            const Box ob = own;
            own = box_union(own, nb);
            cid = create_node<REC>(ns, mc, own, cid, ncid, nrep, sz + nsz, root_step, flags, ob, nb, sz, nsz);
            sz += nsz;
        }
        bool valid = gl < num && (merge || !mutual);
        uint64_t vm = ballot(valid) & gmask;
        uint32_t pos = mask_rank(vm);
        wave_lds_sync();
        if (valid) {
            lds_put_box(s, gbase + pos, own);
            s.cid[gbase + pos] = cid;
            s.rep[gbase + pos] = rep;
            s.sz[gbase + pos] = sz;
        }
        num -= mcount;
        wave_lds_sync();
        if (gl < num) {
            own = lds_get_box(s, slot);
            cid = s.cid[slot];
            rep = s.rep[slot];
            sz = s.sz[slot];
        } else {
            cid = kInvalid;
        }
    }

    // StoreIndices (BinaryBuilder.cu:59-69), as records: the num clusters
    // left, the count in the first record's representative word.
    if (gl < num) {
        store_rec(ns, recs, l_start + gl, own, cid, gl == 0 ? num : rep);
        store_agent(&csizes[l_start + gl], sz);
    }
}

template <typename K, int GW, bool BATCH>
__global__ __launch_bounds__(kHplocThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_hploc(const K* __restrict__ keys, const uint32_t* cluster,
                                                         uint32_t* recs, uint32_t* csizes, uint32_t* parent,
                                                         NodeStore ns, const uint32_t* __restrict__ mesh_lo,
                                                         uint32_t meshes, RadiusParams rp, uint32_t* flags,
                                                         uint32_t* next_chunk, uint32_t first, uint32_t chunk) {
    __shared__ HplocLds lds[kHplocWaves];
    constexpr uint32_t C = GW / 2;
    const uint32_t tid = threadIdx.x;
    HplocLds& s = lds[tid >> 6];
    const uint32_t lane = tid & 63;
    const uint32_t n = ns.n;

    // Persistent waves. The sorted positions are cut into kHplocRegions
    // regions, one per XCD (blocks are dispatched to XCDs round-robin, so
    // block b runs on XCD b % kHplocRegions and the merges of a region stay
    // in one L2). A wave starts positions in order, a batch of free lanes at
    // a time, so that it keeps most lanes busy even though most climbs end
    // after a step or two. Its first `first` positions are fixed by its
    // index in the region; then it takes chunks of `chunk` positions from
    // the region's counter (dynamic chunks keep the waves level where the
    // merge cost varies along the array, e.g. with per-segment radii), and
    // once its region is used up and all its lanes are idle, from the next
    // regions'.
    const uint32_t region_len = (((n + kHplocRegions - 1) / kHplocRegions) + 63u) & ~63u;
    uint32_t region = blockIdx.x % kHplocRegions, tried = 0;
    uint32_t next = 0, end = 0;  // the wave's current chunk (wave-uniform)
    bool more = true;            // chunks may be left (wave-uniform)
    bool have_pf = false;        // lane 0's pf holds the next counter value of the home region
    uint32_t pf = 0;
    {
        const uint32_t j = (blockIdx.x / kHplocRegions) * kHplocWaves + (tid >> 6);
        const uint64_t lo = (uint64_t)region * region_len;
        const uint64_t hi_r = min((uint64_t)n, lo + region_len);
        const uint64_t c = lo + (uint64_t)j * first;
        if (c < hi_r) {
            next = (uint32_t)c;
            end = (uint32_t)min(hi_r, c + first);
        }
    }

    uint32_t left = 0, right = 0, split = 0;
    bool active = false;
    MeshCtx mc{0u, n, 0u, 0u};  // this lane's mesh (the whole array for a single build)
    uint32_t hi = n - 1;

    s.mlane[lane] = lane;
    s.mclaim[lane] = kInvalid;
    wave_lds_sync();

    uint32_t g, gl, gbase;
    uint64_t gmask;
    if (GW == 32) {
        g = lane >> 5;
        gl = lane & 31;
        gbase = g * 32;
        gmask = g ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
    } else {
        g = 0;
        gl = lane;
        gbase = 0;
        gmask = ~0ull;
    }

    for (;;) {
        // Start the next positions on the free lanes (at least kRefill of
        // them, or all lanes free).
        const uint64_t free_lanes = ballot(!active);
        const uint32_t nfree = (uint32_t)__popcll(free_lanes);
        while (more && next >= end && (tried == 0 ? nfree >= kRefill : nfree == 64u)) {
            // the region's dynamic chunks start after the first positions of
            // its waves
            const uint32_t waves_r =
                gridDim.x > region ? ((gridDim.x - region + kHplocRegions - 1) / kHplocRegions) * kHplocWaves : 0u;
            const uint64_t lo = (uint64_t)region * region_len;
            const uint64_t hi_r = min((uint64_t)n, lo + region_len);
            uint32_t* ctr = &next_chunk[region * kHplocCounterStride];
            const uint64_t base = lo + (uint64_t)waves_r * first;
            uint32_t c = 0;
            if (have_pf) {
                c = pf;  // lane 0's, taken with the stores of the refill that used up the last chunk
                have_pf = false;
            } else if (lane == 0) {
                // When stealing, a plain read first: the counter only grows,
                // so a region it shows used up is used up, and idle waves add
                // no atomics to the counters still in use.
                c = tried ? load_agent(ctr) : 0u;
                if (base + c < hi_r) c = atomic_add_agent(ctr, chunk);
            }
            c = __builtin_amdgcn_readfirstlane(c);
            const uint64_t pos = base + c;
            if (pos < hi_r) {
                next = (uint32_t)pos;
                end = (uint32_t)min(hi_r, pos + chunk);
            } else if (++tried == kHplocRegions) {
                more = false;
            } else {
                region = (region + 1) % kHplocRegions;
            }
        }
        if (next < end && (nfree >= kRefill || free_lanes == ~0ull)) {
            const uint32_t i = next + mask_rank(free_lanes);
            if (!active && i < end) {
                left = right = i;
                split = 0;
                active = true;
                if constexpr (BATCH) {
                    const uint32_t m = (uint32_t)((uint64_t)keys[i] >> 30);
                    if (m < meshes) {
                        const uint32_t lo = mesh_lo[m];
                        mc = MeshCtx{lo, mesh_lo[m + 1] - lo, m, 0u};
                        mc.base = n + lo - m - mc.cnt;
                    } else {
                        atomic_or_agent(flags, kFlagBadIndex);
                        active = false;
                    }
                    hi = mc.lo + mc.cnt - 1;
                }
            }
            next = min(end, next + nfree);
            if (next >= end && more && tried == 0 && !have_pf) {
                // this batch uses up the chunk: take the next one now, so
                // that the atomic's latency hides behind the climb
                if (lane == 0) pf = atomic_add_agent(&next_chunk[region * kHplocCounterStride], chunk);
                have_pf = true;
            }
        }
        if (!ballot(active)) {
            if (!more && next >= end) break;
            continue;
        }
        bool to_right = false;
        if (active) {
            if constexpr (BATCH) {
                const uint64_t* k64 = reinterpret_cast<const uint64_t*>(keys);
                to_right = (left == mc.lo) ||
                           (right != hi && delta_batch(k64, right, right + 1, mc.lo) <
                                               delta_batch(k64, left - 1, left, mc.lo));
            } else {
                to_right = (left == 0) ||
                           (right != n - 1 && delta_k(keys, right, right + 1) < delta_k(keys, left - 1, left));
            }
        }
        // Two lanes of this wave that reach the same parent slot in this step
        // (the range ending at it going right, the one starting after it
        // going left) meet in LDS instead of exchanging through memory: the
        // left-going lane hands its range end over and stops. Each slot has
        // one candidate per side, so a lane found through the slot hash is
        // the partner iff it posted the same slot; every other lane takes
        // the exchange. The wave's stores are complete (end of the previous
        // step), so the continuing lane reads the partner's clusters as it
        // would after an exchange. Skipped (wave-uniform) unless the wave
        // has lanes going both ways, as on the serial climb near the root.
        if (ballot(active && to_right) != 0 && ballot(active && !to_right) != 0) {
            const uint32_t slot = to_right ? right : left - 1;
            s.mslot[lane] = (active && to_right) ? slot : kInvalid;
            if (active && to_right) s.mlane[slot & 63u] = lane;
            wave_lds_sync();
            if (active && !to_right) {
                const uint32_t p = s.mlane[slot & 63u] & 63u;
                if (s.mslot[p] == slot) {
                    s.mclaim[p] = right;
                    active = false;
                }
            }
            wave_lds_sync();
        }
        bool handed = false;
        if (active && to_right) {
            const uint32_t c = s.mclaim[lane];
            if (c != kInvalid) {
                s.mclaim[lane] = kInvalid;
                split = right + 1;
                right = c;
                handed = true;
            }
        }
        if (active && !handed) {
            uint32_t prev;
            if (to_right) {
                prev = atomic_xchg_agent(&parent[right], left);
                if (prev != kInvalid) {
                    if (prev > hi || prev < mc.lo) { atomic_or_agent(flags, kFlagBadIndex); prev = hi; }
                    split = right + 1;
                    right = prev;
                }
            } else {
                prev = atomic_xchg_agent(&parent[left - 1], right);
                if (prev != kInvalid) {
                    if (prev > hi || prev < mc.lo) { atomic_or_agent(flags, kFlagBadIndex); prev = mc.lo; }
                    split = left;
                    left = prev;
                }
            }
            if (prev == kInvalid) active = false;
        }
        uint32_t size = right - left + 1;
        bool final = active && size == mc.cnt;
        uint64_t tasks = ballot(active && (size > C || final));
        if (!tasks) continue;

        compiler_barrier();  // cluster/node reads stay after the exchanges
        while (tasks) {
            uint32_t t0 = (uint32_t)__builtin_ctzll(tasks);
            tasks &= tasks - 1;
            uint32_t t1 = 64;
            if (GW == 32 && tasks) {
                t1 = (uint32_t)__builtin_ctzll(tasks);
                tasks &= tasks - 1;
            }
            uint32_t my = (g == 0) ? t0 : t1;
            bool has = my < 64;
            const uint32_t tsel = (g && t1 < 64) ? t1 : t0;  // lane whose task this group runs
            uint32_t l0 = __builtin_amdgcn_readlane(left, t0), s0 = __builtin_amdgcn_readlane(split, t0);
            uint32_t r0 = __builtin_amdgcn_readlane(right, t0);
            uint32_t f0 = __builtin_amdgcn_readlane((uint32_t)final, t0);
            uint32_t l1 = l0, s1 = s0, r1 = r0, f1 = f0;
            if (t1 < 64) {
                l1 = __builtin_amdgcn_readlane(left, t1);
                s1 = __builtin_amdgcn_readlane(split, t1);
                r1 = __builtin_amdgcn_readlane(right, t1);
                f1 = __builtin_amdgcn_readlane((uint32_t)final, t1);
            }
            uint32_t tl = g ? l1 : l0, ts = g ? s1 : s0, tr = g ? r1 : r0;
            bool tf = (g ? f1 : f0) != 0;
            MeshCtx tm = mc;
            if constexpr (BATCH) {
                // the task's mesh (lanes of the other group run another task)
                tm.lo = __shfl(mc.lo, (int)tsel);
                tm.cnt = __shfl(mc.cnt, (int)tsel);
                tm.mesh = __shfl(mc.mesh, (int)tsel);
                tm.base = __shfl(mc.base, (int)tsel);
            }
            ploc_merge<GW, !BATCH>(s, has, tl, ts, tr + 1, tf, gl, gbase, gmask, cluster, recs, csizes, ns, tm, rp, flags);
        }
        // Merge outputs are agent-scope atomic stores; they must be complete
        // before this wave's next parent exchange hands the subtree over.
        wait_memory_ops();
    }
}

// One resident block per slot of the chip (persistent waves): at most
// kHplocBlocksPerCu per CU and no more than the kernel's occupancy allows (a
// block that is not resident would start its fixed positions late), fewer
// when the input gives a wave less than 64 positions.
template <typename Kernel>
static uint32_t hploc_grid(Kernel kernel, uint32_t n, uint32_t blocks_per_cu) {
    static int per_cu = 0;  // per kernel; the same on every gfx950 device
    if (per_cu < 1 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kHplocThreads, 0) != hipSuccess || per_cu < 1))
        per_cu = 1;
    const uint32_t want = blocks_per_cu ? min(blocks_per_cu, kHplocBlocksPerCu) : kHplocBlocksPerCu;
    const uint32_t resident = device_cu_count() * min((uint32_t)per_cu, want);
    const uint32_t most = (n + kHplocThreads - 1) / kHplocThreads;
    return most < resident ? most : resident;
}

// Positions per wave: three quarters of a wave's share fixed, the rest in
// chunks of an eighth of a share (multiples of 64, chunks at most
// kHplocMaxChunk). At 1.1M primitives a half share fixed was 15% slower
// (chunk grabs on the critical path) and at 28.1M a full share fixed was 5%
// slower (waves uneven).
struct HplocChunks {
    uint32_t first, chunk;
};
static HplocChunks hploc_chunks(uint32_t n, uint32_t grid) {
    const uint64_t share = (n + (uint64_t)grid * kHplocWaves - 1) / ((uint64_t)grid * kHplocWaves);
    auto round64 = [](uint64_t v) { return v < 64 ? 64ull : (v + 63) & ~63ull; };
    const uint64_t first = round64(share * 3 / 4), chunk = round64(share / 8);
    return HplocChunks{(uint32_t)first, (uint32_t)(chunk > kHplocMaxChunk ? kHplocMaxChunk : chunk)};
}

template <typename K, int GW, bool BATCH>
static hipError_t launch_hploc_kernel(const K* keys, uint32_t* cluster, uint32_t n, const uint32_t* mesh_lo,
                                      uint32_t meshes, const NodeStore& ns, RadiusParams rp, uint32_t blocks_per_cu,
                                      HplocScratch sc, hipStream_t s) {
    auto kernel = k_hploc<K, GW, BATCH>;
    const uint32_t grid = hploc_grid(kernel, n, blocks_per_cu);
    const HplocChunks cp = hploc_chunks(n, grid);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kHplocThreads), 0, s, keys, cluster, sc.recs, sc.csizes, sc.parent,
                       ns, mesh_lo, meshes, rp, sc.flags, sc.next_chunk, cp.first, cp.chunk);
    return hipGetLastError();
}

hipError_t launch_hploc(const void* keys, uint32_t key_bits, uint32_t* cluster, uint32_t n, nxb_bvh2_node* nodes,
                        uint32_t merge_mode, RadiusParams rp, uint32_t blocks_per_cu, HplocScratch sc,
                        hipStream_t s) {
    if (n == 0) return hipSuccess;
    NodeStore ns{nodes, 2ull * n - 1 < (1ull << 27), sc.moved, sc.moved_parent, sc.moved_size, sc.nsize, n, 2 * n - 1};
    if (sc.rec8 && sc.moved_rec) {
        ns.rec8 = sc.rec8;
        ns.moved_rec = sc.moved_rec;
    }
    const bool w64 = merge_mode == NXB_MERGE_WAVE64;
    if (key_bits == 32) {
        const uint32_t* k = static_cast<const uint32_t*>(keys);
        return w64 ? launch_hploc_kernel<uint32_t, 64, false>(k, cluster, n, nullptr, 1u, ns, rp, blocks_per_cu, sc, s)
                   : launch_hploc_kernel<uint32_t, 32, false>(k, cluster, n, nullptr, 1u, ns, rp, blocks_per_cu, sc, s);
    }
    const uint64_t* k = static_cast<const uint64_t*>(keys);
    return w64 ? launch_hploc_kernel<uint64_t, 64, false>(k, cluster, n, nullptr, 1u, ns, rp, blocks_per_cu, sc, s)
               : launch_hploc_kernel<uint64_t, 32, false>(k, cluster, n, nullptr, 1u, ns, rp, blocks_per_cu, sc, s);
}

hipError_t launch_hploc_batch(const uint64_t* keys, uint32_t* cluster, uint32_t n, const uint32_t* mesh_lo,
                              uint32_t meshes, nxb_bvh2_node* nodes, uint32_t merge_mode, RadiusParams rp,
                              uint32_t blocks_per_cu, HplocScratch sc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    NodeStore ns{nodes, 2ull * n - meshes < (1ull << 27), sc.moved, sc.moved_parent, sc.moved_size, sc.nsize, n,
                 2 * n - meshes};
    return merge_mode == NXB_MERGE_WAVE64
               ? launch_hploc_kernel<uint64_t, 64, true>(keys, cluster, n, mesh_lo, meshes, ns, rp, blocks_per_cu, sc, s)
               : launch_hploc_kernel<uint64_t, 32, true>(keys, cluster, n, mesh_lo, meshes, ns, rp, blocks_per_cu, sc, s);
}

}  // namespace nxb
