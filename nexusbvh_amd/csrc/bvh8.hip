// bvh8.hip - BVH2 -> compressed 8-wide BVH (Ylitie et al. 2017).
//
// Replaces BuildBVH8Kernel and its helpers (NexusBVH/src/Cuda/WideConverter.cu:
// GetCost 25-30, GreedyAssignment 106-153, CreateBVH8Node 156-206,
// CreateBVH8SingleLeaf 209-222, kernel 225-414).
//
// Per BVH8 node: collapse up to 8 BVH2 descendants (push both children of
// the current node, the smaller-area one into the popped slot, then pop the
// highest-index inner slot), greedy octant slot assignment, then 8-bit
// child-box quantization against the node's box.
//
// Scheduling (gfx950): persistent waves. A work item is one BVH8 *inner*
// node; its slot index is its BVH8 node index, so work[k] holds the BVH2
// node that BVH8 node k is built from. Lanes take node indices in increasing
// order from an atomic counter and wait on work[k] until the parent (always
// a smaller index, already taken by a running lane) publishes it with an
// agent-scope store. The producing lane keeps its first inner child and
// builds it next (depth-first, no publish/poll round trip on that path),
// marking the slot kTaken so whoever drew it moves on. Leaf children are
// written to primIdx by the parent directly, so leaves cost no work item.
// Leaf ranges in primIdx are fixed top-down from subtree primitive counts
// (no atomics). Node indices and completion share one 64-bit counter
// (finished << 32 | allocated), updated once per wave iteration for all its
// lanes (a hot address takes only ~10^8 atomics/s on MI355X); waiting lanes
// stop when finished == allocated. Waits are bounded by spin_limit; hitting
// it sets the timeout flag and exits.
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

constexpr int kBvh8Threads = 256;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
// work[k] value of a node that its producer builds itself (depth-first).
constexpr uint32_t kTaken = 0xFFFFFFFEu;

__device__ __forceinline__ uint32_t nib(uint32_t x, uint32_t i) { return (x >> (4 * i)) & 0xFu; }
__device__ __forceinline__ uint32_t bits_below(uint32_t x, uint32_t i) {
    return (uint32_t)__popc(x & ((1u << i) - 1u));
}
// CeilLog2 / InvPow2 (BuilderUtils.h:335-347).
__device__ __forceinline__ uint32_t ceil_log2(float x) {
    uint32_t ix = __float_as_uint(x);
    return ((ix >> 23) & 0xFFu) + ((ix & 0x7FFFFFu) != 0);
}
__device__ __forceinline__ float inv_pow2(uint32_t e) { return __uint_as_float((uint32_t)(254 - (int)e) << 23); }

// GetCost (WideConverter.cu:25-30): exact +-1 products, FMA-neutral.
__device__ __forceinline__ float slot_cost(uint32_t s, float ox, float oy, float oz) {
    float sx = ((s >> 2) & 1) ? -1.0f : 1.0f;
    float sy = ((s >> 1) & 1) ? -1.0f : 1.0f;
    float sz = (s & 1) ? -1.0f : 1.0f;
    return sx * ox + sy * oy + sz * oz;
}

__device__ __forceinline__ uint32_t with_nib(uint32_t x, uint32_t i, uint32_t v) {
    return (x & ~(0xFu << (4 * i))) | (v << (4 * i));
}

constexpr uint32_t kAuctionMaxBids = 4096;   // per epsilon phase
constexpr uint32_t kAuctionMaxPhases = 16;

// Auction slot assignment (the reference's compiled-out alternative to the
// greedy one, WideConverter.cu:52-102 with its cost table at 336-361;
// Bertsekas' auction algorithm with epsilon scaling, theta = 8). Children
// bid for slots with reward cost(c, s) - price(s); the winner's price rises
// by (best - second best) + epsilon and an outbid child bids again. Bounded:
// at most kAuctionMaxPhases phases of kAuctionMaxBids bids; children left
// unassigned (or non-finite costs) take the first free slots in child order.
// ox/oy/oz: (parent min + max) - (child min + max) per child.
__device__ __forceinline__ uint32_t auction_assign(const Box& pb, const float ox[8], const float oy[8],
                                                   const float oz[8], uint32_t n) {
    const float max_dim = fmaxf(fmaxf(pb.hx - pb.lx, pb.hy - pb.ly), pb.hz - pb.lz);
    const float scale = (max_dim == 0.0f ? 0.0f : 10.0f / max_dim) * 0.5f;
    float cx[8], cy[8], cz[8];
    float max_cost = -3.402823466e+38f;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        cx[k] = ox[k] * scale;
        cy[k] = oy[k] * scale;
        cz[k] = oz[k] * scale;
        const float cost = fabsf(cx[k]) + fabsf(cy[k]) + fabsf(cz[k]);
        if (k < n && cost > max_cost) max_cost = cost;
    }
    float prices[8];
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) prices[s] = 0.0f;
    const float threshold = 1.0f / (float)n;
    float epsilon = fmaxf(max_cost, threshold);
    uint32_t assign = 0xFFFFFFFFu;
    bool done = !(epsilon < INFINITY);  // non-finite costs: first free slots
    for (uint32_t phase = 0; !done && epsilon >= threshold && phase < kAuctionMaxPhases; ++phase) {
        assign = 0xFFFFFFFFu;
        uint32_t bidders = 0x76543210u, left = n;
        for (uint32_t bids = 0; left > 0; ++bids) {
            if (bids == kAuctionMaxBids) { done = true; break; }
            const uint32_t c = nib(bidders, --left);
            float ccx = 0.0f, ccy = 0.0f, ccz = 0.0f;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k)
                if (k == c) { ccx = cx[k]; ccy = cy[k]; ccz = cz[k]; }
            float win = -3.402823466e+38f, second = -3.402823466e+38f;
            uint32_t ws = 0xFu;
#pragma unroll
            for (uint32_t t = 0; t < 8; ++t) {
                const float reward = slot_cost(t, ccx, ccy, ccz) - prices[t];
                if (reward > win) { second = win; win = reward; ws = t; }
                else if (reward > second) second = reward;
            }
            if (ws == 0xFu) { done = true; break; }
            float raise = (win - second) + epsilon;
            uint32_t prev = 0xFu;
#pragma unroll
            for (uint32_t t = 0; t < 8; ++t)
                if (t == ws) { prices[t] = prices[t] + raise; prev = nib(assign, t); }
            assign = with_nib(assign, ws, c);
            if (prev != 0xFu) bidders = with_nib(bidders, left++, prev);
        }
        epsilon *= 0.125f;
    }
    // children without a slot (bid limit, non-finite costs): first free slots
    for (uint32_t k = 0; k < n; ++k) {
        bool placed = false;
        for (uint32_t t = 0; t < 8; ++t) placed |= nib(assign, t) == k;
        if (placed) continue;
        for (uint32_t t = 0; t < 8; ++t)
            if (nib(assign, t) == 0xFu) { assign = with_nib(assign, t, k); break; }
    }
    return assign;
}

__device__ __forceinline__ uint32_t q_floor(float v) { return ((uint32_t)floorf(v)) & 0xFFu; }
__device__ __forceinline__ uint32_t q_ceil(float v) { return ((uint32_t)ceilf(v)) & 0xFFu; }

// Boxes of the children of the node each lane collapses, in LDS (child t of
// lane i at [t][i]): loaded once during the collapse, read again by the slot
// assignment and the quantization without another global load.
struct ChildBoxes {
    float4 a[8][kBvh8Threads];  // lx, ly, lz, hx
    float2 b[8][kBvh8Threads];  // hy, hz
};
__device__ __forceinline__ void put_child_box(ChildBoxes& cb, uint32_t t, uint32_t tid, const Box& b) {
    cb.a[t][tid] = make_float4(b.lx, b.ly, b.lz, b.hx);
    cb.b[t][tid] = make_float2(b.hy, b.hz);
}
__device__ __forceinline__ Box get_child_box(const ChildBoxes& cb, uint32_t t, uint32_t tid) {
    const float4 a = cb.a[t][tid];
    const float2 c = cb.b[t][tid];
    Box b;
    b.lx = a.x; b.ly = a.y; b.lz = a.z; b.hx = a.w; b.hy = c.x; b.hz = c.y;
    return b;
}

// CreateBVH8Node (WideConverter.cu:156-206) into 20 words; empty slots keep
// all-zero meta and quantization bytes.
__device__ __forceinline__ void write_bvh8_node(nxb_bvh8_node* out, uint32_t k, const Box& pb, const ChildBoxes& cbx,
                                                uint32_t tid, uint32_t child_base, uint32_t prim_base,
                                                uint32_t assign, uint32_t inner_mask, uint32_t leaf_mask) {
    const float qstep = 1.0f / 255.0f;
    uint32_t e0 = ceil_log2((pb.hx - pb.lx) * qstep) & 0xFFu;
    uint32_t e1 = ceil_log2((pb.hy - pb.ly) * qstep) & 0xFFu;
    uint32_t e2 = ceil_log2((pb.hz - pb.lz) * qstep) & 0xFFu;
    float ix = inv_pow2(e0), iy = inv_pow2(e1), iz = inv_pow2(e2);
    uint32_t meta[2] = {0, 0}, qlx[2] = {0, 0}, qly[2] = {0, 0}, qlz[2] = {0, 0};
    uint32_t qhx[2] = {0, 0}, qhy[2] = {0, 0}, qhz[2] = {0, 0};
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        uint32_t a = nib(assign, i);
        if (a == 0xFu) continue;
        uint32_t m;
        if (inner_mask & (1u << i)) m = (1u << 5) | (24 + i);
        else m = (1u << 5) | bits_below(leaf_mask, i);
        const Box cb = get_child_box(cbx, a, tid);
        const uint32_t w = i >> 2, sh = 8 * (i & 3);
        meta[w] |= m << sh;
        qlx[w] |= q_floor((cb.lx - pb.lx) * ix) << sh;
        qly[w] |= q_floor((cb.ly - pb.ly) * iy) << sh;
        qlz[w] |= q_floor((cb.lz - pb.lz) * iz) << sh;
        qhx[w] |= q_ceil((cb.hx - pb.lx) * ix) << sh;
        qhy[w] |= q_ceil((cb.hy - pb.ly) * iy) << sh;
        qhz[w] |= q_ceil((cb.hz - pb.lz) * iz) << sh;
    }
    uint4* o = reinterpret_cast<uint4*>(out + k);
    o[0] = make_uint4(__float_as_uint(pb.lx), __float_as_uint(pb.ly), __float_as_uint(pb.lz),
                      e0 | (e1 << 8) | (e2 << 16) | (inner_mask << 24));
    o[1] = make_uint4(child_base, prim_base, meta[0], meta[1]);
    o[2] = make_uint4(qlx[0], qlx[1], qly[0], qly[1]);
    o[3] = make_uint4(qlz[0], qlz[1], qhx[0], qhx[1]);
    o[4] = make_uint4(qhy[0], qhy[1], qhz[0], qhz[1]);
}

// Collapse + slot assignment of one inner BVH2 node (WideConverter.cu:284-379).
struct Collapsed {
    Box pb;              // the node's box
    uint32_t child[8];   // BVH2 children in push order
    uint32_t cleft[8];   // their left fields (kInvalid for a leaf)
    uint32_t cright[8];  // their right fields (primitive index of a leaf)
    uint32_t count;
    uint32_t assign;     // nibble s = child index in slot s, 0xF = empty
    uint32_t im, lm;     // inner / leaf slot masks
};

// GreedyAssignment (WideConverter.cu:106-153), or the auction, of the
// c.count collapsed children; inner: their inner-child bits by child index.
__device__ __forceinline__ void assign_slots(Collapsed& c, const ChildBoxes& cbx, uint32_t tid, uint32_t slot_mode,
                                             uint32_t inner) {
    const uint32_t count = c.count;
    const Box& pb = c.pb;
    const float pcx = pb.lx + pb.hx, pcy = pb.ly + pb.hy, pcz = pb.lz + pb.hz;
    uint32_t assign = 0xFFFFFFFFu;
    if (slot_mode == NXB_SLOT_AUCTION) {
        float ox[8], oy[8], oz[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            ox[k] = oy[k] = oz[k] = 0.0f;
            if (k < count) {
                const Box cb = get_child_box(cbx, k, tid);
                ox[k] = pcx - (cb.lx + cb.hx);
                oy[k] = pcy - (cb.ly + cb.hy);
                oz[k] = pcz - (cb.lz + cb.hz);
            }
        }
        assign = auction_assign(pb, ox, oy, oz, count);
    }
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        if (slot_mode == NXB_SLOT_AUCTION || k >= count) break;
        const Box cb = get_child_box(cbx, k, tid);
        float ox = pcx - (cb.hx + cb.lx), oy = pcy - (cb.hy + cb.ly), oz = pcz - (cb.hz + cb.lz);
        float best_cost = -3.402823466e+38f;
        uint32_t best = 0xFu, first_free = 0xFu;
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) {
            if (nib(assign, s) != 0xFu) continue;
            if (first_free == 0xFu) first_free = s;
            float cost = slot_cost(s, ox, oy, oz);
            if (cost > best_cost) { best_cost = cost; best = s; }
        }
        if (best == 0xFu) best = first_free;
        assign = (assign & ~(0xFu << (4 * best))) | (k << (4 * best));
    }
    uint32_t im = 0, lm = 0;
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
        uint32_t a = nib(assign, s);
        if (a == 0xFu) continue;
        uint32_t bit = (inner >> a) & 1u;
        im |= bit << s;
        lm |= (bit ^ 1u) << s;
    }
    c.assign = assign;
    c.im = im;
    c.lm = lm;
}

// have_lr: x's children (gl, gr) come with the work item, so their loads do
// not wait for x's (whose box is needed only after the expansion).
__device__ __forceinline__ bool collapse_node(const nxb_bvh2_node* b2, uint32_t node_count, uint32_t x,
                                              Collapsed& c, ChildBoxes& cbx, uint32_t tid, uint32_t slot_mode,
                                              bool have_lr = false, uint32_t gl = 0, uint32_t gr = 0) {
    uint32_t xl, xr;
    load_node(b2, x, c.pb, xl, xr);
    if (have_lr) {
        xl = gl;
        xr = gr;
    }
    if (xl == kInvalid) return false;  // a leaf is never a work item
#pragma unroll
    for (int t = 0; t < 8; ++t) { c.child[t] = 0; c.cright[t] = 0; c.cleft[t] = 0; }
    uint32_t inner = 0, count = 0;
    int msb = 0;
    uint32_t lr0 = xl, lr1 = xr;
    for (;;) {
        if (lr0 >= node_count || lr1 >= node_count) return false;
        Box b0, b1;
        uint32_t l0, r0, l1, r1;
        load_node(b2, lr0, b0, l0, r0);
        load_node(b2, lr1, b1, l1, r1);
        bool first_is_1 = !(box_area(b0) < box_area(b1));
        uint32_t ia = (uint32_t)msb, ib = count + 1;
        uint32_t na = first_is_1 ? lr1 : lr0, nb = first_is_1 ? lr0 : lr1;
        uint32_t la = first_is_1 ? l1 : l0, lb = first_is_1 ? l0 : l1;
        uint32_t ra = first_is_1 ? r1 : r0, rb = first_is_1 ? r0 : r1;
        put_child_box(cbx, ia, tid, first_is_1 ? b1 : b0);
        put_child_box(cbx, ib, tid, first_is_1 ? b0 : b1);
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t) {
            if (t == ia) { c.child[t] = na; c.cleft[t] = la; c.cright[t] = ra; }
            if (t == ib) { c.child[t] = nb; c.cleft[t] = lb; c.cright[t] = rb; }
        }
        if (la != kInvalid) inner |= 1u << ia;
        if (lb != kInvalid) inner |= 1u << ib;
        count += 2;
        msb = inner ? 31 - __clz(inner) : -1;
        if (msb < 0 || count == 8) break;
        inner &= ~(1u << msb);
        --count;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t)
            if (t == (uint32_t)msb) { lr0 = c.cleft[t]; lr1 = c.cright[t]; }
    }
    c.count = count;
    assign_slots(c, cbx, tid, slot_mode, inner);
    return true;
}

// Collapse of internal node x from the wide records of a BVH8 build (record
// x - n: both children's boxes, ids and primitive counts). Same push/pop
// order as collapse_node; each expansion is one 64-byte load, the node's box
// is the union of its children's (as H-PLOC formed it, bit for bit), and the
// children's primitive counts come with them: cleft[t] holds child t's count
// if it is internal (id >= n), kInvalid for a leaf; cright[t] its id (the
// primitive index of a leaf).
__device__ __forceinline__ void load_rec8(const uint4* rec, uint32_t i, Box& l, Box& r, uint32_t& li, uint32_t& ri,
                                          uint32_t& ls, uint32_t& rs) {
    const uint4* p = rec + 4ull * i;
    const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
    l.lx = __uint_as_float(q0.x); l.ly = __uint_as_float(q0.y); l.lz = __uint_as_float(q0.z);
    l.hx = __uint_as_float(q0.w); l.hy = __uint_as_float(q1.x); l.hz = __uint_as_float(q1.y);
    r.lx = __uint_as_float(q1.z); r.ly = __uint_as_float(q1.w); r.lz = __uint_as_float(q2.x);
    r.hx = __uint_as_float(q2.y); r.hy = __uint_as_float(q2.z); r.hz = __uint_as_float(q2.w);
    li = q3.x; ri = q3.y; ls = q3.z; rs = q3.w;
}

__device__ __forceinline__ bool collapse_rec(const uint4* rec, uint32_t n, uint32_t node_count, uint32_t x,
                                             Collapsed& c, ChildBoxes& cbx, uint32_t tid, uint32_t slot_mode) {
    if (x < n || x >= node_count) return false;
#pragma unroll
    for (int t = 0; t < 8; ++t) { c.child[t] = 0; c.cright[t] = 0; c.cleft[t] = 0; }
    uint32_t inner = 0, count = 0;
    int msb = 0;
    uint32_t cur = x;
    for (;;) {
        Box b0, b1;
        uint32_t i0, i1, s0, s1;
        load_rec8(rec, cur - n, b0, b1, i0, i1, s0, s1);
        if (cur == x) c.pb = box_union(b0, b1);
        if (i0 >= node_count || i1 >= node_count) return false;
        bool first_is_1 = !(box_area(b0) < box_area(b1));
        uint32_t ia = (uint32_t)msb, ib = count + 1;
        uint32_t na = first_is_1 ? i1 : i0, nb = first_is_1 ? i0 : i1;
        uint32_t sa = first_is_1 ? s1 : s0, sb = first_is_1 ? s0 : s1;
        put_child_box(cbx, ia, tid, first_is_1 ? b1 : b0);
        put_child_box(cbx, ib, tid, first_is_1 ? b0 : b1);
        const bool ia_inner = na >= n, ib_inner = nb >= n;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t) {
            if (t == ia) { c.child[t] = na; c.cleft[t] = ia_inner ? sa : kInvalid; c.cright[t] = na; }
            if (t == ib) { c.child[t] = nb; c.cleft[t] = ib_inner ? sb : kInvalid; c.cright[t] = nb; }
        }
        if (ia_inner) inner |= 1u << ia;
        if (ib_inner) inner |= 1u << ib;
        count += 2;
        msb = inner ? 31 - __clz(inner) : -1;
        if (msb < 0 || count == 8) break;
        inner &= ~(1u << msb);
        --count;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t)
            if (t == (uint32_t)msb) cur = c.child[t];
    }
    c.count = count;
    assign_slots(c, cbx, tid, slot_mode, inner);
    return true;
}

// Primitive counts of a node's inner children per output slot, from the
// record collapse (0 for leaf and free slots).
__device__ __forceinline__ void rec_child_sizes(const Collapsed& c, uint32_t sz[8]) {
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
        sz[s] = 0;
        const uint32_t a = nib(c.assign, s);
        if (a == 0xFu || !(c.im & (1u << s))) continue;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t)
            if (t == a) sz[s] = c.cleft[t];
    }
}

// Publishes the inner children of BVH8 node k (the first one is kept by this
// lane, depth-first), writes its leaves to primIdx and writes the node.
// Primitive ranges are fixed top-down: the subtree of a node owns a range of
// primIdx as long as its primitive count; the node's own leaves take the
// front, then each inner child (in slot order) takes a range of its size.
// Primitive counts of a node's inner children, per output slot (0 for leaf
// and free slots). The level kernel loads them before the block allocates
// the child ids, so that their latency overlaps that allocation.
__device__ __forceinline__ void load_child_sizes(const Collapsed& c, const uint32_t* nsize, uint32_t n,
                                                 uint32_t node_count, uint32_t sz[8]) {
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
        sz[s] = 0;
        const uint32_t a = nib(c.assign, s);
        if (a == 0xFu || !(c.im & (1u << s))) continue;
        uint32_t ch = 0;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t)
            if (t == a) ch = c.child[t];
        if (ch >= n && ch < node_count) sz[s] = nsize[ch - n];
    }
}

template <bool KEEP = true>
__device__ __forceinline__ void emit_node(const ChildBoxes& cbx, uint32_t tid, const Collapsed& c, uint32_t k,
                                          uint32_t child_base, uint32_t prim_base, uint32_t leaf_count,
                                          const uint32_t sz[8], nxb_bvh8_node* out, uint32_t* prim_idx,
                                          uint32_t n, uint32_t node_count, uint32_t capacity, uint64_t* work,
                                          uint64_t* work2, uint32_t* flags,
                                          uint32_t& keep_slot, uint32_t& keep_x, uint32_t& keep_base) {
    keep_slot = kNoSlot;
    keep_x = kInvalid;
    keep_base = 0;
    uint32_t next_base = prim_base + leaf_count;
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
        uint32_t a = nib(c.assign, s);
        if (a == 0xFu) continue;
        uint32_t ch = 0, pl = 0, pr = 0;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t)
            if (t == a) { ch = c.child[t]; pl = c.cleft[t]; pr = c.cright[t]; }
        if (c.im & (1u << s)) {
            uint32_t slot = child_base + bits_below(c.im, s);
            const uint32_t size = sz[s];
            if (slot >= capacity || size < 2 || next_base + size > n) {
                atomic_or_agent(flags, kFlagBadIndex);
            } else if (!KEEP) {
                // read by the next level's kernel
                work[slot] = ((uint64_t)next_base << 32) | ch;
                if (work2) work2[slot] = ((uint64_t)pr << 32) | pl;
            } else if (keep_slot == kNoSlot) {
                keep_slot = slot;
                keep_x = ch;
                keep_base = next_base;
                store_agent64(&work[slot], ((uint64_t)next_base << 32) | kTaken);
            } else {
                store_agent64(&work[slot], ((uint64_t)next_base << 32) | ch);
            }
            next_base += size;
        } else {
            uint32_t p = prim_base + bits_below(c.lm, s);
            if (p < n) prim_idx[p] = pr;
            else atomic_or_agent(flags, kFlagBadIndex);
        }
    }
    write_bvh8_node(out, k, c.pb, cbx, tid, child_base, leaf_count ? prim_base : 0u, c.assign, c.im, c.lm);
}

// Inclusive sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(v, off);
        if (lane >= (uint32_t)off) v += y;
    }
    return v;
}

// n leaves; node_count = 2n - 1 for one BVH2, 2n - (trees) for a forest
// whose roots seed work[0..trees).
template <uint32_t SLOT>
__global__ __launch_bounds__(kBvh8Threads) void k_bvh8(const nxb_bvh2_node* __restrict__ b2, uint32_t n,
                                                       uint32_t node_count, nxb_bvh8_node* __restrict__ out,
                                                       uint32_t* prim_idx, uint32_t capacity, uint64_t* work,
                                                       unsigned long long* ctr, uint32_t* grab,
                                                       const uint32_t* __restrict__ nsize, uint32_t* flags,
                                                       uint32_t spin_limit) {
    __shared__ ChildBoxes cbx;
    const uint32_t tid = threadIdx.x;
    if (n == 1) {
        // CreateBVH8SingleLeaf (WideConverter.cu:209-222).
        if (blockIdx.x == 0 && tid == 0) {
            Box pb = load_box(b2, 0);
            put_child_box(cbx, 0, tid, pb);
            prim_idx[0] = 0;
            write_bvh8_node(out, 0, pb, cbx, tid, 0, 0, 0xFFFFFFF0u, 0u, 1u);
        }
        return;
    }
    uint32_t slot = kNoSlot;  // BVH8 node this lane builds next
    uint32_t x = kInvalid;    // its BVH2 node, once known
    uint32_t pbase = 0;       // first primIdx entry of its subtree
    bool done = false;
    uint32_t spins = 0, idle = 0, witer = 0;
    for (;;) {
        if (!ballot(!done)) break;
        bool need = !done && slot == kNoSlot;
        uint64_t nm = ballot(need);
        if (nm) {
            uint32_t leader = (uint32_t)__builtin_ctzll(nm);
            uint32_t base = 0;
            if (lane_id() == leader) base = atomic_add_agent(grab, (uint32_t)__popcll(nm));
            base = __shfl(base, (int)leader);
            if (need) slot = base + mask_rank(nm);
        }
        bool waiting = false;
        if (!done && x == kInvalid) {
            if (slot >= capacity) {
                done = true;
            } else {
                uint64_t v = load_agent64(&work[slot]);
                uint32_t lo = (uint32_t)v;
                if (lo == kTaken) {
                    slot = kNoSlot;  // its producer builds it
                    spins = 0;
                } else if (lo != kInvalid) {
                    x = lo;
                    pbase = (uint32_t)(v >> 32);
                } else {
                    waiting = true;
                    if (++spins >= spin_limit) {
                        atomic_or_agent(flags, kFlagTimeout);
                        done = true;
                        waiting = false;
                    }
                }
            }
        }
        // Completion: every allocated node finished (one load per wave).
        ++witer;
        if ((witer & 7u) == 0 && ballot(waiting)) {
            unsigned long long cv = 0;
            if (lane_id() == 0) cv = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t lo = __shfl((uint32_t)cv, 0), hi = __shfl((uint32_t)(cv >> 32), 0);
            if (waiting && lo == hi) done = true;
        }
        bool have = !done && x != kInvalid;
        if (!ballot(have)) {
            if (!ballot(!done && slot == kNoSlot)) {
                // the whole wave waits on unpublished nodes
                if (++idle > 8) __builtin_amdgcn_s_sleep(24);
                else __builtin_amdgcn_s_sleep(2);
            }
            continue;
        }
        idle = 0;
        Collapsed c;
        bool ok = false;
        if (have) {
            spins = 0;
            ok = x < node_count && collapse_node(b2, node_count, x, c, cbx, tid, SLOT);
            if (!ok) atomic_or_agent(flags, kFlagBadIndex);
        }
        // One atomic per wave iteration: allocate the inner children of all
        // lanes and count their nodes as finished.
        uint32_t ic = ok ? (uint32_t)__popc(c.im) : 0u;
        uint32_t packed = (ic << 16) | (have ? 1u : 0u);
        uint32_t incl = wave_inclusive_sum(packed);
        uint32_t total = __shfl(incl, 63);
        uint32_t excl = incl - packed;
        unsigned long long old = 0;
        if (lane_id() == 0)
            old = __hip_atomic_fetch_add(ctr, ((unsigned long long)(total & 0xFFFFu) << 32) | (total >> 16),
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t node_base = __shfl((uint32_t)old, 0);
        if (have) {
            uint32_t ks = kNoSlot, kx = kInvalid, kb = 0;
            if (ok) {
                uint32_t child_base = ic ? node_base + (excl >> 16) : 0u;
                uint32_t sz[8];
                load_child_sizes(c, nsize, n, node_count, sz);
                emit_node(cbx, tid, c, slot, child_base, pbase, c.count - ic, sz, out, prim_idx, n, node_count,
                          capacity, work, nullptr, flags, ks, kx, kb);
            }
            slot = ks;
            x = kx;
            pbase = kb;
        }
    }
}

// ---- level-synchronous construction (single builds) -------------------------
//
// One kernel per BVH8 level: every node of the level is collapsed by one
// lane; a block allocates the inner children of all its lanes with one
// atomic add on the next level's counter (block scan of the counts). Level
// l holds the ids [base_l, base_l + cnt[l]) with base_l = cnt[0] + ... +
// cnt[l-1], so a level's kernel finds its range from counters that earlier
// kernels completed, and no lane ever waits on another (no polling, no spin
// limit). The total node count is the sum of the counters.
//
// Inputs: a whole build (one mesh) collapses the wide records its H-PLOC
// wrote instead of BVH2 nodes: one aligned 64-byte load per expansion (both
// children's boxes, ids and primitive counts) replaces two 32-byte node
// loads (leaves in input order, a line each) and the size lookups. At 28.1M
// triangles this took the collapse from 1.94 to 1.31 ms. The stage entry
// point and batched forests collapse BVH2 nodes.
constexpr uint32_t kMaxLevels = kBvh8LevelWords - 1;  // BVH8 levels; each consumes at least one BVH2 level

// Exclusive scan over a 256-thread block; *total gets the block sum.
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63) tmp[w] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBvh8Threads / 64; ++k) {
        const uint32_t t = tmp[k];
        if (k < w) before += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

// The top levels of one tree in one block: the level counters initialized
// (cnt[0] = 1, the root; the rest 0), then levels 0..kBvh8TopLevels-1 (at most 1, 8 and 64
// nodes) one after another, the children's ids taken from a running count
// instead of a global atomic. Saves the launches and counter round trips of
// the levels whose time is all latency.
// REC: collapse from the wide records (rec) instead of the BVH2 nodes; no
// nsize, no work2.
template <uint32_t SLOT, bool REC>
__global__ __launch_bounds__(kBvh8Threads) void k_bvh8_top(const nxb_bvh2_node* __restrict__ b2, uint32_t n,
                                                           uint32_t node_count, nxb_bvh8_node* __restrict__ out,
                                                           uint32_t* prim_idx, uint32_t capacity, uint64_t* work,
                                                           uint64_t* work2, uint32_t* cnt,
                                                           const uint32_t* __restrict__ nsize, uint32_t* flags,
                                                           const uint4* __restrict__ rec) {
    __shared__ ChildBoxes cbx;
    __shared__ uint32_t scan_tmp[kBvh8Threads / 64];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i <= kMaxLevels; i += kBvh8Threads) cnt[i] = i == 0 ? 1u : 0u;
    uint32_t lo = 0, m = 1;
    for (uint32_t level = 0; level < kBvh8TopLevels && m != 0; ++level) {
        const uint32_t hi = lo + m;
        uint32_t next = 0;
        for (uint32_t k0 = lo; k0 < hi; k0 += kBvh8Threads) {
            const uint32_t k = k0 + tid;
            Collapsed c;
            bool ok = false;
            uint32_t pbase = 0;
            if (k < hi) {
                uint32_t x = node_count - 1;  // the root, primitive base 0
                uint64_t lr = 0;
                if (level > 0) {
                    // written by this block's previous level (agent-scope
                    // loads: not through the CU's L1)
                    const uint64_t v = load_agent64(&work[k]);
                    if constexpr (!REC) lr = load_agent64(&work2[k]);
                    x = (uint32_t)v;
                    pbase = (uint32_t)(v >> 32);
                }
                if constexpr (REC)
                    ok = collapse_rec(rec, n, node_count, x, c, cbx, tid, SLOT);
                else
                    ok = x < node_count && collapse_node(b2, node_count, x, c, cbx, tid, SLOT, level > 0,
                                                         (uint32_t)lr, (uint32_t)(lr >> 32));
                if (!ok) atomic_or_agent(flags, kFlagBadIndex);
            }
            const uint32_t ic = ok ? (uint32_t)__popc(c.im) : 0u;
            uint32_t sz[8];
            if (ok) {
                if constexpr (REC) rec_child_sizes(c, sz);
                else load_child_sizes(c, nsize, n, node_count, sz);
            }
            uint32_t total = 0;
            const uint32_t excl = block_scan_256(ic, scan_tmp, &total);
            if (ok) {
                uint32_t ks, kx, kb;
                emit_node<false>(cbx, tid, c, k, ic ? hi + next + excl : 0u, pbase, c.count - ic, sz, out, prim_idx,
                                 n, node_count, capacity, work, REC ? nullptr : work2, flags, ks, kx, kb);
            }
            next += total;
        }
        if (tid == 0) cnt[level + 1] = next;
        __syncthreads();  // the level's work items are written before the next level reads them
        lo = hi;
        m = next;
    }
}

template <uint32_t SLOT, bool REC>
__global__ __launch_bounds__(kBvh8Threads) void k_bvh8_level(const nxb_bvh2_node* __restrict__ b2, uint32_t n,
                                                             uint32_t node_count, nxb_bvh8_node* __restrict__ out,
                                                             uint32_t* prim_idx, uint32_t capacity, uint64_t* work,
                                                             uint64_t* work2, uint32_t* cnt, uint32_t level,
                                                             const uint32_t* __restrict__ nsize, uint32_t* flags,
                                                             const uint4* __restrict__ rec) {
    __shared__ ChildBoxes cbx;
    __shared__ uint32_t scan_tmp[kBvh8Threads / 64];
    __shared__ uint32_t sbase;
    const uint32_t tid = threadIdx.x;
    const uint32_t m = cnt[level];
    if (m == 0) return;
    uint32_t lo = 0;
    for (uint32_t j = 0; j < level; ++j) lo += cnt[j];
    const uint32_t hi = lo + m;
    if (level + 1 > kMaxLevels) {
        if (blockIdx.x == 0 && tid == 0) atomic_or_agent(flags, kFlagBadIndex);
        return;
    }
    for (uint32_t k0 = lo + blockIdx.x * kBvh8Threads; k0 < hi; k0 += gridDim.x * kBvh8Threads) {
        const uint32_t k = k0 + tid;
        Collapsed c;
        bool ok = false;
        uint32_t pbase = 0;
        if (k < hi) {
            const uint64_t v = work[k];
            const uint32_t x = (uint32_t)v;
            pbase = (uint32_t)(v >> 32);
            if constexpr (REC) {
                ok = collapse_rec(rec, n, node_count, x, c, cbx, tid, SLOT);
            } else {
                const bool have_lr = level > 0;  // written with work[k] by the level above
                const uint64_t lr = have_lr ? work2[k] : 0ull;
                ok = x < node_count && collapse_node(b2, node_count, x, c, cbx, tid, SLOT, have_lr, (uint32_t)lr,
                                                     (uint32_t)(lr >> 32));
            }
            if (!ok) atomic_or_agent(flags, kFlagBadIndex);
        }
        const uint32_t ic = ok ? (uint32_t)__popc(c.im) : 0u;
        uint32_t sz[8];
        if (ok) {
            if constexpr (REC) rec_child_sizes(c, sz);
            else load_child_sizes(c, nsize, n, node_count, sz);
        }
        uint32_t total = 0;
        const uint32_t excl = block_scan_256(ic, scan_tmp, &total);
        if (tid == 0) sbase = total ? hi + atomic_add_agent(&cnt[level + 1], total) : 0u;
        __syncthreads();
        const uint32_t base = sbase;
        if (ok) {
            uint32_t ks, kx, kb;
            emit_node<false>(cbx, tid, c, k, ic ? base + excl : 0u, pbase, c.count - ic, sz, out, prim_idx, n,
                             node_count, capacity, work, REC ? nullptr : work2, flags, ks, kx, kb);
        }
    }
}

// `count` consecutive values per lane from one counter per tree (node ids,
// primIdx slots): each run of consecutive lanes with the same tree takes its
// values with one atomic (issued by the run's last lane) and every lane gets
// the first of its `count`. Lanes with tree == kInvalid take none. Whole wave
// only.
__device__ __forceinline__ uint32_t alloc_tree_ids(uint32_t* ctr, uint32_t tree, uint32_t count) {
    const uint32_t lane = lane_id();
    const uint32_t prev = __shfl_up(tree, 1);
    const uint64_t heads = ballot(lane == 0 || prev != tree);
    const uint32_t first = 63u - (uint32_t)__clzll((long long)(heads & (~0ull >> (63u - lane))));
    uint32_t v = count;  // inclusive sum over the run
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off);
        if (lane >= first + off) v += t;
    }
    const bool tail = lane == 63u || ((heads >> (lane + 1u)) & 1ull);
    const uint64_t tails = ballot(tail);
    const uint32_t last = (uint32_t)__ffsll((long long)(tails & (~0ull << lane))) - 1u;
    uint32_t base = 0;
    if (tail && tree != kInvalid && v != 0) base = atomic_add_agent(&ctr[tree], v);
    return __shfl(base, last) + v - count;
}

// Work items of one block's top levels (k_bvh8_tree_top), in LDS: BVH2
// node, the node's children, tree and node id.
constexpr uint32_t kTopQueue = kBvh8Threads;
struct TopQueue {
    uint32_t x[kTopQueue], l[kTopQueue], r[kTopQueue], tree[kTopQueue], self[kTopQueue];
};

// One BVH8 node of tree `tree` (batched builds): node `self` of the tree's
// region out, children's work items at queue positions qbase.., their node
// ids nbase.. (tree-local), its leaves at primIdx slots prim_base.. (local to
// the tree; each node takes its own from the tree's counter, so no subtree
// sizes are read). The children go to the level queue in work/work2/work3
// (positions below capacity), or with lq to a block's LDS queue (positions
// below kTopQueue).
// ti = {first primitive, primitive count, region start, region capacity}.
__device__ __forceinline__ void emit_tree_node(const ChildBoxes& cbx, uint32_t tid, const Collapsed& c,
                                               uint32_t tree, uint32_t self, uint32_t qbase, uint32_t nbase,
                                               uint32_t prim_base, uint4 ti, nxb_bvh8_node* out,
                                               uint32_t* prim_idx, uint32_t capacity, uint64_t* work,
                                               uint64_t* work2, uint64_t* work3, uint32_t* flags,
                                               TopQueue* lq = nullptr) {
    const uint32_t ic = (uint32_t)__popc(c.im);
    const uint32_t leaf_count = c.count - ic;
    if (self >= ti.w || (ic != 0 && nbase + ic > ti.w) || (leaf_count != 0 && prim_base + leaf_count > ti.y)) {
        atomic_or_agent(flags, kFlagBadIndex);
        return;
    }
    bool bad = false;
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
        const uint32_t a = nib(c.assign, s);
        if (a == 0xFu) continue;
        uint32_t ch = 0, pl = 0, pr = 0;
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t)
            if (t == a) { ch = c.child[t]; pl = c.cleft[t]; pr = c.cright[t]; }
        if (c.im & (1u << s)) {
            const uint32_t rank = bits_below(c.im, s);
            const uint32_t q = qbase + rank;
            if (q >= (lq ? kTopQueue : capacity)) {
                bad = true;
            } else if (lq) {
                // read by this block's next level
                lq->x[q] = ch;
                lq->l[q] = pl;
                lq->r[q] = pr;
                lq->tree[q] = tree;
                lq->self[q] = nbase + rank;
            } else {
                // read by the next level's kernel
                work[q] = ch;
                work2[q] = ((uint64_t)pr << 32) | pl;
                work3[q] = ((uint64_t)tree << 32) | (nbase + rank);
            }
        } else {
            const uint32_t p = prim_base + bits_below(c.lm, s);
            if (p < ti.y) prim_idx[ti.x + p] = pr;
            else bad = true;
        }
    }
    if (bad) atomic_or_agent(flags, kFlagBadIndex);
    write_bvh8_node(out + ti.z, self, c.pb, cbx, tid, ic ? nbase : 0u, leaf_count ? prim_base : 0u, c.assign, c.im,
                    c.lm);
}

// A level of a forest whose trees each own an output region (batched
// builds). Queue position k (level order, as in k_bvh8_level) holds work[k] =
// its BVH2 node, work2[k] = that node's children and work3[k] = (tree << 32 |
// tree-local node id); tree t writes its nodes to out + info[t].z and its
// primitives to prim_idx + info[t].x, so each tree's BVH8 is contiguous with
// root 0 and needs no compaction afterwards. Node ids come from ids[t] and
// each node's primIdx slots for its leaves from prims[t].
template <uint32_t SLOT>
__global__ __launch_bounds__(kBvh8Threads) void k_bvh8_tree_level(
    const nxb_bvh2_node* __restrict__ b2, uint32_t node_count, nxb_bvh8_node* __restrict__ out, uint32_t* prim_idx,
    uint32_t capacity, uint64_t* work, uint64_t* work2, uint64_t* work3, uint32_t* cnt, uint32_t level,
    const uint4* __restrict__ info, uint32_t trees, uint32_t* ids, uint32_t* prims, uint32_t* flags) {
    __shared__ ChildBoxes cbx;
    __shared__ uint32_t scan_tmp[kBvh8Threads / 64];
    __shared__ uint32_t sbase;
    const uint32_t tid = threadIdx.x;
    const uint32_t m = cnt[level];
    if (m == 0) return;
    uint32_t lo = 0;
    for (uint32_t j = 0; j < level; ++j) lo += cnt[j];
    const uint32_t hi = lo + m;
    if (level + 1 > kMaxLevels) {
        if (blockIdx.x == 0 && tid == 0) atomic_or_agent(flags, kFlagBadIndex);
        return;
    }
    for (uint32_t k0 = lo + blockIdx.x * kBvh8Threads; k0 < hi; k0 += gridDim.x * kBvh8Threads) {
        const uint32_t k = k0 + tid;
        Collapsed c;
        bool ok = false;
        uint32_t tree = kInvalid, self = 0;
        uint4 ti = make_uint4(0, 0, 0, 0);
        if (k < hi && k < capacity) {
            const uint64_t v = work[k], t3 = work3[k];
            const uint32_t x = (uint32_t)v;
            self = (uint32_t)t3;
            const uint32_t t = (uint32_t)(t3 >> 32);
            ti = info[min(t, trees - 1u)];
            const bool have_lr = level > 0;  // written with work[k] by the level above
            const uint64_t lr = have_lr ? work2[k] : 0ull;
            ok = t < trees && x < node_count &&
                 collapse_node(b2, node_count, x, c, cbx, tid, SLOT, have_lr, (uint32_t)lr, (uint32_t)(lr >> 32));
            if (ok) tree = t;
        }
        if (k < hi && !ok) atomic_or_agent(flags, kFlagBadIndex);
        const uint32_t ic = ok ? (uint32_t)__popc(c.im) : 0u;
        const uint32_t leaves = ok ? c.count - ic : 0u;
        uint32_t total = 0;
        const uint32_t excl = block_scan_256(ic, scan_tmp, &total);
        // lanes without inner children or leaves stay in their tree's run
        // (count 0), so a wave whose nodes lie in one tree takes its ids and
        // its primIdx slots with one atomic each
        const uint32_t nbase = alloc_tree_ids(ids, tree, ic);
        const uint32_t pbase = alloc_tree_ids(prims, tree, leaves);
        if (tid == 0) sbase = total ? hi + atomic_add_agent(&cnt[level + 1], total) : 0u;
        __syncthreads();
        const uint32_t qbase = sbase + excl;
        if (ok)
            emit_tree_node(cbx, tid, c, tree, self, qbase, nbase, pbase, ti, out, prim_idx, capacity, work, work2,
                           work3, flags);
    }
}

// Levels 0..kTreeTopLevels-1 of a forest with per-tree regions, one block
// per kTopTrees trees: at most 4 roots, 32 level-1 and 256 level-2 nodes,
// so each level is one pass of the block's lanes. The block seeds its roots,
// keeps the work items of levels 1 and 2 in LDS and numbers those levels'
// nodes with LDS counters (no global atomics, no launches between these
// levels), then appends level 3's work items to the level queue from
// position 0 (cnt[0..2] stay 0, cnt[3] counts them) and hands its trees'
// node and primIdx counts to ids[] and prims[] for the level kernels.
constexpr uint32_t kTopTrees = 4;

template <uint32_t SLOT>
__global__ __launch_bounds__(kBvh8Threads) void k_bvh8_tree_top(
    const nxb_bvh2_node* __restrict__ b2, uint32_t n, uint32_t node_count, nxb_bvh8_node* __restrict__ out,
    uint32_t* prim_idx, uint32_t capacity, uint64_t* work, uint64_t* work2, uint64_t* work3, uint32_t* cnt,
    const uint4* __restrict__ info, uint32_t trees, const uint32_t* __restrict__ mesh_lo, uint32_t* ids,
    uint32_t* prims, uint32_t* flags) {
    __shared__ ChildBoxes cbx;
    __shared__ TopQueue tq[2];
    __shared__ uint32_t scan_tmp[kBvh8Threads / 64];
    __shared__ uint32_t s_ids[kTopTrees], s_prims[kTopTrees];
    __shared__ uint32_t s_cnt[2];
    __shared__ uint32_t sbase;
    const uint32_t tid = threadIdx.x;
    const uint32_t t0 = blockIdx.x * kTopTrees;
    if (t0 >= trees) return;
    const uint32_t nt = min(kTopTrees, trees - t0);
    if (tid < nt) {
        // the root of tree t is its last internal node, n + lo - t + cnt - 2
        const uint32_t t = t0 + tid;
        const uint32_t lo = mesh_lo[t], c = mesh_lo[t + 1] - lo;
        tq[0].x[tid] = n + lo - t + c - 2;
        tq[0].l[tid] = 0;
        tq[0].r[tid] = 0;
        tq[0].tree[tid] = t;
        tq[0].self[tid] = 0;
        s_ids[tid] = 1;
        s_prims[tid] = 0;
    }
    if (tid == 0) s_cnt[0] = nt;
    __syncthreads();
    for (uint32_t level = 0; level < kTreeTopLevels; ++level) {
        const uint32_t cur = level & 1u;
        const uint32_t m = s_cnt[cur];
        const bool last = level + 1 == kTreeTopLevels;
        Collapsed c;
        bool ok = false;
        uint32_t tree = kInvalid, self = 0;
        uint4 ti = make_uint4(0, 0, 0, 0);
        if (tid < m) {
            const uint32_t x = tq[cur].x[tid], t = tq[cur].tree[tid];
            self = tq[cur].self[tid];
            ti = info[min(t, trees - 1u)];
            ok = t >= t0 && t < t0 + nt && x < node_count &&
                 collapse_node(b2, node_count, x, c, cbx, tid, SLOT, level > 0, tq[cur].l[tid], tq[cur].r[tid]);
            if (ok) tree = t;
            else atomic_or_agent(flags, kFlagBadIndex);
        }
        const uint32_t ic = ok ? (uint32_t)__popc(c.im) : 0u;
        const uint32_t leaves = ok ? c.count - ic : 0u;
        uint32_t total = 0;
        const uint32_t excl = block_scan_256(ic, scan_tmp, &total);
        uint32_t nbase = 0, pbase = 0;
        if (ic) nbase = atomicAdd(&s_ids[tree - t0], ic);
        if (leaves) pbase = atomicAdd(&s_prims[tree - t0], leaves);
        if (tid == 0) {
            if (last) sbase = total ? atomic_add_agent(&cnt[kTreeTopLevels], total) : 0u;
            else s_cnt[cur ^ 1u] = total;
        }
        __syncthreads();  // this level's queue is read; sbase and the next count are set
        if (ok)
            emit_tree_node(cbx, tid, c, tree, self, last ? sbase + excl : excl, nbase, pbase, ti, out, prim_idx,
                           capacity, work, work2, work3, flags, last ? nullptr : &tq[cur ^ 1u]);
        __syncthreads();  // the next level's queue is written
    }
    if (tid < nt) {
        ids[t0 + tid] = s_ids[tid];
        prims[t0 + tid] = s_prims[tid];
    }
}

// Primitive counts of the internal nodes of any BVH2: parents first, then a
// bottom-up climb where the second child to arrive at a node sums the two
// counts (distinct-address atomics only).
__global__ void k_parents(const nxb_bvh2_node* __restrict__ b2, uint32_t n, uint32_t* parent, uint32_t* arrive,
                          uint32_t* flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) parent[2 * n - 2] = kInvalid;
    if (i < n - 1) {
        uint32_t x = n + i;
        arrive[i] = 0;
        uint32_t l = b2[x].left, r = b2[x].right;
        if (l < 2 * n - 1 && r < 2 * n - 1) {
            parent[l] = x;
            parent[r] = x;
        } else {
            atomic_or_agent(flags, kFlagBadIndex);
        }
    }
}

__global__ void k_sizes(const nxb_bvh2_node* __restrict__ b2, uint32_t n, const uint32_t* __restrict__ parent,
                        uint32_t* arrive, uint32_t* nsize) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = parent[i];
    for (uint32_t guard = 0; p != kInvalid && p >= n && p < 2 * n - 1 && guard < n; ++guard) {
        wait_memory_ops();
        if (atomic_add_agent(&arrive[p - n], 1u) == 0) return;  // first arrival: the sibling finishes p
        compiler_barrier();
        uint32_t l = b2[p].left, r = b2[p].right;
        uint32_t sl = l < n ? 1u : load_agent(&nsize[l - n]);
        uint32_t sr = r < n ? 1u : load_agent(&nsize[r - n]);
        store_agent(&nsize[p - n], sl + sr);
        p = parent[p];
    }
}

hipError_t launch_subtree_sizes(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t* parent, uint32_t* arrive,
                                uint32_t* nsize, uint32_t* flags, hipStream_t s) {
    if (n < 2) return hipSuccess;
    uint32_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_parents, dim3(g), dim3(256), 0, s, bvh2, n, parent, arrive, flags);
    hipLaunchKernelGGL(k_sizes, dim3(g), dim3(256), 0, s, bvh2, n, parent, arrive, nsize);
    return hipGetLastError();
}

__global__ void k_bvh8_init(unsigned long long* ctr, uint32_t* grab, uint64_t* work, uint32_t root) {
    *ctr = 1;   // one node allocated (the root), none finished
    *grab = 0;
    work[0] = root;  // primitive base 0
}

hipError_t launch_bvh8_prep(uint32_t n, uint32_t capacity, Bvh8Scratch sc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(sc.work), kInvalid, 2ull * capacity, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bvh8_init, dim3(1), dim3(1), 0, s, sc.ctr, sc.grab, sc.work, 2 * n - 2);
    return hipGetLastError();
}

hipError_t launch_bvh8_run(const nxb_bvh2_node* bvh2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                           uint32_t capacity, uint32_t spin_limit, Bvh8Scratch sc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // Fewer, busier waves: lanes of a wave that build nodes in the same
    // iteration share one counter atomic.
    uint32_t lanes = capacity / 16;
    lanes = max(lanes, 4096u);
    lanes = min(lanes, device_cu_count() * 512u);
    lanes = min(lanes, capacity);
    uint32_t grid = (lanes + kBvh8Threads - 1) / kBvh8Threads;
    if (n == 1) grid = 1;
    hipLaunchKernelGGL(sc.slot_mode == NXB_SLOT_AUCTION ? k_bvh8<NXB_SLOT_AUCTION> : k_bvh8<NXB_SLOT_GREEDY>,
                       dim3(grid), dim3(kBvh8Threads), 0, s, bvh2, n, 2 * n - 1, out, prim_idx, capacity, sc.work,
                       sc.ctr, sc.grab, sc.nsize, sc.flags, spin_limit);
    return hipGetLastError();
}

hipError_t launch_bvh8_forest(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t trees, nxb_bvh8_node* out,
                              uint32_t* prim_idx, uint32_t capacity, uint32_t lanes, uint32_t spin_limit,
                              Bvh8Scratch sc, hipStream_t s) {
    if (n < 2) return hipErrorInvalidValue;
    lanes = min(max(lanes, 256u), capacity);
    uint32_t grid = (lanes + kBvh8Threads - 1) / kBvh8Threads;
    hipLaunchKernelGGL(sc.slot_mode == NXB_SLOT_AUCTION ? k_bvh8<NXB_SLOT_AUCTION> : k_bvh8<NXB_SLOT_GREEDY>,
                       dim3(grid), dim3(kBvh8Threads), 0, s, bvh2, n, 2 * n - trees, out, prim_idx, capacity,
                       sc.work, sc.ctr, sc.grab, sc.nsize, sc.flags, spin_limit);
    return hipGetLastError();
}

hipError_t launch_bvh8_levels(const nxb_bvh2_node* bvh2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                              uint32_t capacity, Bvh8Scratch sc, uint32_t* cnt, uint32_t from, uint32_t count,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (n == 1) {
        if (from == 0)
            hipLaunchKernelGGL(k_bvh8<NXB_SLOT_GREEDY>, dim3(1), dim3(kBvh8Threads), 0, s, bvh2, n, 1u, out,
                               prim_idx, capacity, sc.work, sc.ctr, sc.grab, sc.nsize, sc.flags, 1u);
        return hipGetLastError();
    }
    const bool rec = sc.rec8 != nullptr;
    if (!rec && !sc.work2) return hipErrorInvalidValue;
    const bool auction = sc.slot_mode == NXB_SLOT_AUCTION;
    // three 48 KB blocks per CU; a level of m nodes uses ceil(m / 256) blocks
    const uint32_t grid = min((capacity + kBvh8Threads - 1) / kBvh8Threads, device_cu_count() * 3u);
    if (from == 0) {
        auto top = auction ? (rec ? k_bvh8_top<NXB_SLOT_AUCTION, true> : k_bvh8_top<NXB_SLOT_AUCTION, false>)
                           : (rec ? k_bvh8_top<NXB_SLOT_GREEDY, true> : k_bvh8_top<NXB_SLOT_GREEDY, false>);
        hipLaunchKernelGGL(top, dim3(1), dim3(kBvh8Threads), 0, s, bvh2, n, 2 * n - 1, out, prim_idx, capacity,
                           sc.work, sc.work2, cnt, sc.nsize, sc.flags, sc.rec8);
        from = kBvh8TopLevels;
        count = count > kBvh8TopLevels ? count - kBvh8TopLevels : 0u;
    }
    auto level = auction ? (rec ? k_bvh8_level<NXB_SLOT_AUCTION, true> : k_bvh8_level<NXB_SLOT_AUCTION, false>)
                         : (rec ? k_bvh8_level<NXB_SLOT_GREEDY, true> : k_bvh8_level<NXB_SLOT_GREEDY, false>);
    const uint32_t end = min(from + count, kMaxLevels);
    for (uint32_t l = from; l < end; ++l)
        hipLaunchKernelGGL(level, dim3(grid), dim3(kBvh8Threads), 0, s, bvh2, n, 2 * n - 1, out, prim_idx, capacity,
                           sc.work, sc.work2, cnt, l, sc.nsize, sc.flags, sc.rec8);
    return hipGetLastError();
}

// Forest of `trees` BVH2 trees whose roots the caller put in work[0..trees).
__global__ void k_bvh8_forest_level_init(uint32_t* cnt, uint32_t trees) {
    for (uint32_t i = threadIdx.x; i <= kMaxLevels; i += blockDim.x) cnt[i] = i == 0 ? trees : 0u;
}

// Node total into the low word of *ctr; a non-empty level past the last one
// launched means the forest was deeper than the launch covered
// (kFlagDeeper: the caller launches again with more levels).
__global__ void k_bvh8_level_total(const uint32_t* cnt, uint32_t levels, unsigned long long* ctr, uint32_t* flags) {
    uint32_t total = 0;
    for (uint32_t l = 0; l < levels; ++l) total += cnt[l];
    *ctr = total;
    if (cnt[levels] != 0) atomic_or_agent(flags, kFlagDeeper);
}

hipError_t launch_bvh8_forest_levels(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t trees, nxb_bvh8_node* out,
                                     uint32_t* prim_idx, uint32_t capacity, uint32_t levels, Bvh8Scratch sc,
                                     uint32_t* cnt, hipStream_t s) {
    if (n < 2 || levels == 0 || levels > kMaxLevels || !sc.work2) return hipErrorInvalidValue;
    const uint32_t grid = min((capacity + kBvh8Threads - 1) / kBvh8Threads, device_cu_count() * 3u);
    hipLaunchKernelGGL(k_bvh8_forest_level_init, dim3(1), dim3(128), 0, s, cnt, trees);
    auto level = sc.slot_mode == NXB_SLOT_AUCTION ? k_bvh8_level<NXB_SLOT_AUCTION, false>
                                                  : k_bvh8_level<NXB_SLOT_GREEDY, false>;
    for (uint32_t l = 0; l < levels; ++l)
        hipLaunchKernelGGL(level, dim3(grid), dim3(kBvh8Threads), 0, s, bvh2, n, 2 * n - trees, out, prim_idx,
                           capacity, sc.work, sc.work2, cnt, l, sc.nsize, sc.flags, static_cast<const uint4*>(nullptr));
    hipLaunchKernelGGL(k_bvh8_level_total, dim3(1), dim3(1), 0, s, cnt, levels, sc.ctr, sc.flags);
    return hipGetLastError();
}

hipError_t launch_bvh8_tree_levels(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t trees, nxb_bvh8_node* out,
                                   uint32_t* prim_idx, uint32_t capacity, uint32_t levels, Bvh8Scratch sc,
                                   uint32_t* cnt, uint64_t* work3, const uint4* info, const uint32_t* mesh_lo,
                                   uint32_t* ids, uint32_t* prims, hipStream_t s) {
    if (n < 2 || trees == 0 || levels < kTreeTopLevels || levels > kMaxLevels || !sc.work2 || !work3 || !info ||
        !mesh_lo || !ids || !prims)
        return hipErrorInvalidValue;
    const uint32_t grid = min((capacity + kBvh8Threads - 1) / kBvh8Threads, device_cu_count() * 3u);
    const bool auction = sc.slot_mode == NXB_SLOT_AUCTION;
    hipLaunchKernelGGL(k_bvh8_forest_level_init, dim3(1), dim3(128), 0, s, cnt, 0u);
    auto top = auction ? k_bvh8_tree_top<NXB_SLOT_AUCTION> : k_bvh8_tree_top<NXB_SLOT_GREEDY>;
    hipLaunchKernelGGL(top, dim3((trees + kTopTrees - 1) / kTopTrees), dim3(kBvh8Threads), 0, s, bvh2, n,
                       2 * n - trees, out, prim_idx, capacity, sc.work, sc.work2, work3, cnt, info, trees, mesh_lo,
                       ids, prims, sc.flags);
    auto level = auction ? k_bvh8_tree_level<NXB_SLOT_AUCTION> : k_bvh8_tree_level<NXB_SLOT_GREEDY>;
    for (uint32_t l = kTreeTopLevels; l < levels; ++l)
        hipLaunchKernelGGL(level, dim3(grid), dim3(kBvh8Threads), 0, s, bvh2, 2 * n - trees, out, prim_idx,
                           capacity, sc.work, sc.work2, work3, cnt, l, info, trees, ids, prims, sc.flags);
    hipLaunchKernelGGL(k_bvh8_level_total, dim3(1), dim3(1), 0, s, cnt, levels, sc.ctr, sc.flags);
    return hipGetLastError();
}

bool bvh8_levels_done(uint32_t n, const uint32_t* cnt_host, uint32_t launched, uint32_t* node_total) {
    if (n <= 1) {
        *node_total = n;
        return true;
    }
    // counts of levels 0..launched are final; an empty one ends the tree
    uint32_t total = 0, l = 0;
    for (; l <= launched && l <= kMaxLevels && cnt_host[l] != 0; ++l) total += cnt_host[l];
    *node_total = total;
    return l <= launched;
}

hipError_t launch_bvh8(const nxb_bvh2_node* bvh2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                       uint32_t capacity, uint32_t spin_limit, Bvh8Scratch sc, hipStream_t s) {
    hipError_t e = launch_bvh8_prep(n, capacity, sc, s);
    if (e != hipSuccess) return e;
    return launch_bvh8_run(bvh2, n, out, prim_idx, capacity, spin_limit, sc, s);
}

}  // namespace nxb
