// nxb_internal.h - launchers shared between the kernel files and the host
// orchestration (api.hip). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/nxb.h"

namespace nxb {

struct Box;

struct BoundsScratch {
    Box* partials;        // max_blocks boxes
    uint32_t* done;       // last-block-done counter, zero between builds
    uint32_t max_blocks;
};

// Global digit counts are kept in kHistCopies copies (block b adds into copy
// b % kHistCopies) so that per-address atomic traffic is divided by the copy
// count; readers sum the copies. One copy holds up to 11 passes of 512 digits.
constexpr uint32_t kHistCopies = 8;
constexpr uint32_t kHistCopyWords = 11 * 512;

struct SortScratch {
    uint32_t* hist;          // [kHistCopies][kHistCopyWords]: per copy [passes][radix] digit counts
    uint32_t* counts;        // [radix][tiles] per-tile digit counts, then output positions
    size_t counts_words;     // capacity of counts
};

struct RadiusParams {
    uint32_t radius;        // fixed radius (1..16)
    const uint8_t* table;   // per-segment radii or null
    uint32_t seg_log2;
    uint32_t table_len;
};

struct HplocScratch {
    uint32_t* parent;         // [n], filled with kInvalid before the kernel
    uint32_t* recs;           // [8n] cluster records: box, id, representative (no init needed)
    uint32_t* csizes;         // [n] cluster primitive counts (no init needed)
    nxb_bvh2_node* moved;     // one spare node slot
    uint32_t* moved_parent;   // one word
    uint32_t* moved_size;     // one word
    uint32_t* nsize;          // [n-1] output: primitives under each internal node
    uint32_t* flags;
    uint32_t* next_chunk;     // kHplocCounterWords words, zero before the kernel: chunk counters
    // BVH8 builds of one mesh: [n-1] wide records of the internal nodes (both
    // children's boxes, ids and primitive counts, 64 bytes, index id - n) and
    // one spare record for the node parked under the root's id; null: none.
    uint4* rec8 = nullptr;
    uint4* moved_rec = nullptr;
};

// Words of the H-PLOC chunk counters (one per XCD region, 128 bytes apart).
constexpr uint32_t kHplocCounterWords = 256;

struct Bvh8Scratch {
    uint64_t* work;           // [capacity] (prim base << 32 | BVH2 index) per BVH8 node
    unsigned long long* ctr;  // (finished << 32 | allocated) BVH8 nodes
    uint32_t* grab;           // work grab counter (own cache line)
    uint32_t* flags;
    const uint32_t* nsize;    // [n-1] primitives under each internal BVH2 node
    uint64_t* work2;          // [capacity] (right << 32 | left) children of each work item's BVH2 node
                              // (level-synchronous collapse only)
    uint32_t slot_mode;       // nxb_slot_assignment of the child slots
    const uint4* rec8 = nullptr;  // wide records of the internal nodes (level-synchronous collapse of
                                  // one mesh): read instead of the BVH2 nodes, nsize and work2
};

// indices (nullable): three vertex indices per triangle into prims, which
// then holds vertex_count vertices of 3 floats (type must be triangles).
hipError_t launch_scene_bounds(const void* prims, nxb_prim_type type, uint32_t n, nxb_bvh2_node* nodes,
                               nxb_aabb* scene, uint32_t* flags, BoundsScratch sc, hipStream_t s,
                               const uint32_t* indices = nullptr, uint32_t vertex_count = 0);
hipError_t launch_morton(const nxb_bvh2_node* nodes, uint32_t n, const nxb_aabb* scene, uint32_t bits,
                         void* keys, uint32_t* values, hipStream_t s);
// Morton codes plus the digit histograms of every pass of a sort plan
// (zeroed beforehand by launch_sort_prepare).
struct SortPlan;
hipError_t launch_morton_hist(const nxb_bvh2_node* nodes, uint32_t n, const nxb_aabb* scene, void* keys,
                              const SortPlan& pl, uint32_t* hist, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t s);
// One launch: hdr[0..hdr_words) = 0, parent[0..n) = kInvalid, and the first
// hist_words words of every digit-histogram copy = 0.
hipError_t launch_build_prep(uint32_t* hdr, uint32_t hdr_words, uint32_t* parent, uint32_t n, uint32_t* hist,
                             uint32_t hist_words, hipStream_t s);

// Stable LSD sort of (key, value) on key bits [begin_bit, end_bit).
// values_in may be null (iota). Result lands in keys_out/values_out.
struct SortPlan {
    uint32_t key_bits, begin, end, bits, passes, tiles;
    uint32_t wide;  // 1: wide tile shape (large inputs)
};
SortPlan plan_sort(uint32_t n, uint32_t key_bits, uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits);
// Digit width used when the caller leaves it open: 7 for 32-bit keys; 8 for
// 64-bit keys, 9 once the input takes the wide tile shape.
uint32_t default_digit_bits(uint32_t n, uint32_t key_bits);
size_t sort_scratch_bytes(uint32_t n, uint32_t key_bits, uint32_t digit_bits);
// Zeroes the digit histograms of a plan.
hipError_t launch_sort_prepare(const SortPlan& pl, SortScratch sc, hipStream_t s);
// Digit histograms of every pass (also produced by the Morton kernel).
hipError_t launch_sort_histogram(const void* keys, uint32_t n, const SortPlan& pl, SortScratch sc, hipStream_t s);
hipError_t launch_sort_passes(const void* keys_in, const uint32_t* values_in, void* keys_out, uint32_t* values_out,
                              void* keys_tmp, uint32_t* values_tmp, uint32_t n, const SortPlan& pl, SortScratch sc,
                              hipStream_t s);
hipError_t launch_radix_sort(const void* keys_in, const uint32_t* values_in, void* keys_out,
                             uint32_t* values_out, void* keys_tmp, uint32_t* values_tmp, uint32_t n,
                             uint32_t key_bits, uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits,
                             SortScratch sc, hipStream_t s);

// Most persistent H-PLOC blocks per CU (two waves each, six waves per SIMD
// at the kernel's register budget); blocks_per_cu 0 takes this many.
constexpr uint32_t kHplocMaxBlocksPerCu = 12;
hipError_t launch_hploc(const void* keys, uint32_t key_bits, uint32_t* cluster, uint32_t n,
                        nxb_bvh2_node* nodes, uint32_t merge_mode, RadiusParams rp, uint32_t blocks_per_cu,
                        HplocScratch sc, hipStream_t s);
// Batched H-PLOC: keys are (mesh << 32 | 32-bit Morton) sorted, mesh m owns
// sorted positions [mesh_lo[m], mesh_lo[m+1]). Leaves nodes[0..n) hold every
// primitive; the internal nodes of mesh m follow those of meshes < m from
// nodes[n] on, children as global node ids. moved/moved_parent/moved_size
// hold one entry per mesh.
hipError_t launch_hploc_batch(const uint64_t* keys, uint32_t* cluster, uint32_t n, const uint32_t* mesh_lo,
                              uint32_t meshes, nxb_bvh2_node* nodes, uint32_t merge_mode, RadiusParams rp,
                              uint32_t blocks_per_cu, HplocScratch sc, hipStream_t s);

hipError_t launch_bvh8_prep(uint32_t n, uint32_t capacity, Bvh8Scratch sc, hipStream_t s);
// Primitive counts of the internal nodes of an arbitrary BVH2 (root 2n-2).
// parent: [2n-1] scratch, arrive: [n-1] scratch, nsize: [n-1] output.
hipError_t launch_subtree_sizes(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t* parent, uint32_t* arrive,
                                uint32_t* nsize, uint32_t* flags, hipStream_t s);
hipError_t launch_bvh8_run(const nxb_bvh2_node* bvh2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                           uint32_t capacity, uint32_t spin_limit, Bvh8Scratch sc, hipStream_t s);
// Level-synchronous BVH8 of one BVH2 (root 2n-2): enqueues the kernels of
// levels [from, from + count) (from == 0 also initializes; count is then at
// least kBvh8TopLevels). cnt: kBvh8LevelWords
// device words of level counters. After reading them back (levels
// 0..launched are final), bvh8_levels_done tells whether the tree is
// complete and sets the node count; if not, enqueue more levels.
constexpr uint32_t kBvh8LevelWords = 97;
constexpr uint32_t kBvh8FirstLevels = 24;
constexpr uint32_t kBvh8TopLevels = 3;  // levels 0..2 run in one single-block launch (from == 0)
hipError_t launch_bvh8_levels(const nxb_bvh2_node* bvh2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                              uint32_t capacity, Bvh8Scratch sc, uint32_t* cnt, uint32_t from, uint32_t count,
                              hipStream_t s);
bool bvh8_levels_done(uint32_t n, const uint32_t* cnt_host, uint32_t launched, uint32_t* node_total);
// Level-synchronous BVH8 of a forest (roots in work[0..trees), as for
// launch_bvh8_forest) with a fixed number of levels; the node total ends in
// the low word of *sc.ctr, and a tree deeper than `levels` sets the
// bad-index flag.
hipError_t launch_bvh8_forest_levels(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t trees, nxb_bvh8_node* out,
                                     uint32_t* prim_idx, uint32_t capacity, uint32_t levels, Bvh8Scratch sc,
                                     uint32_t* cnt, hipStream_t s);
// As launch_bvh8_forest_levels for the forest of a batch (tree t's root is
// its last internal node, mesh_lo[] as for launch_hploc_batch), but each
// tree t writes its own contiguous BVH8 (root 0, tree-local child and
// primitive bases) to out + info[t].z with at most info[t].w nodes, and its
// primitives to prim_idx + info[t].x (info[t].y of them). Levels
// 0..kTreeTopLevels-1 run in one launch, a block per few trees; their level
// counters stay 0, so the depth of the forest is kTreeTopLevels plus the
// number of non-empty counters from cnt[kTreeTopLevels] on. levels >=
// kTreeTopLevels. Each node takes the primIdx slots of its leaves from its
// tree's counter prims[t] (allocation order; no subtree sizes are read).
// After the levels ids[t] is tree t's node count and prims[t] its primitive
// count.
constexpr uint32_t kTreeTopLevels = 3;
hipError_t launch_bvh8_tree_levels(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t trees, nxb_bvh8_node* out,
                                   uint32_t* prim_idx, uint32_t capacity, uint32_t levels, Bvh8Scratch sc,
                                   uint32_t* cnt, uint64_t* work3, const uint4* info, const uint32_t* mesh_lo,
                                   uint32_t* ids, uint32_t* prims, hipStream_t s);
hipError_t launch_bvh8(const nxb_bvh2_node* bvh2, uint32_t n, nxb_bvh8_node* out, uint32_t* prim_idx,
                       uint32_t capacity, uint32_t spin_limit, Bvh8Scratch sc, hipStream_t s);
// BVH8 of a forest (n leaves, `trees` BVH2s, work[0..trees) seeded with the
// roots and their first primitive) with `lanes` persistent lanes.
hipError_t launch_bvh8_forest(const nxb_bvh2_node* bvh2, uint32_t n, uint32_t trees, nxb_bvh8_node* out,
                              uint32_t* prim_idx, uint32_t capacity, uint32_t lanes, uint32_t spin_limit,
                              Bvh8Scratch sc, hipStream_t s);

hipError_t launch_segment_features(const void* keys, uint32_t key_bits, const uint32_t* perm,
                                   const nxb_bvh2_node* leaves, uint32_t n, uint32_t seg_log2, float* feat,
                                   hipStream_t s);
hipError_t launch_segment_sah(const nxb_bvh2_node* nodes, uint32_t n, uint32_t seg_log2, const nxb_aabb* scene,
                              double* seg_cost, hipStream_t s);
hipError_t launch_mlp_table(const float* params, const float* feat, uint32_t count, float min_confidence,
                            uint8_t* table, float* logits, uint32_t* fallbacks, hipStream_t s);
hipError_t launch_mlp_grad(const float* params, const float* feat, const uint8_t* labels, uint32_t count,
                           float* grads, float* loss, hipStream_t s);
hipError_t launch_adam(float* params, const float* grads, float* m, float* v, uint32_t step, float lr, float beta1,
                       float beta2, float eps, hipStream_t s);

// Closest-hit traversal of a BVH8 (one lane per ray).
// indices (nullable): tris is then vertex_count vertices of three floats and
// indices three vertex indices per triangle.
hipError_t launch_trace(const nxb_bvh8_node* nodes, uint32_t node_count, const uint32_t* prim_idx,
                        uint32_t prim_count, const float* tris, const nxb_ray* rays, uint32_t ray_count,
                        nxb_hit* hits, uint32_t* stats, hipStream_t s, const uint32_t* indices = nullptr,
                        uint32_t vertex_count = 0);

hipError_t launch_sah_bvh2(const nxb_bvh2_node* nodes, uint32_t count, const nxb_aabb* scene, double* acc,
                           hipStream_t s);
// BVH2 SAH from the wide records of a BVH8 build (n >= 2 primitives).
hipError_t launch_sah_rec8(const uint4* rec, uint32_t n, const nxb_aabb* scene, double* acc, hipStream_t s);
hipError_t launch_sah_bvh8(const nxb_bvh8_node* nodes, uint32_t count, const nxb_aabb* scene, double* acc,
                           hipStream_t s);

// Number of compute units of the current device (cached).
uint32_t device_cu_count();

}  // namespace nxb
