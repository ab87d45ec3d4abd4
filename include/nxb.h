/*
 * nxb.h - C ABI of the MI355X-native NexusBVH construction pipeline (libnxb.so).
 *
 * Plain C: fixed-layout POD types, plain pointers and sizes, status codes.
 * Every entry point that replaces a reference function cites it as
 * `file:line` relative to the NexusBVH source tree.
 *
 * Memory conventions (same as the reference, NexusBVH/src/BVHBuilder.cpp):
 *   - primitive buffers are DEVICE pointers, borrowed, never freed here;
 *   - returned BVH structs are by-value handles that own device buffers and
 *     are released with nxb_free_device_bvh2 / nxb_free_device_bvh8;
 *   - nxb_bvh2_to_host allocates host memory released by nxb_free_host_bvh2
 *     (nxb_bvh8_to_host: nxb_free_host_bvh8).
 * Differences: errors are returned as nxb_status (the reference prints and
 * calls exit(99)), and every call takes an explicit hipStream_t (NULL = the
 * null stream). Build calls are blocking, as in the reference.
 */
#ifndef NXB_H_
#define NXB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t is a pointer to an opaque struct; declared here so that this
 * header does not pull in the HIP headers. */
typedef struct ihipStream_t* nxb_stream_t;

#define NXB_INVALID_IDX 0xFFFFFFFFu  /* NexusBVH/include/NXB/BVH.h:8 */

typedef enum nxb_status {
    NXB_SUCCESS = 0,
    NXB_ERROR_INVALID_VALUE = 1,     /* null pointer, bad enum, bad size      */
    NXB_ERROR_OUT_OF_MEMORY = 2,     /* device or host allocation failed     */
    NXB_ERROR_NONFINITE_INPUT = 3,   /* a vertex/box coordinate is NaN/Inf   */
    NXB_ERROR_TOO_LARGE = 4,         /* primitive count above NXB_MAX_PRIMS  */
    NXB_ERROR_TIMEOUT = 5,           /* a bounded device spin hit its limit  */
    NXB_ERROR_HIP = 6,               /* HIP runtime error                    */
    NXB_ERROR_NOT_SUPPORTED = 7      /* option combination not implemented   */
} nxb_status;

/* 2^30: node ids (up to 2n-2) and the byte offsets of the device passes
 * stay within 32 bits. */
#define NXB_MAX_PRIMS (1u << 30)

/* ---- geometry (layouts of NexusBVH/include/NXB/{AABB,Triangle}.h) ------ */
typedef struct nxb_float3 { float x, y, z; } nxb_float3;                 /* 12 B */
typedef struct nxb_aabb { nxb_float3 bmin, bmax; } nxb_aabb;             /* 24 B, AABB.h:8-54 */
typedef struct nxb_triangle { nxb_float3 v0, v1, v2; } nxb_triangle;     /* 36 B, Triangle.h:10-48 */

/* Order of NXB::PrimType (BVH.h:10-14). */
typedef enum nxb_prim_type { NXB_PRIM_AABB = 0, NXB_PRIM_TRIANGLE = 1 } nxb_prim_type;

/* BVH2::Node (BVH.h:18-27): leaf <=> left == NXB_INVALID_IDX, right = primIdx. */
typedef struct nxb_bvh2_node {
    nxb_aabb bounds;
    uint32_t left;
    uint32_t right;
} nxb_bvh2_node;                                                         /* 32 B */

/* BVH8::NodeExplicit (BVH.h:40-66), Ylitie et al. compressed wide node. */
typedef struct nxb_bvh8_node {
    nxb_float3 p;          /* origin of the child quantization grid            */
    uint8_t e[3];          /* biased exponents of the grid scale per axis      */
    uint8_t imask;         /* bit i set <=> slot i holds an inner node         */
    uint32_t child_base;   /* index of the first inner child node              */
    uint32_t prim_base;    /* index of the first leaf in the primIdx array     */
    uint8_t meta[8];       /* inner: 0x20|(24+i); leaf: 0x20|rank; empty: 0    */
    uint8_t qlox[8], qloy[8], qloz[8];
    uint8_t qhix[8], qhiy[8], qhiz[8];
} nxb_bvh8_node;                                                         /* 80 B */

/* BVH2 (BVH.h:16-35). Nodes 0..n-1 are the leaves in input order, internal
 * nodes are n..2n-2 and the root is node 2n-2. */
typedef struct nxb_bvh2 {
    nxb_bvh2_node* nodes;
    uint32_t node_count;
    uint32_t prim_count;
    nxb_aabb bounds;
} nxb_bvh2;

/* BVH8 (BVH.h:38-94). Root is node 0. */
typedef struct nxb_bvh8 {
    nxb_bvh8_node* nodes;
    uint32_t node_count;
    uint32_t* prim_idx;
    uint32_t prim_count;
    nxb_aabb bounds;
} nxb_bvh8;

/* H-PLOC merge schedule. */
typedef enum nxb_merge_mode {
    /* The reference schedule (BinaryBuilder.cu:9-10,42-197): at most 16
     * clusters loaded per child, merge down to 16, nearest neighbours inside
     * a 32-cluster window. Bit-for-bit the reference structure. */
    NXB_MERGE_WAVE32 = 0,
    /* Native CDNA schedule: 64-cluster window, 32 per child, threshold 32. */
    NXB_MERGE_WAVE64 = 1
} nxb_merge_mode;

/* Per-segment policies: a merge task over sorted positions [l, r) with
 * r - l <= 2^segment_log2 takes the radius of the segment of l; wider tasks
 * (the upper tree levels) and table entries 0 take search_radius. */
typedef enum nxb_radius_policy {
    NXB_RADIUS_FIXED = 0,      /* search_radius for every merge                          */
    NXB_RADIUS_TABLE = 1,      /* caller's per-segment radius table                      */
    NXB_RADIUS_LEARNED = 2     /* per-segment radius proposed by the MLP (d_mlp_params)  */
} nxb_radius_policy;

/* Learned radius MLP: 16 segment features -> 64 -> 64 -> 4 classes (radius
 * 2, 4, 8, 16), ReLU. FP32 parameters, row-major [in][out]:
 * W1 16x64 | b1 64 | W2 64x64 | b2 64 | W3 64x4 | b3 4. */
#define NXB_MLP_FEATURES 16
#define NXB_MLP_HIDDEN 64
#define NXB_MLP_CLASSES 4
#define NXB_MLP_PARAMS 5508

/* Extends BuildConfig (NexusBVH/include/NXB/BuildConfig.h:6-12). */
typedef struct nxb_build_config {
    uint32_t prioritize_speed;  /* BuildConfig::prioritizeSpeed: nonzero = 32-bit Morton */
    uint32_t morton_bits;       /* 0 = follow prioritize_speed, else 32 or 64            */
    uint32_t search_radius;     /* H-PLOC radius, 1..16 (reference: 8)                   */
    uint32_t merge_mode;        /* nxb_merge_mode                                        */
    uint32_t radius_policy;     /* nxb_radius_policy                                     */
    uint32_t segment_log2;      /* table policy: segment size = 2^segment_log2 sorted prims */
    const uint8_t* d_radius_table; /* table policy: device array, one radius per segment */
    uint32_t radius_table_len;  /* number of entries in d_radius_table                   */
    uint32_t sort_digit_bits;   /* 0 = auto, else 6..9                                   */
    uint32_t spin_limit;        /* 0 = default watchdog for bounded device spins (batches) */
    uint32_t metrics_flags;     /* NXB_METRICS_* bits                                    */
    const float* d_mlp_params;  /* learned policy: device array of NXB_MLP_PARAMS floats */
    float mlp_min_confidence;   /* learned policy: below this class probability use r=8  */
    uint32_t slot_assignment;   /* nxb_slot_assignment: BVH8 child slot order            */
    uint32_t hploc_blocks_per_cu; /* H-PLOC persistent blocks (two waves each) per CU:   */
                                  /* 0 = default (12, six waves per SIMD), else 1..12     */
} nxb_build_config;

/* BVH8 child slot assignment. GREEDY is the reference's (WideConverter.cu:
 * 106-153); AUCTION is its compiled-out alternative (WideConverter.cu:6,
 * 52-102, 336-361): Bertsekas' auction with epsilon scaling on the same cost
 * table, here with at most 16 epsilon phases of at most 4096 bids (children
 * still unassigned then take the first free slots in child order). Slot
 * order changes traversal order only: the nodes, primitives and SAH are the
 * same as GREEDY's, and the collapse is several times slower. */
typedef enum nxb_slot_assignment { NXB_SLOT_GREEDY = 0, NXB_SLOT_AUCTION = 1 } nxb_slot_assignment;

/* metrics_flags: record stage times only (no SAH kernels, no extra sync). */
#define NXB_METRICS_TIMES_ONLY 1u

/* BVHBuildMetrics (NexusBVH/include/NXB/BVHBuildMetrics.h:36-50): same nine
 * floats in the same order, times in ms. Costs: the SAH of the reference
 * (Eval.cu:12-80); here each per-node FP32 term is summed in FP64 and the
 * sum is rounded to float. */
typedef struct nxb_metrics {
    float compute_scene_bounds_time;
    float compute_morton_codes_time;
    float radix_sort_time;
    float bvh_build_time;
    float bvh8_conversion_time;
    float total_time;
    float bvh2_cost;
    float bvh8_cost;
    float average_child_per_node;
} nxb_metrics;

/* Fills *cfg with the reference defaults (64-bit Morton, radius 8, wave32). */
void nxb_default_config(nxb_build_config* cfg);

const char* nxb_status_string(nxb_status s);

/* ---- builder entry points (NexusBVH/include/NXB/BVHBuilder.h:18-55) ----- */

/* Replaces NXB::BuildBVH2<PrimT> (BVHBuilder.h:18-19; BVHBuilder.cpp:114-171).
 * metrics may be NULL (then no per-stage events are recorded). */
nxb_status nxb_build_bvh2(const void* d_prims, nxb_prim_type type, uint32_t prim_count,
                          const nxb_build_config* cfg, nxb_bvh2* out,
                          nxb_metrics* metrics, nxb_stream_t stream);

/* Replaces NXB::BuildBVH8<PrimT> (BVHBuilder.h:30-31; BVHBuilder.cpp:173-267). */
nxb_status nxb_build_bvh8(const void* d_prims, nxb_prim_type type, uint32_t prim_count,
                          const nxb_build_config* cfg, nxb_bvh8* out,
                          nxb_metrics* metrics, nxb_stream_t stream);

/* Allocation-free BVH2 build. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid allocation). Builds what
 * nxb_build_bvh2 builds (the same nodes, byte for byte) into the caller's
 * d_nodes (2n-1 nodes); out->nodes is d_nodes, out->bounds the scene box.
 * Scratch is a static device array of NXB_NOALLOC_SCRATCH_MIB MiB per GPU
 * (held until the process ends); no device or pinned memory is allocated and
 * no timing events are created. At most NXB_NOALLOC_MAX_PRIMS primitives
 * (NXB_ERROR_TOO_LARGE above). All calls on one GPU share its array: calls
 * that overlap on one GPU (other threads, other streams before this one has
 * returned) race. Blocking, like nxb_build_bvh2. */
#define NXB_NOALLOC_MAX_PRIMS (1u << 19)
#define NXB_NOALLOC_SCRATCH_MIB 48u
nxb_status nxb_build_bvh2_noalloc(const void* d_prims, nxb_prim_type type, uint32_t prim_count,
                                  const nxb_build_config* cfg, nxb_bvh2_node* d_nodes, nxb_bvh2* out,
                                  nxb_stream_t stream);

/* Indexed triangle meshes (no counterpart in the reference, whose
 * primitives are Triangle/AABB records, BVHBuilder.cpp:296-300): triangle i
 * is the vertices d_indices[3i..3i+2] of d_vertices (vertex_count vertices of
 * 3 floats). The result equals nxb_build_bvh2/8 on the expanded triangles;
 * an index >= vertex_count returns NXB_ERROR_INVALID_VALUE. */
nxb_status nxb_build_bvh2_indexed(const float* d_vertices, uint32_t vertex_count, const uint32_t* d_indices,
                                  uint32_t tri_count, const nxb_build_config* cfg, nxb_bvh2* out,
                                  nxb_metrics* metrics, nxb_stream_t stream);
nxb_status nxb_build_bvh8_indexed(const float* d_vertices, uint32_t vertex_count, const uint32_t* d_indices,
                                  uint32_t tri_count, const nxb_build_config* cfg, nxb_bvh8* out,
                                  nxb_metrics* metrics, nxb_stream_t stream);

/* Replaces NXB::ToHost(BVH2) (BVHBuilder.h:40; BVHBuilder.cpp:269-278). */
nxb_status nxb_bvh2_to_host(const nxb_bvh2* device_bvh, nxb_bvh2* host_bvh);
/* Replaces NXB::FreeHostBVH(BVH2) (BVHBuilder.h:45; BVHBuilder.cpp:280-283). */
void nxb_free_host_bvh2(nxb_bvh2* host_bvh);
/* ToHost for a BVH8 (no counterpart: the reference's ToHost takes a BVH2
 * only, BVHBuilder.cpp:269-278): copies nodes and primIdx into host memory
 * released by nxb_free_host_bvh8. Synchronous, like nxb_bvh2_to_host. */
nxb_status nxb_bvh8_to_host(const nxb_bvh8* device_bvh, nxb_bvh8* host_bvh);
void nxb_free_host_bvh8(nxb_bvh8* host_bvh);
/* Replaces NXB::FreeDeviceBVH(BVH2) (BVHBuilder.h:50; BVHBuilder.cpp:285-288). */
nxb_status nxb_free_device_bvh2(nxb_bvh2* device_bvh);
/* Replaces NXB::FreeDeviceBVH(BVH8) (BVHBuilder.h:55; BVHBuilder.cpp:290-294). */
nxb_status nxb_free_device_bvh8(nxb_bvh8* device_bvh);

/* ---- stage entry points (each a stage of BVHBuilder.cpp) --------------- *
 * All take device pointers; they enqueue on `stream` and return without
 * synchronizing unless stated. Without a workspace of the caller's they use
 * the library's per-device scratch buffer; work on it is ordered across
 * streams (a later build or stage call waits on the device for the earlier
 * asynchronous stage calls). */

/* Scratch bytes needed by the stage entry points for prim_count primitives;
 * a larger buffer may be passed. */
size_t nxb_stage_workspace_bytes(uint32_t prim_count, uint32_t morton_bits);

/* ComputeSceneBoundsKernel (Setup.cu:12-39): writes leaf nodes[0..n) and the
 * scene box. d_flags (one uint32) gets bit 0 set if a coordinate is not
 * finite; it must be zero on entry. */
nxb_status nxb_compute_scene_bounds(const void* d_prims, nxb_prim_type type, uint32_t n,
                                    nxb_bvh2_node* d_nodes, nxb_aabb* d_scene,
                                    uint32_t* d_flags, nxb_stream_t stream);

/* ComputeMortonCodesKernel (Setup.cu:42-59): keys from leaf-box centroids
 * normalized to the scene box; d_values (nullable) gets the iota 0..n-1. */
nxb_status nxb_compute_morton(const nxb_bvh2_node* d_nodes, uint32_t n, const nxb_aabb* d_scene,
                              uint32_t morton_bits, void* d_keys, uint32_t* d_values,
                              nxb_stream_t stream);

/* RadixSort (Setup.cu:62-112): stable LSD sort of (key, value) on key bits
 * [begin_bit, end_bit). Inputs are read from d_keys_in/d_values_in (when
 * d_values_in is NULL the values are the iota 0..n-1), outputs written to
 * d_keys_out/d_values_out; all four buffers are distinct. digit_bits is
 * 6..9, or 0 for the library's choice. */
nxb_status nxb_radix_sort_pairs(const void* d_keys_in, const uint32_t* d_values_in,
                                void* d_keys_out, uint32_t* d_values_out, uint32_t n,
                                uint32_t key_bits, uint32_t begin_bit, uint32_t end_bit,
                                uint32_t digit_bits, void* d_workspace, size_t workspace_bytes,
                                nxb_stream_t stream);

/* BuildBVH2Kernel (BinaryBuilder.cu:200-273): H-PLOC over sorted keys.
 * d_sorted_values is the sorted primitive permutation (read, not modified).
 * d_nodes[0..n) must hold the leaves;
 * internal nodes n..2n-2 are written, root = 2n-2. */
nxb_status nxb_build_hploc(const void* d_sorted_keys, uint32_t* d_sorted_values, uint32_t n,
                           uint32_t key_bits, const nxb_build_config* cfg,
                           nxb_bvh2_node* d_nodes, void* d_workspace, size_t workspace_bytes,
                           nxb_stream_t stream);

/* BuildBVH8Kernel (WideConverter.cu:225-414): collapse a BVH2 (root 2n-2)
 * into d_bvh8 (capacity ceil((4n-1)/7) nodes) and d_prim_idx (n entries).
 * Synchronizes the stream and returns the node count in *out_node_count.
 * d_workspace (nullable) needs nxb_stage_workspace_bytes(n, 32) bytes; a
 * buffer sized for 64-bit keys also serves. */
nxb_status nxb_collapse_bvh8(const nxb_bvh2_node* d_bvh2, uint32_t n, nxb_bvh8_node* d_bvh8,
                             uint32_t* d_prim_idx, uint32_t* out_node_count,
                             uint32_t spin_limit, void* d_workspace, size_t workspace_bytes,
                             nxb_stream_t stream);

/* ComputeBVH2CostKernel (Eval.cu:12-33): SAH of a device BVH2, FP32 per-node
 * terms accumulated in FP64. Synchronizes the stream. */
nxb_status nxb_sah_bvh2(const nxb_bvh2* device_bvh, double* out_cost, nxb_stream_t stream);
/* ComputeBVH8CostKernel (Eval.cu:36-80). Synchronizes the stream. */
nxb_status nxb_sah_bvh8(const nxb_bvh8* device_bvh, double* out_cost, nxb_stream_t stream);

/* ---- ray traversal of a BVH8 (no counterpart in the reference) ---------- *
 * The consumer of the compressed wide node (BVH.h:39-96; Ylitie, Karras and
 * Laine 2017): closest-hit queries against the triangles a BVH8 was built
 * from. */
typedef struct nxb_ray {
    nxb_float3 origin;
    float tmin;
    nxb_float3 dir;
    float tmax;
} nxb_ray;                                                                /* 32 B */

typedef struct nxb_hit {
    float t;        /* hit distance in units of |dir|; tmax when nothing is hit */
    uint32_t prim;  /* triangle index (input order), NXB_INVALID_IDX when none   */
    float u, v;     /* barycentrics of the hit (v1 weight u, v2 weight v)        */
} nxb_hit;                                                                /* 16 B */

/* Closest hit of each ray (d_rays[ray_count] -> d_hits[ray_count]) among the
 * triangles d_tris (9 floats each, the array the BVH8 was built from),
 * Moller-Trumbore intersection in FP32 without contraction; a hit counts for
 * tmin <= t <= tmax, and equal distances go to the smaller triangle index.
 * d_stats (nullable) receives per ray the BVH8 nodes visited and the
 * triangles tested. Enqueued on the stream; does not synchronize. */
nxb_status nxb_trace_closest(const nxb_bvh8* device_bvh, const float* d_tris, const nxb_ray* d_rays,
                             uint32_t ray_count, nxb_hit* d_hits, uint32_t* d_stats, nxb_stream_t stream);

/* The same for a BVH8 of an indexed mesh (nxb_build_bvh8_indexed): the
 * triangles are read through d_indices (three per triangle) from d_vertices
 * (vertex_count vertices of three floats); hits and stats equal those of
 * nxb_trace_closest on the expanded triangles. A triangle with an index past
 * vertex_count is never hit. No counterpart in the reference. */
nxb_status nxb_trace_closest_indexed(const nxb_bvh8* device_bvh, const float* d_vertices, uint32_t vertex_count,
                                     const uint32_t* d_indices, const nxb_ray* d_rays, uint32_t ray_count,
                                     nxb_hit* d_hits, uint32_t* d_stats, nxb_stream_t stream);

/* ---- learned search radius (no counterpart in the reference) ----------- *
 * Segments are the sorted positions [s << segment_log2, (s+1) << segment_log2);
 * segment_log2 is 8..20. */

/* 16 features per segment from the sorted keys, the sorted permutation and
 * the leaf nodes (nodes[0..n)); d_features holds ceil(n / 2^segment_log2) * 16
 * floats. */
nxb_status nxb_segment_features(const void* d_sorted_keys, const uint32_t* d_sorted_values,
                                const nxb_bvh2_node* d_leaves, uint32_t n, uint32_t key_bits,
                                uint32_t segment_log2, float* d_features, nxb_stream_t stream);

/* Per-segment SAH of a BVH2 built by this library (internal node id
 * n - 1 + p is charged to the segment of sorted position p). d_seg_cost
 * (ceil(n / 2^segment_log2) doubles) is accumulated into, not cleared. */
nxb_status nxb_segment_sah(const nxb_bvh2_node* d_nodes, uint32_t n, const nxb_aabb* d_scene,
                           uint32_t segment_log2, double* d_seg_cost, nxb_stream_t stream);

/* MLP inference: radius table (2, 4, 8 or 16 per segment; 8 when the
 * arg-max probability is below min_confidence). d_logits (count * 4 floats)
 * and d_fallbacks (one uint32, accumulated) may be NULL. */
nxb_status nxb_mlp_radius_table(const float* d_params, const float* d_features, uint32_t count,
                                float min_confidence, uint8_t* d_table, float* d_logits,
                                uint32_t* d_fallbacks, nxb_stream_t stream);

/* Mean cross-entropy of the MLP over count labelled samples (labels are
 * class indices 0..3): adds the loss to *d_loss and its gradient to d_grads
 * (NXB_MLP_PARAMS floats); both must be zeroed by the caller. */
nxb_status nxb_mlp_gradients(const float* d_params, const float* d_features, const uint8_t* d_labels,
                             uint32_t count, float* d_grads, float* d_loss, nxb_stream_t stream);

/* One Adam step (step >= 1) on d_params with moment buffers d_m, d_v. */
nxb_status nxb_mlp_adam_step(float* d_params, const float* d_grads, float* d_m, float* d_v, uint32_t step,
                             float lr, float beta1, float beta2, float eps, nxb_stream_t stream);

/* Segments and radius-8 fallbacks of the last build with the learned policy. */
void nxb_last_learned_stats(uint32_t* segments, uint32_t* fallbacks);

/* ---- batched builds (no counterpart in the reference) ------------------ *
 * Many independent meshes stored back to back in one primitive buffer: mesh
 * i is primitives [h_offsets[i], h_offsets[i+1]) (at least 2 each). A plan
 * owns the workspace and the outputs and records bounds, Morton, sort,
 * H-PLOC and BVH8 for all meshes into one hipGraph. Each mesh's BVH8 equals
 * nxb_build_bvh8 of that mesh alone (32-bit Morton keys, fixed radius). */
typedef struct nxb_batch_plan nxb_batch_plan;

/* cfg: morton_bits must select 32-bit keys; search_radius, merge_mode and
 * sort_digit_bits apply to every mesh. bvh8_lanes: 0 = level-synchronous
 * BVH8 collapse (one kernel per level; each mesh's nodes go straight to its
 * own region of nxb_bvh8_capacity(count) nodes); nonzero = persistent
 * collapse kernel with that many lanes, followed by a per-mesh compaction of
 * its nodes. d_prims must stay valid while the plan is used. */
nxb_status nxb_batch_plan_create(const void* d_prims, nxb_prim_type type, uint32_t mesh_count,
                                 const uint32_t* h_offsets, const nxb_build_config* cfg,
                                 uint32_t bvh8_lanes, nxb_batch_plan** plan);
/* Builds every mesh (one graph launch) and waits. With stage_ms != NULL the
 * stages are enqueued one by one instead and timed: bounds, Morton, sort,
 * H-PLOC, BVH8 (with the compaction of the persistent kernel), total. */
nxb_status nxb_batch_plan_run(nxb_batch_plan* plan, nxb_stream_t stream, float* stage_ms);
/* Mesh i's BVH8 from the last run, as a view into the plan's buffers (valid
 * until the next run or destroy; do not free it). */
nxb_status nxb_batch_plan_mesh(const nxb_batch_plan* plan, uint32_t mesh, nxb_bvh8* out);
/* BVH8 nodes of all meshes in the last run. */
uint32_t nxb_batch_plan_node_total(const nxb_batch_plan* plan);
void nxb_batch_plan_destroy(nxb_batch_plan* plan);

/* Capacity of the BVH8 node array for n primitives: ceil((4n-1)/7)
 * (BVHBuilder.cpp:182-184). */
uint32_t nxb_bvh8_capacity(uint32_t n);

/* Library version string. */
const char* nxb_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NXB_H_ */
