// batch.hip - BVH8 builds of many independent meshes in one pipeline.
//
// No counterpart in the reference (it builds one mesh per call). Meshes are
// stored back to back; mesh m owns primitives [lo_m, lo_m + cnt_m). Every
// stage runs once for all meshes:
//   bounds   per-mesh scene boxes (order-preserving integer atomics, one
//            per wave when the wave lies inside one mesh); leaves keep the
//            mesh-local primitive index;
//   Morton   32-bit codes against the mesh's box, key = mesh << 30 | code
//            (a code has 30 bits);
//   sort     one stable radix sort of the composite keys: the meshes stay
//            grouped and each mesh is sorted exactly as alone;
//   H-PLOC   batched kernel (hploc.hip), each merge task inside its mesh;
//   BVH8     level-synchronous over the forest (one root per mesh), each
//            mesh writing its nodes straight into its own region of
//            nxb_bvh8_capacity(count) nodes with mesh-local node ids
//            (one counter per mesh) and primitive bases; or the persistent
//            kernel over the forest, then a stable per-mesh compaction of
//            its nodes with mesh-local child/primitive bases.
// Each mesh's BVH8 equals the one nxb_build_bvh8 builds for that mesh alone
// (32-bit Morton keys). A plan owns the workspace and the outputs and
// records the whole sequence into one hipGraph, so a rebuild of all meshes
// is a single graph launch.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/nxb.h"
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

constexpr int kBatchThreads = 256;
constexpr uint32_t kBatchHplocChunks = 32;  // first word of the H-PLOC chunk counters after the flags

// Largest m with mesh_lo[m] <= g (mesh_lo[0] = 0 <= g < mesh_lo[meshes]).
__device__ __forceinline__ uint32_t mesh_of(const uint32_t* __restrict__ mesh_lo, uint32_t meshes, uint32_t g) {
    uint32_t lo = 0, hi = meshes;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (mesh_lo[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr uint32_t kBatchBoundsTile = 1024;  // primitives per tile (4 per thread)
constexpr uint32_t kBatchTileMeshes = 64;    // meshes of a tile reduced in LDS (later ones: global atomics)

// Leaf nodes and per-mesh boxes. A block takes tiles of kBatchBoundsTile
// primitives: the primitives come in through LDS with 16-byte loads and the
// leaf nodes go out through it as contiguous 16-byte stores (as the
// single-scene bounds kernel). The boxes of the tile's first
// kBatchTileMeshes meshes are reduced with LDS atomics on ordered integers,
// then each mesh the tile touches gets one global atomic per bound.
template <int PF>
__global__ __launch_bounds__(kBatchThreads) void k_batch_bounds(const float* __restrict__ prims, uint32_t n,
                                                                const uint32_t* __restrict__ mesh_lo,
                                                                uint32_t meshes, nxb_bvh2_node* __restrict__ nodes,
                                                                uint32_t* bounds_ord, uint32_t* flags) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    constexpr uint32_t T = kBatchBoundsTile, PER = T / kBatchThreads;
    __shared__ __attribute__((aligned(16))) float tile[T * (PF > 8 ? PF : 8)];
    __shared__ uint32_t lo_local[kBatchTileMeshes + 1];  // mesh_lo[m0 + q], ~0u past the last mesh
    __shared__ uint32_t ord[kBatchTileMeshes * 6];
    __shared__ uint32_t sm0;
    const uint32_t tid = threadIdx.x;
    const bool aligned = (reinterpret_cast<uintptr_t>(prims) & 15u) == 0;
    const uint32_t tiles = (n + T - 1) / T;
    bool bad = false;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint32_t base = t * T;
        const uint32_t count = min(T, n - base);
        const float* src = prims + (size_t)base * PF;
        __syncthreads();
        if (tid == 0) sm0 = mesh_of(mesh_lo, meshes, base);
        if (aligned && count == T) {
            const f4v* s4 = reinterpret_cast<const f4v*>(src);
            f4v* t4 = reinterpret_cast<f4v*>(tile);
#pragma unroll
            for (uint32_t k = 0; k < (T * PF / 4) / kBatchThreads; ++k)
                t4[k * kBatchThreads + tid] = __builtin_nontemporal_load(s4 + k * kBatchThreads + tid);
        } else {
            for (uint32_t k = tid; k < count * PF; k += kBatchThreads) tile[k] = src[k];
        }
        for (uint32_t k = tid; k < kBatchTileMeshes * 6; k += kBatchThreads) ord[k] = (k % 6) < 3 ? ~0u : 0u;
        __syncthreads();
        const uint32_t m0 = sm0;
        if (tid <= kBatchTileMeshes) lo_local[tid] = m0 + tid <= meshes ? mesh_lo[m0 + tid] : ~0u;
        __syncthreads();
        Box bx[PER];
        uint32_t local[PER];
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            const uint32_t j = k * kBatchThreads + tid;
            const uint32_t g = base + j;
            const bool valid = j < count;
            bx[k] = Box{INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            local[k] = 0;
            uint32_t q = kBatchTileMeshes;  // tile-local mesh, kBatchTileMeshes: past the LDS table
            if (valid) {
                bx[k] = prim_box<PF>(tile + j * PF, bad);
                if (g < lo_local[kBatchTileMeshes]) {
                    // largest q with lo_local[q] <= g
                    uint32_t a = 0, z = kBatchTileMeshes;
                    while (z - a > 1) {
                        const uint32_t mid = (a + z) >> 1;
                        if (lo_local[mid] <= g) a = mid;
                        else z = mid;
                    }
                    q = a;
                    local[k] = g - lo_local[a];
                } else {
                    const uint32_t m = mesh_of(mesh_lo, meshes, g);
                    local[k] = g - mesh_lo[m];
                    uint32_t* o = bounds_ord + (size_t)m * 6;
                    atomicMin(o + 0, f2ord(bx[k].lx)); atomicMin(o + 1, f2ord(bx[k].ly));
                    atomicMin(o + 2, f2ord(bx[k].lz)); atomicMax(o + 3, f2ord(bx[k].hx));
                    atomicMax(o + 4, f2ord(bx[k].hy)); atomicMax(o + 5, f2ord(bx[k].hz));
                }
            }
            const uint32_t q0 = __builtin_amdgcn_readfirstlane(q);
            if (q0 < kBatchTileMeshes && ballot(valid && q != q0) == 0) {
                // the wave's valid elements lie in one mesh: reduce, then one lane
                Box b = bx[k];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    Box o;
                    o.lx = __shfl_xor(b.lx, off); o.ly = __shfl_xor(b.ly, off); o.lz = __shfl_xor(b.lz, off);
                    o.hx = __shfl_xor(b.hx, off); o.hy = __shfl_xor(b.hy, off); o.hz = __shfl_xor(b.hz, off);
                    b = box_union(b, o);
                }
                if (lane_id() == 0) {
                    uint32_t* o = ord + q0 * 6;
                    atomicMin(o + 0, f2ord(b.lx)); atomicMin(o + 1, f2ord(b.ly)); atomicMin(o + 2, f2ord(b.lz));
                    atomicMax(o + 3, f2ord(b.hx)); atomicMax(o + 4, f2ord(b.hy)); atomicMax(o + 5, f2ord(b.hz));
                }
            } else if (valid && q < kBatchTileMeshes) {
                uint32_t* o = ord + q * 6;
                atomicMin(o + 0, f2ord(bx[k].lx)); atomicMin(o + 1, f2ord(bx[k].ly));
                atomicMin(o + 2, f2ord(bx[k].lz)); atomicMax(o + 3, f2ord(bx[k].hx));
                atomicMax(o + 4, f2ord(bx[k].hy)); atomicMax(o + 5, f2ord(bx[k].hz));
            }
        }
        __syncthreads();
        f4v* t4 = reinterpret_cast<f4v*>(tile);
#pragma unroll
        for (uint32_t k = 0; k < PER; ++k) {
            const uint32_t j = k * kBatchThreads + tid;
            if (j < count) {
                t4[2 * j] = f4v{bx[k].lx, bx[k].ly, bx[k].lz, bx[k].hx};
                t4[2 * j + 1] = f4v{bx[k].hy, bx[k].hz, __uint_as_float(kInvalid), __uint_as_float(local[k])};
            }
        }
        __syncthreads();
        f4v* dst = reinterpret_cast<f4v*>(nodes + base);
#pragma unroll
        for (uint32_t k = 0; k < 2 * PER; ++k) {
            const uint32_t q = k * kBatchThreads + tid;
            if (q < 2 * count) dst[q] = t4[q];
        }
        // meshes of the tile reduced in LDS: one global atomic per bound
        for (uint32_t k = tid; k < kBatchTileMeshes * 6; k += kBatchThreads) {
            const uint32_t q = k / 6, c = k % 6;
            if (lo_local[q] < base + count && lo_local[q + 1] > base && m0 + q < meshes) {
                uint32_t* o = bounds_ord + (size_t)(m0 + q) * 6 + c;
                if (c < 3) atomicMin(o, ord[k]);
                else atomicMax(o, ord[k]);
            }
        }
    }
    if (bad) atomic_or_agent(flags, kFlagNonFinite);
}

__global__ void k_batch_bounds_init(uint32_t* bounds_ord, uint32_t meshes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < meshes * 6) bounds_ord[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

__device__ __forceinline__ uint32_t spread10_b(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ uint32_t quantize_b(float v, float scale) {
    float q = v * scale;
    if (!(q > 0.0f)) return 0u;
    q = fminf(q, scale);
    return (uint32_t)q;
}

// ComputeMortonCodesKernel (Setup.cu:42-59) with 32-bit codes against each
// mesh's own box; key = mesh << 30 | code (codes have 30 bits, so the sort
// window holds no always-zero bits), fused with the digit counts of every
// pass of the sort plan (begin_bit, end_bit, passes; BITS per digit), as the
// single-mesh Morton kernel does: a grid-stride loop counts into LDS, and each
// block adds its counts once into histogram copy blockIdx % kHistCopies
// (zeroed beforehand). The first threads also write the mesh boxes as
// nxb_aabb.
constexpr uint32_t kBatchHistMaxPasses = kHistCopyWords / 512;

template <int BITS>
__global__ __launch_bounds__(kBatchThreads) void k_batch_morton_hist(
    const nxb_bvh2_node* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ mesh_lo, uint32_t meshes,
    const uint32_t* __restrict__ bounds_ord, uint64_t* __restrict__ keys, nxb_aabb* scenes, uint32_t begin_bit,
    uint32_t end_bit, uint32_t passes, uint32_t* hist) {
    constexpr uint32_t RADIX = 1u << BITS;
    __shared__ uint32_t sh[kBatchHistMaxPasses * RADIX];
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += kBatchThreads) sh[i] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * kBatchThreads;
    const uint32_t g0 = blockIdx.x * kBatchThreads + threadIdx.x;
    for (uint32_t i = g0; i < meshes; i += stride) {
        const uint32_t* o = bounds_ord + (size_t)i * 6;
        scenes[i] = nxb_aabb{{ord2f(o[0]), ord2f(o[1]), ord2f(o[2])}, {ord2f(o[3]), ord2f(o[4]), ord2f(o[5])}};
    }
    for (uint32_t i = g0; i < n; i += stride) {
        // Most waves lie in one mesh: find it once from the wave's first
        // position (wave-uniform, scalar loads); a wave that crosses a mesh
        // boundary searches per lane.
        const uint32_t first = __builtin_amdgcn_readfirstlane(i);
        const uint32_t last = min(n, first + 64u) - 1u;
        uint32_t m = mesh_of(mesh_lo, meshes, first);
        if (last >= mesh_lo[m + 1]) m = mesh_of(mesh_lo, meshes, i);
        const uint32_t* o = bounds_ord + (size_t)m * 6;
        const float slx = ord2f(o[0]), sly = ord2f(o[1]), slz = ord2f(o[2]);
        const float ex = ord2f(o[3]) - slx, ey = ord2f(o[4]) - sly, ez = ord2f(o[5]) - slz;
        Box b = load_box(nodes, i);
        float cx = (b.lx + b.hx) * 0.5f;
        float cy = (b.ly + b.hy) * 0.5f;
        float cz = (b.lz + b.hz) * 0.5f;
        float nx = (cx - slx) / ex;
        float ny = (cy - sly) / ey;
        float nz = (cz - slz) / ez;
        uint32_t code = spread10_b(quantize_b(nx, 1023.0f)) | (spread10_b(quantize_b(ny, 1023.0f)) << 1) |
                        (spread10_b(quantize_b(nz, 1023.0f)) << 2);
        const uint64_t key = ((uint64_t)m << 30) | code;
        keys[i] = key;
        for (uint32_t p = 0; p < passes; ++p) {
            const uint32_t shift = begin_bit + p * BITS;
            const uint32_t w = min((uint32_t)BITS, end_bit - shift);
            hist_add(sh + p * RADIX, (uint32_t)(key >> shift) & ((1u << w) - 1u));
        }
    }
    __syncthreads();
    uint32_t* copy = hist + (blockIdx.x % kHistCopies) * kHistCopyWords;
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += kBatchThreads) {
        const uint32_t c = sh[i];
        if (c) atomic_add_agent(&copy[i], c);
    }
}

// Seeds the BVH8 work array with one root per mesh: the root of mesh m is
// its last internal node, n + lo - m + cnt - 2, and owns primIdx from lo.
__global__ void k_batch_bvh8_init(unsigned long long* ctr, uint32_t* grab, uint64_t* work,
                                  const uint32_t* __restrict__ mesh_lo, uint32_t meshes, uint32_t n) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0) {
        *ctr = meshes;  // allocated = meshes, finished = 0
        *grab = 0;
    }
    if (m < meshes) {
        const uint32_t lo = mesh_lo[m], cnt = mesh_lo[m + 1] - lo;
        work[m] = ((uint64_t)lo << 32) | (n + lo - m + cnt - 2);
    }
}

// Mesh of every built BVH8 node (from the first primitive of its subtree,
// kept in the high half of its work slot); padding slots get `meshes`.
__global__ void k_batch_node_mesh(const uint64_t* __restrict__ work, const unsigned long long* ctr, uint32_t capacity,
                                  const uint32_t* __restrict__ mesh_lo, uint32_t meshes, uint32_t* keys) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= capacity) return;
    const uint32_t total = (uint32_t)*ctr;
    keys[k] = k < total ? mesh_of(mesh_lo, meshes, (uint32_t)(work[k] >> 32)) : meshes;
}

// First position of each mesh in the node keys sorted by mesh (every mesh
// has at least one node); start[meshes] = number of nodes.
__global__ void k_batch_starts(const uint32_t* __restrict__ sorted, uint32_t capacity, uint32_t meshes,
                               uint32_t* start) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= capacity) return;
    const uint32_t k = sorted[p];
    if ((p == 0 || sorted[p - 1] != k) && k <= meshes) start[k] = p;
    if (p == capacity - 1 && k != meshes) start[meshes] = capacity;
}

__global__ void k_batch_old2new(const uint32_t* __restrict__ sorted_old, const unsigned long long* ctr,
                                uint32_t* old2new) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)*ctr) old2new[sorted_old[p]] = p;
}

// Node p of the compacted array is old node sorted_old[p] of mesh
// sorted_mesh[p]; its child base becomes mesh-local, its primitive base too.
__global__ void k_batch_compact(const nxb_bvh8_node* __restrict__ src, const uint32_t* __restrict__ sorted_old,
                                const uint32_t* __restrict__ sorted_mesh, const unsigned long long* ctr,
                                const uint32_t* __restrict__ old2new, const uint32_t* __restrict__ start,
                                const uint32_t* __restrict__ mesh_lo, uint32_t meshes, nxb_bvh8_node* dst,
                                uint32_t* flags) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)*ctr) return;
    const uint32_t old = sorted_old[p], m = sorted_mesh[p];
    if (m >= meshes) { atomic_or_agent(flags, kFlagBadIndex); return; }
    const uint4* s = reinterpret_cast<const uint4*>(src + old);
    uint4 w0 = s[0], w1 = s[1], w2 = s[2], w3 = s[3], w4 = s[4];
    const uint32_t imask = w0.w >> 24;
    if (imask) w1.x = old2new[w1.x] - start[m];
    // leaf slots: meta byte set, not an inner slot
    uint32_t leaves = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t meta = ((i < 4 ? w1.z : w1.w) >> (8 * (i & 3))) & 0xFFu;
        if (meta && !(imask & (1u << i))) ++leaves;
    }
    if (leaves) w1.y -= mesh_lo[m];
    uint4* d = reinterpret_cast<uint4*>(dst + p);
    d[0] = w0; d[1] = w1; d[2] = w2; d[3] = w3; d[4] = w4;
}

namespace {

nxb_status from_hip_b(hipError_t e) {
    if (e == hipSuccess) return NXB_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return NXB_ERROR_OUT_OF_MEMORY;
    return NXB_ERROR_HIP;
}
#define NXB_HIPB(expr)                                      \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return from_hip_b(e_);        \
    } while (0)

uint32_t bit_width(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }
size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

}  // namespace nxb

using namespace nxb;

struct nxb_batch_plan {
    int device = 0;
    const void* prims = nullptr;
    nxb_prim_type type = NXB_PRIM_TRIANGLE;
    uint32_t meshes = 0, n = 0, cap8 = 0, lanes = 0, spin = 0, levels = 0;
    uint32_t max_levels = 0;   // BVH8 levels of the deepest possible forest
    bool depth_known = false;  // levels was fitted to the forest of a run
    uint32_t key_end = 0, node_key_bits = 0, digit_bits = 7;
    nxb_build_config cfg{};
    std::vector<uint32_t> h_lo;        // mesh_count + 1
    std::vector<uint32_t> h_first;     // per-mesh first node in nodes8
    std::vector<uint32_t> h_count;     // per-mesh BVH8 node counts of the last run
    uint32_t node_total = 0;           // sum of h_count
    std::vector<nxb_aabb> h_scene;     // per-mesh boxes of the last run
    uint8_t* ws = nullptr;             // workspace
    nxb_bvh8_node* nodes8 = nullptr;   // output nodes (mesh m from h_first[m])
    uint32_t* prim_idx = nullptr;      // output primIdx (mesh m at lo_m, mesh-local values)
    // carved workspace
    uint32_t *mesh_lo, *bounds_ord, *flags, *parent, *recs, *csizes, *nsize, *moved_parent, *moved_size;
    nxb_aabb* scenes;
    nxb_bvh2_node *nodes2, *moved;
    uint64_t *keys_a, *keys_b, *keys_tmp, *work, *work2;
    uint32_t *vals_b, *vals_tmp, *hist, *counts, *start, *old2new, *nkeys, *nkeys_out, *nvals_out;
    nxb_bvh8_node* tmp8;
    unsigned long long* ctr;
    uint32_t* grab;
    uint32_t* level_cnt;  // level counters of the level-synchronous BVH8
    uint64_t* work3;      // (mesh << 32 | node id) per queue position (level-synchronous BVH8)
    uint4* info;          // per mesh: first primitive, primitive count, first node, node capacity
    uint32_t* ids;        // per-mesh node counters (level-synchronous BVH8)
    uint32_t* pslots;     // per-mesh primIdx slot counters (level-synchronous BVH8)
    size_t counts_words = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

namespace {

// Enqueues every stage on s (no host synchronisation, no memset nodes).
nxb_status enqueue_batch(nxb_batch_plan& p, hipStream_t s, bool events) {
    const uint32_t n = p.n, M = p.meshes;
    if (events) NXB_HIPB(hipEventRecord(p.ev[0], s));
    // flags, then the H-PLOC chunk counters from word kBatchHplocChunks
    NXB_HIPB(launch_fill_u32(p.flags, 0u, kBatchHplocChunks + kHplocCounterWords, s));
    // the sort's digit histograms, counted by the Morton kernel
    const SortPlan pl = plan_sort(n, 64, 2, p.key_end, p.digit_bits);
    if (pl.passes > kBatchHistMaxPasses) return NXB_ERROR_INVALID_VALUE;
    SortScratch ss{p.hist, p.counts, p.counts_words};
    NXB_HIPB(launch_sort_prepare(pl, ss, s));
    NXB_HIPB(launch_fill_u32(p.parent, kInvalid, n, s));
    hipLaunchKernelGGL(k_batch_bounds_init, dim3((M * 6 + 255) / 256), dim3(256), 0, s, p.bounds_ord, M);
    const uint32_t gb = min((n + kBatchBoundsTile - 1) / kBatchBoundsTile, 2048u);
    if (p.type == NXB_PRIM_TRIANGLE)
        hipLaunchKernelGGL(k_batch_bounds<9>, dim3(gb), dim3(kBatchThreads), 0, s,
                           static_cast<const float*>(p.prims), n, p.mesh_lo, M, p.nodes2, p.bounds_ord, p.flags);
    else
        hipLaunchKernelGGL(k_batch_bounds<6>, dim3(gb), dim3(kBatchThreads), 0, s,
                           static_cast<const float*>(p.prims), n, p.mesh_lo, M, p.nodes2, p.bounds_ord, p.flags);
    NXB_HIPB(hipGetLastError());
    if (events) NXB_HIPB(hipEventRecord(p.ev[1], s));
    // Morton keys with the digit counts of every sort pass (one block per
    // 3072 keys, 2048 to 8192 blocks, as for a single mesh)
    const uint32_t gm = min((n + kBatchThreads - 1) / kBatchThreads, max(2048u, min(n / 3072u, 8192u)));
    switch (pl.bits) {
        case 6: hipLaunchKernelGGL(k_batch_morton_hist<6>, dim3(gm), dim3(kBatchThreads), 0, s, p.nodes2, n, p.mesh_lo, M, p.bounds_ord, p.keys_a, p.scenes, pl.begin, pl.end, pl.passes, p.hist); break;
        case 7: hipLaunchKernelGGL(k_batch_morton_hist<7>, dim3(gm), dim3(kBatchThreads), 0, s, p.nodes2, n, p.mesh_lo, M, p.bounds_ord, p.keys_a, p.scenes, pl.begin, pl.end, pl.passes, p.hist); break;
        case 8: hipLaunchKernelGGL(k_batch_morton_hist<8>, dim3(gm), dim3(kBatchThreads), 0, s, p.nodes2, n, p.mesh_lo, M, p.bounds_ord, p.keys_a, p.scenes, pl.begin, pl.end, pl.passes, p.hist); break;
        case 9: hipLaunchKernelGGL(k_batch_morton_hist<9>, dim3(gm), dim3(kBatchThreads), 0, s, p.nodes2, n, p.mesh_lo, M, p.bounds_ord, p.keys_a, p.scenes, pl.begin, pl.end, pl.passes, p.hist); break;
        default: return NXB_ERROR_INVALID_VALUE;
    }
    NXB_HIPB(hipGetLastError());
    if (events) NXB_HIPB(hipEventRecord(p.ev[2], s));
    NXB_HIPB(launch_sort_passes(p.keys_a, nullptr, p.keys_b, p.vals_b, p.keys_tmp, p.vals_tmp, n, pl, ss, s));
    if (events) NXB_HIPB(hipEventRecord(p.ev[3], s));
    RadiusParams rp{p.cfg.search_radius, nullptr, 12, 0};
    // subtree sizes only for the persistent BVH8 kernel (the level-synchronous
    // collapse takes primIdx slots from per-mesh counters)
    HplocScratch hs{p.parent, p.recs, p.csizes, p.moved, p.moved_parent, p.moved_size, p.lanes ? p.nsize : nullptr,
                    p.flags, p.flags + kBatchHplocChunks};
    NXB_HIPB(launch_hploc_batch(p.keys_b, p.vals_b, n, p.mesh_lo, M, p.nodes2, p.cfg.merge_mode, rp,
                                p.cfg.hploc_blocks_per_cu, hs, s));
    if (events) NXB_HIPB(hipEventRecord(p.ev[4], s));
    Bvh8Scratch bs{p.work, p.ctr, p.grab, p.flags, p.nsize, p.work2, p.cfg.slot_assignment};
    if (p.lanes == 0) {
        // every queue position a level reads was written by the level above
        NXB_HIPB(launch_bvh8_tree_levels(p.nodes2, n, M, p.nodes8, p.prim_idx, p.cap8, p.levels, bs, p.level_cnt,
                                         p.work3, p.info, p.mesh_lo, p.ids, p.pslots, s));
        if (events) NXB_HIPB(hipEventRecord(p.ev[5], s));
        return NXB_SUCCESS;
    }
    NXB_HIPB(launch_fill_u32(reinterpret_cast<uint32_t*>(p.work), kInvalid, 2ull * p.cap8, s));
    hipLaunchKernelGGL(k_batch_bvh8_init, dim3((M + 255) / 256), dim3(256), 0, s, p.ctr, p.grab, p.work, p.mesh_lo,
                       M, n);
    NXB_HIPB(launch_bvh8_forest(p.nodes2, n, M, p.tmp8, p.prim_idx, p.cap8, p.lanes, p.spin, bs, s));
    const uint32_t gc = (p.cap8 + 255) / 256;
    hipLaunchKernelGGL(k_batch_node_mesh, dim3(gc), dim3(256), 0, s, p.work, p.ctr, p.cap8, p.mesh_lo, M, p.nkeys);
    NXB_HIPB(launch_radix_sort(p.nkeys, nullptr, p.nkeys_out, p.nvals_out, p.keys_tmp, p.vals_tmp, p.cap8, 32, 0,
                               p.node_key_bits, 8, ss, s));
    hipLaunchKernelGGL(k_batch_starts, dim3(gc), dim3(256), 0, s, p.nkeys_out, p.cap8, M, p.start);
    hipLaunchKernelGGL(k_batch_old2new, dim3(gc), dim3(256), 0, s, p.nvals_out, p.ctr, p.old2new);
    hipLaunchKernelGGL(k_batch_compact, dim3(gc), dim3(256), 0, s, p.tmp8, p.nvals_out, p.nkeys_out, p.ctr, p.old2new,
                       p.start, p.mesh_lo, M, p.nodes8, p.flags);
    NXB_HIPB(hipGetLastError());
    if (events) NXB_HIPB(hipEventRecord(p.ev[5], s));
    return NXB_SUCCESS;
}

// (Re-)records every stage into the plan's graph (p.levels BVH8 levels).
nxb_status record_graph(nxb_batch_plan& p) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    p.exec = nullptr;
    p.graph = nullptr;
    hipError_t e = hipStreamBeginCapture(p.capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return from_hip_b(e);
    nxb_status st = enqueue_batch(p, p.capture_stream, false);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(p.capture_stream, &graph);
    p.graph = graph;
    if (st != NXB_SUCCESS) return st;
    if (e == hipSuccess) e = hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0);
    return from_hip_b(e);
}

// BVH8 levels recorded once a run has shown the forest's depth: the empty
// levels past it cost a launch each (the bound for any forest is 32 plus the
// bit width of the largest mesh, ~49, against ~20 for typical meshes).
constexpr uint32_t kBatchLevelMargin = 2;

void release(nxb_batch_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->ws) (void)hipFree(p->ws);
    if (p->nodes8) (void)hipFree(p->nodes8);
    if (p->prim_idx) (void)hipFree(p->prim_idx);
    delete p;
}

}  // namespace

extern "C" {

nxb_status nxb_batch_plan_create(const void* d_prims, nxb_prim_type type, uint32_t mesh_count,
                                 const uint32_t* h_offsets, const nxb_build_config* cfgp, uint32_t bvh8_lanes,
                                 nxb_batch_plan** out) {
    if (!out) return NXB_ERROR_INVALID_VALUE;
    *out = nullptr;
    if (!d_prims || !h_offsets || mesh_count == 0) return NXB_ERROR_INVALID_VALUE;
    if (type != NXB_PRIM_TRIANGLE && type != NXB_PRIM_AABB) return NXB_ERROR_INVALID_VALUE;
    if (h_offsets[0] != 0) return NXB_ERROR_INVALID_VALUE;
    for (uint32_t m = 0; m < mesh_count; ++m)
        if (h_offsets[m + 1] < h_offsets[m] + 2) return NXB_ERROR_INVALID_VALUE;  // every mesh >= 2 primitives
    const uint32_t n = h_offsets[mesh_count];
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    nxb_build_config cfg;
    nxb_default_config(&cfg);
    if (cfgp) cfg = *cfgp;
    const uint32_t bits = cfg.morton_bits ? cfg.morton_bits : (cfg.prioritize_speed ? 32u : 64u);
    if (bits != 32) return NXB_ERROR_NOT_SUPPORTED;  // composite keys carry 32-bit codes
    if (cfg.search_radius < 1 || cfg.search_radius > 16) return NXB_ERROR_INVALID_VALUE;
    if (cfg.merge_mode != NXB_MERGE_WAVE32 && cfg.merge_mode != NXB_MERGE_WAVE64) return NXB_ERROR_INVALID_VALUE;
    if (cfg.radius_policy != NXB_RADIUS_FIXED) return NXB_ERROR_NOT_SUPPORTED;
    if (cfg.sort_digit_bits != 0 && (cfg.sort_digit_bits < 6 || cfg.sort_digit_bits > 9))
        return NXB_ERROR_INVALID_VALUE;
    if (cfg.slot_assignment != NXB_SLOT_GREEDY && cfg.slot_assignment != NXB_SLOT_AUCTION)
        return NXB_ERROR_INVALID_VALUE;
    if (cfg.hploc_blocks_per_cu > kHplocMaxBlocksPerCu) return NXB_ERROR_INVALID_VALUE;

    nxb_batch_plan* p = new nxb_batch_plan();
    p->prims = d_prims;
    p->type = type;
    p->meshes = mesh_count;
    p->n = n;
    p->cfg = cfg;
    p->h_lo.assign(h_offsets, h_offsets + mesh_count + 1);
    p->h_scene.assign(mesh_count, nxb_aabb{});
    p->spin = cfg.spin_limit ? cfg.spin_limit : (1u << 22);
    p->key_end = 30 + max(1u, bit_width(mesh_count - 1));
    // unless set: 7-bit digits, or 8-bit ones where that takes a pass fewer
    // (bits [2, key_end): 39 of them for 2,048 meshes, 6 passes of 7 or 5 of 8)
    const uint32_t window = p->key_end - 2;
    p->digit_bits = cfg.sort_digit_bits ? cfg.sort_digit_bits
                                        : ((window + 7) / 8 < (window + 6) / 7 ? 8u : 7u);
    p->node_key_bits = max(1u, bit_width(mesh_count));  // node keys go up to mesh_count (padding)
    uint64_t cap = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) cap += nxb_bvh8_capacity(h_offsets[m + 1] - h_offsets[m]);
    if (cap > 0xFFFFFFF0ull) { delete p; return NXB_ERROR_TOO_LARGE; }
    p->cap8 = (uint32_t)cap;
    // bvh8_lanes == 0: level-synchronous BVH8 with enough levels for the
    // deepest possible tree (each BVH8 level consumes at least one BVH2
    // level; an LBVH over 32-bit codes of m primitives is at most 30 +
    // ceil(log2 m) deep); else the persistent kernel with that many lanes
    uint32_t max_cnt = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) max_cnt = max(max_cnt, h_offsets[m + 1] - h_offsets[m]);
    p->lanes = bvh8_lanes;
    p->max_levels = min(kBvh8LevelWords - 1, 32u + bit_width(max_cnt));
    p->levels = p->max_levels;

    hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) { release(p); return from_hip_b(e); }
    // workspace layout (the compaction buffers only for the persistent
    // kernel, the per-mesh queue words only for the level-synchronous one)
    const size_t N = n, M = mesh_count, C8 = p->cap8;
    const bool compact = bvh8_lanes != 0;
    const size_t sort_n = compact ? max(N, C8) : N, CC = compact ? C8 : 0;
    p->counts_words = sort_scratch_bytes((uint32_t)sort_n, 64, 9) / 4;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += al(bytes ? bytes : 1); return o; };
    size_t o_lo = take((M + 1) * 4), o_bounds = take(M * 24), o_scenes = take(M * sizeof(nxb_aabb)),
           o_flags = take((kBatchHplocChunks + kHplocCounterWords) * 4), o_nodes2 = take((2 * N - M) * sizeof(nxb_bvh2_node)),
           o_moved = take(M * sizeof(nxb_bvh2_node)), o_mpar = take(M * 4), o_msz = take(M * 4),
           o_ka = take(N * 8), o_kb = take(N * 8), o_kt = take(sort_n * 8), o_vb = take(N * 4),
           o_vt = take(sort_n * 4), o_hist = take(kHistCopies * kHistCopyWords * 4),
           o_counts = take(p->counts_words * 4), o_parent = take(N * 4), o_recs = take(N * 32),
           o_csz = take(N * 4), o_nsize = take(N * 4), o_work = take(C8 * 8), o_work2 = take(C8 * 8), o_ctr = take(64), o_grab = take(64), o_lvl = take(kBvh8LevelWords * 4),
           o_tmp8 = take(CC * sizeof(nxb_bvh8_node)), o_nkeys = take(CC * 4), o_nko = take(CC * 4),
           o_nvo = take(CC * 4), o_start = take((M + 1) * 4), o_o2n = take(CC * 4),
           o_work3 = take(compact ? 0 : C8 * 8), o_info = take(M * sizeof(uint4)), o_ids = take(M * 4),
           o_pslots = take(M * 4);
    e = hipMalloc((void**)&p->ws, off);
    if (e == hipSuccess) e = hipMalloc((void**)&p->nodes8, C8 * sizeof(nxb_bvh8_node));
    if (e == hipSuccess) e = hipMalloc((void**)&p->prim_idx, N * 4);
    if (e != hipSuccess) { release(p); return from_hip_b(e); }
    uint8_t* b = p->ws;
    p->mesh_lo = reinterpret_cast<uint32_t*>(b + o_lo);
    p->bounds_ord = reinterpret_cast<uint32_t*>(b + o_bounds);
    p->scenes = reinterpret_cast<nxb_aabb*>(b + o_scenes);
    p->flags = reinterpret_cast<uint32_t*>(b + o_flags);
    p->nodes2 = reinterpret_cast<nxb_bvh2_node*>(b + o_nodes2);
    p->moved = reinterpret_cast<nxb_bvh2_node*>(b + o_moved);
    p->moved_parent = reinterpret_cast<uint32_t*>(b + o_mpar);
    p->moved_size = reinterpret_cast<uint32_t*>(b + o_msz);
    p->keys_a = reinterpret_cast<uint64_t*>(b + o_ka);
    p->keys_b = reinterpret_cast<uint64_t*>(b + o_kb);
    p->keys_tmp = reinterpret_cast<uint64_t*>(b + o_kt);
    p->vals_b = reinterpret_cast<uint32_t*>(b + o_vb);
    p->vals_tmp = reinterpret_cast<uint32_t*>(b + o_vt);
    p->hist = reinterpret_cast<uint32_t*>(b + o_hist);
    p->counts = reinterpret_cast<uint32_t*>(b + o_counts);
    p->parent = reinterpret_cast<uint32_t*>(b + o_parent);
    p->recs = reinterpret_cast<uint32_t*>(b + o_recs);
    p->csizes = reinterpret_cast<uint32_t*>(b + o_csz);
    p->nsize = reinterpret_cast<uint32_t*>(b + o_nsize);
    p->work = reinterpret_cast<uint64_t*>(b + o_work);
    p->work2 = reinterpret_cast<uint64_t*>(b + o_work2);
    p->ctr = reinterpret_cast<unsigned long long*>(b + o_ctr);
    p->level_cnt = reinterpret_cast<uint32_t*>(b + o_lvl);
    p->grab = reinterpret_cast<uint32_t*>(b + o_grab);
    p->tmp8 = reinterpret_cast<nxb_bvh8_node*>(b + o_tmp8);
    p->nkeys = reinterpret_cast<uint32_t*>(b + o_nkeys);
    p->nkeys_out = reinterpret_cast<uint32_t*>(b + o_nko);
    p->nvals_out = reinterpret_cast<uint32_t*>(b + o_nvo);
    p->start = reinterpret_cast<uint32_t*>(b + o_start);
    p->old2new = reinterpret_cast<uint32_t*>(b + o_o2n);
    p->work3 = reinterpret_cast<uint64_t*>(b + o_work3);
    p->info = reinterpret_cast<uint4*>(b + o_info);
    p->ids = reinterpret_cast<uint32_t*>(b + o_ids);
    p->pslots = reinterpret_cast<uint32_t*>(b + o_pslots);
    // mesh m's nodes: region [first, first + capacity) of nodes8
    std::vector<uint4> h_info(M);
    p->h_first.assign(M, 0);
    p->h_count.assign(M, 0);
    uint32_t first = 0;
    for (uint32_t m = 0; m < mesh_count; ++m) {
        const uint32_t cnt = h_offsets[m + 1] - h_offsets[m], c8 = nxb_bvh8_capacity(cnt);
        h_info[m] = make_uint4(h_offsets[m], cnt, first, c8);
        p->h_first[m] = first;
        first += c8;
    }
    e = hipMemcpy(p->mesh_lo, h_offsets, (M + 1) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->info, h_info.data(), M * sizeof(uint4), hipMemcpyHostToDevice);
    for (auto& x : p->ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableSystemFence);  // timing only
    if (e != hipSuccess) { release(p); return from_hip_b(e); }

    // Record the whole sequence into one graph.
    e = hipStreamCreate(&p->capture_stream);
    if (e != hipSuccess) { release(p); return from_hip_b(e); }
    nxb_status st = record_graph(*p);
    if (st != NXB_SUCCESS) {
        release(p);
        return st;
    }
    *out = p;
    return NXB_SUCCESS;
}

nxb_status nxb_batch_plan_run(nxb_batch_plan* p, nxb_stream_t stream, float* stage_ms) {
    if (!p) return NXB_ERROR_INVALID_VALUE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (stage_ms) {
        nxb_status st = enqueue_batch(*p, s, true);
        if (st) return st;
    } else {
        NXB_HIPB(hipGraphLaunch(p->exec, s));
    }
    uint32_t flags[4];
    uint32_t level_cnt[kBvh8LevelWords];
    const bool fit_levels = p->lanes == 0 && !p->depth_known;
    // node counts (level-synchronous) or compacted starts (persistent kernel),
    // kept only if the run succeeds
    std::vector<uint32_t> nodes(p->meshes + 1, 0);
    if (p->lanes == 0)
        NXB_HIPB(hipMemcpyAsync(nodes.data(), p->ids, p->meshes * 4, hipMemcpyDeviceToHost, s));
    else
        NXB_HIPB(hipMemcpyAsync(nodes.data(), p->start, (p->meshes + 1) * 4, hipMemcpyDeviceToHost, s));
    NXB_HIPB(hipMemcpyAsync(p->h_scene.data(), p->scenes, p->meshes * sizeof(nxb_aabb), hipMemcpyDeviceToHost, s));
    NXB_HIPB(hipMemcpyAsync(flags, p->flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    if (fit_levels) NXB_HIPB(hipMemcpyAsync(level_cnt, p->level_cnt, sizeof(level_cnt), hipMemcpyDeviceToHost, s));
    NXB_HIPB(hipStreamSynchronize(s));
    if (flags[0] & kFlagNonFinite) return NXB_ERROR_NONFINITE_INPUT;
    if (flags[0] & kFlagTimeout) return NXB_ERROR_TIMEOUT;
    if (flags[0] & kFlagBadIndex) return NXB_ERROR_HIP;
    if (flags[0] & kFlagDeeper) {
        // the primitives changed since the levels were fitted: record the
        // bound for any forest again and build once more
        if (p->levels >= p->max_levels) return NXB_ERROR_HIP;
        p->levels = p->max_levels;
        p->depth_known = true;
        nxb_status st = record_graph(*p);
        if (st) return st;
        return nxb_batch_plan_run(p, stream, stage_ms);
    }
    if (fit_levels) {
        // the first kTreeTopLevels levels run in one launch and count nothing
        uint32_t depth = kTreeTopLevels;
        while (depth < p->levels && level_cnt[depth] != 0) ++depth;
        p->depth_known = true;
        const uint32_t fitted = min(p->max_levels, depth + kBatchLevelMargin);
        if (fitted < p->levels) {
            p->levels = fitted;
            nxb_status st = record_graph(*p);
            if (st) return st;
        }
    }
    if (stage_ms) {
        for (int i = 0; i < 5; ++i) NXB_HIPB(hipEventElapsedTime(&stage_ms[i], p->ev[i], p->ev[i + 1]));
        NXB_HIPB(hipEventElapsedTime(&stage_ms[5], p->ev[0], p->ev[5]));
    }
    uint32_t total = 0;
    for (uint32_t m = 0; m < p->meshes; ++m) {
        if (p->lanes != 0) {
            p->h_first[m] = nodes[m];
            nodes[m] = nodes[m + 1] - nodes[m];
        }
        p->h_count[m] = nodes[m];
        total += nodes[m];
    }
    p->node_total = total;
    return NXB_SUCCESS;
}

nxb_status nxb_batch_plan_mesh(const nxb_batch_plan* p, uint32_t mesh, nxb_bvh8* out) {
    if (!p || !out || mesh >= p->meshes) return NXB_ERROR_INVALID_VALUE;
    out->nodes = p->nodes8 + p->h_first[mesh];
    out->node_count = p->h_count[mesh];
    out->prim_idx = p->prim_idx + p->h_lo[mesh];
    out->prim_count = p->h_lo[mesh + 1] - p->h_lo[mesh];
    out->bounds = p->h_scene[mesh];
    return NXB_SUCCESS;
}

uint32_t nxb_batch_plan_node_total(const nxb_batch_plan* p) { return p ? p->node_total : 0u; }

void nxb_batch_plan_destroy(nxb_batch_plan* p) { release(p); }

}  // extern "C"
