// setup.hip - scene bounds + leaf initialisation, Morton codes.
//
// ComputeSceneBoundsKernel (NexusBVH/src/Cuda/Setup.cu:12-39) and
// ComputeMortonCodesKernel (Setup.cu:42-59), re-designed for gfx950:
//  * bounds: each 256-thread block stages tiles of 1024 primitives through
//    LDS with coalesced 16-byte loads, writes the 32-byte leaf nodes back
//    through LDS as contiguous 16-byte stores, and reduces its box in
//    registers/LDS. The scene box is
//    combined by a last-block-done reduction over per-block partials (one
//    agent-scope release/acquire per block) instead of per-warp float CAS
//    loops on six global words.
//  * Morton: one pass over the leaf boxes, writes only the keys; the iota
//    values are synthesised by the first radix-sort pass.
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

constexpr int kBoundsThreads = 256;
constexpr int kBoundsTile = 1024;  // primitives per tile (4 per thread)

__device__ __forceinline__ Box wave_reduce_box(Box b) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        Box o;
        o.lx = __shfl_xor(b.lx, off); o.ly = __shfl_xor(b.ly, off); o.lz = __shfl_xor(b.lz, off);
        o.hx = __shfl_xor(b.hx, off); o.hy = __shfl_xor(b.hy, off); o.hz = __shfl_xor(b.hz, off);
        b = box_union(b, o);
    }
    return b;
}

// Per-build scratch layout for the bounds reduction (in BuildScratch):
//   partials[grid] : Box of each block
//   done           : last-block-done counter (left at 0 on exit)
//   indices        : INDEXED only: three vertex indices per triangle into
//                    prims (3 floats per vertex, vertex_count vertices)
template <int PF, bool INDEXED = false>
__global__ __launch_bounds__(kBoundsThreads) void k_scene_bounds(
    const float* __restrict__ prims, uint32_t n, nxb_bvh2_node* __restrict__ nodes,
    Box* __restrict__ partials, uint32_t* done, nxb_aabb* scene, uint32_t* flags,
    const uint32_t* __restrict__ indices, uint32_t vertex_count) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    // primitives in, then the leaf nodes out (8 floats each)
    __shared__ __attribute__((aligned(16))) float tile[kBoundsTile * (PF > 8 ? PF : 8)];
    __shared__ Box wave_box[kBoundsThreads / 64];
    __shared__ uint32_t is_last;
    __shared__ uint32_t any_bad;

    const uint32_t tid = threadIdx.x;
    if (tid == 0) any_bad = 0;
    Box acc = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    bool bad = false;
    const bool aligned = (reinterpret_cast<uintptr_t>(prims) & 15u) == 0;
    const uint32_t tiles = (n + kBoundsTile - 1) / kBoundsTile;

    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint32_t base = t * kBoundsTile;
        const uint32_t count = min((uint32_t)kBoundsTile, n - base);
        const float* src = prims + (size_t)base * PF;
        const uint32_t nf = count * PF;
        __syncthreads();
        if (INDEXED) {
            // triangle j of the tile: its three vertices gathered through the
            // index buffer into the tile's 9 floats (an index past the vertex
            // buffer flags the input and reads vertex 0)
            for (uint32_t j = tid; j < count; j += kBoundsThreads) {
                const uint32_t* ix = indices + (size_t)(base + j) * 3;
#pragma unroll
                for (uint32_t c = 0; c < 3; ++c) {
                    uint32_t v = ix[c];
                    if (v >= vertex_count) {
                        atomic_or_agent(flags, kFlagBadInput);
                        v = 0;
                    }
                    const float* vp = prims + (size_t)v * 3;
                    tile[j * 9 + c * 3 + 0] = vp[0];
                    tile[j * 9 + c * 3 + 1] = vp[1];
                    tile[j * 9 + c * 3 + 2] = vp[2];
                }
            }
        } else if (aligned && count == (uint32_t)kBoundsTile) {
            // kBoundsTile * PF floats is a multiple of 4: whole float4 loads.
            const f4v* s4 = reinterpret_cast<const f4v*>(src);
            f4v* t4 = reinterpret_cast<f4v*>(tile);
#pragma unroll
            for (uint32_t k = 0; k < (kBoundsTile * PF / 4) / kBoundsThreads; ++k)
                t4[k * kBoundsThreads + tid] = __builtin_nontemporal_load(s4 + k * kBoundsThreads + tid);
        } else {
            for (uint32_t k = tid; k < nf; k += kBoundsThreads) tile[k] = src[k];
        }
        __syncthreads();
        // leaf nodes go back through the same LDS tile, so that the global
        // stores are contiguous 16-byte runs (measured 13% faster than each
        // lane storing its own 32-byte node)
        Box bx[kBoundsTile / kBoundsThreads];
#pragma unroll
        for (uint32_t k = 0; k < kBoundsTile / kBoundsThreads; ++k) {
            uint32_t j = k * kBoundsThreads + tid;
            if (j < count) {
                bx[k] = prim_box<PF>(tile + j * PF, bad);
                acc = box_union(acc, bx[k]);
            }
        }
        __syncthreads();
        f4v* t4 = reinterpret_cast<f4v*>(tile);
#pragma unroll
        for (uint32_t k = 0; k < kBoundsTile / kBoundsThreads; ++k) {
            uint32_t j = k * kBoundsThreads + tid;
            if (j < count) {
                t4[2 * j] = f4v{bx[k].lx, bx[k].ly, bx[k].lz, bx[k].hx};
                t4[2 * j + 1] = f4v{bx[k].hy, bx[k].hz, __uint_as_float(kInvalid), __uint_as_float(base + j)};
            }
        }
        __syncthreads();
        f4v* dst = reinterpret_cast<f4v*>(nodes + base);
#pragma unroll
        for (uint32_t k = 0; k < 2 * kBoundsTile / kBoundsThreads; ++k) {
            uint32_t q = k * kBoundsThreads + tid;
            if (q < 2 * count) dst[q] = t4[q];
        }
    }

    acc = wave_reduce_box(acc);
    if ((tid & 63) == 0) wave_box[tid >> 6] = acc;
    if (bad) any_bad = 1;  // benign race: every writer stores 1
    __syncthreads();
    if (tid == 0) {
        Box b = wave_box[0];
#pragma unroll
        for (int w = 1; w < kBoundsThreads / 64; ++w) b = box_union(b, wave_box[w]);
        // The partial goes out with agent-scope stores and is complete before
        // the counter add, so the last block sees it without the L2
        // writeback of a release fence (one per block was a large share of
        // this kernel's time for ~1M primitives).
        uint64_t* pp = reinterpret_cast<uint64_t*>(partials + blockIdx.x);
        store_agent64(pp, (uint64_t)__float_as_uint(b.lx) | ((uint64_t)__float_as_uint(b.ly) << 32));
        store_agent64(pp + 1, (uint64_t)__float_as_uint(b.lz) | ((uint64_t)__float_as_uint(b.hx) << 32));
        store_agent64(pp + 2, (uint64_t)__float_as_uint(b.hy) | ((uint64_t)__float_as_uint(b.hz) << 32));
        if (any_bad) atomic_or_agent(flags, kFlagNonFinite);
        wait_memory_ops();
        uint32_t prev = atomic_add_agent(done, 1u);
        is_last = (prev == gridDim.x - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (!is_last) return;

    // Last block: combine all partials into the scene box (agent-scope loads).
    Box b = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = tid; i < gridDim.x; i += kBoundsThreads) {
        const uint64_t* pp = reinterpret_cast<const uint64_t*>(partials + i);
        const uint64_t a = load_agent64(pp), c = load_agent64(pp + 1), d = load_agent64(pp + 2);
        Box q;
        q.lx = __uint_as_float((uint32_t)a); q.ly = __uint_as_float((uint32_t)(a >> 32));
        q.lz = __uint_as_float((uint32_t)c); q.hx = __uint_as_float((uint32_t)(c >> 32));
        q.hy = __uint_as_float((uint32_t)d); q.hz = __uint_as_float((uint32_t)(d >> 32));
        b = box_union(b, q);
    }
    b = wave_reduce_box(b);
    __syncthreads();
    if ((tid & 63) == 0) wave_box[tid >> 6] = b;
    __syncthreads();
    if (tid == 0) {
        Box r = wave_box[0];
#pragma unroll
        for (int w = 1; w < kBoundsThreads / 64; ++w) r = box_union(r, wave_box[w]);
        scene->bmin = nxb_float3{r.lx, r.ly, r.lz};
        scene->bmax = nxb_float3{r.hx, r.hy, r.hz};
        *done = 0;  // ready for the next build (kernel boundary orders it)
    }
}

// Spreads the low bits of x so that bit b lands at bit 3b.
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// (uint32)(v * scale) with NaN/negatives -> 0 and a clamp at scale
// (BuilderUtils.h:235-253 truncating cast).
__device__ __forceinline__ uint32_t quantize(float v, float scale) {
    float q = v * scale;
    if (!(q > 0.0f)) return 0u;
    q = fminf(q, scale);
    return (uint32_t)q;
}

template <typename K>
__global__ __launch_bounds__(256) void k_morton(const nxb_bvh2_node* __restrict__ nodes, uint32_t n,
                                                const nxb_aabb* __restrict__ scene, K* __restrict__ keys,
                                                uint32_t* __restrict__ values) {
    const float slx = scene->bmin.x, sly = scene->bmin.y, slz = scene->bmin.z;
    const float ex = scene->bmax.x - slx, ey = scene->bmax.y - sly, ez = scene->bmax.z - slz;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Box b = load_box_stream(nodes, i);
        float cx = (b.lx + b.hx) * 0.5f;
        float cy = (b.ly + b.hy) * 0.5f;
        float cz = (b.lz + b.hz) * 0.5f;
        float nx = (cx - slx) / ex;
        float ny = (cy - sly) / ey;
        float nz = (cz - slz) / ez;
        if (sizeof(K) == 4) {
            uint32_t code = spread10(quantize(nx, 1023.0f)) | (spread10(quantize(ny, 1023.0f)) << 1) |
                            (spread10(quantize(nz, 1023.0f)) << 2);
            keys[i] = (K)code;
        } else {
            uint64_t code = spread21(quantize(nx, 2097151.0f)) | (spread21(quantize(ny, 2097151.0f)) << 1) |
                            (spread21(quantize(nz, 2097151.0f)) << 2);
            keys[i] = (K)code;
        }
        if (values) values[i] = i;
    }
}

constexpr int kHistMaxPasses = 11;

// ComputeMortonCodesKernel fused with the digit histograms of every sort pass.
template <typename K, int BITS>
__global__ __launch_bounds__(256) void k_morton_hist(const nxb_bvh2_node* __restrict__ nodes, uint32_t n,
                                                     const nxb_aabb* __restrict__ scene, K* __restrict__ keys,
                                                     uint32_t begin_bit, uint32_t end_bit, uint32_t passes,
                                                     uint32_t* hist) {
    constexpr uint32_t RADIX = 1u << BITS;
    __shared__ uint32_t sh[kHistMaxPasses * RADIX];
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += 256) sh[i] = 0;
    __syncthreads();
    const float slx = scene->bmin.x, sly = scene->bmin.y, slz = scene->bmin.z;
    const float ex = scene->bmax.x - slx, ey = scene->bmax.y - sly, ez = scene->bmax.z - slz;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Box b = load_box_stream(nodes, i);
        float cx = (b.lx + b.hx) * 0.5f;
        float cy = (b.ly + b.hy) * 0.5f;
        float cz = (b.lz + b.hz) * 0.5f;
        float nx = (cx - slx) / ex;
        float ny = (cy - sly) / ey;
        float nz = (cz - slz) / ez;
        K code;
        if (sizeof(K) == 4)
            code = (K)(spread10(quantize(nx, 1023.0f)) | (spread10(quantize(ny, 1023.0f)) << 1) |
                       (spread10(quantize(nz, 1023.0f)) << 2));
        else
            code = (K)(spread21(quantize(nx, 2097151.0f)) | (spread21(quantize(ny, 2097151.0f)) << 1) |
                       (spread21(quantize(nz, 2097151.0f)) << 2));
        keys[i] = code;
        for (uint32_t p = 0; p < passes; ++p) {
            uint32_t shift = begin_bit + p * BITS;
            uint32_t w = min((uint32_t)BITS, end_bit - shift);
            hist_add(sh + p * RADIX, (uint32_t)(code >> shift) & ((1u << w) - 1u));
        }
    }
    __syncthreads();
    uint32_t* copy = hist + (blockIdx.x % kHistCopies) * kHistCopyWords;
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += 256) {
        uint32_t c = sh[i];
        if (c) atomic_add_agent(&copy[i], c);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t count) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

// The scratch set-up of a build in one launch: header words zeroed, the
// H-PLOC parent table filled with kInvalid, and the first hist_words words of
// every digit-histogram copy zeroed.
__global__ void k_build_prep(uint32_t* hdr, uint32_t hdr_words, uint32_t* parent, uint32_t n, uint32_t* hist,
                             uint32_t hist_words) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t; i < hdr_words; i += stride) hdr[i] = 0;
    const size_t hw = (size_t)kHistCopies * hist_words;
    for (size_t i = t; i < hw; i += stride) hist[(i / hist_words) * kHistCopyWords + i % hist_words] = 0;
    for (size_t i = t; i < n; i += stride) parent[i] = kInvalid;
}

// ---- launchers -------------------------------------------------------------

hipError_t launch_scene_bounds(const void* prims, nxb_prim_type type, uint32_t n, nxb_bvh2_node* nodes,
                               nxb_aabb* scene, uint32_t* flags, BoundsScratch sc, hipStream_t s,
                               const uint32_t* indices, uint32_t vertex_count) {
    uint32_t tiles = (n + kBoundsTile - 1) / kBoundsTile;
    uint32_t grid = min(tiles, sc.max_blocks);
    if (grid == 0) grid = 1;
    const float* p = static_cast<const float*>(prims);
    if (indices)
        hipLaunchKernelGGL((k_scene_bounds<9, true>), dim3(grid), dim3(kBoundsThreads), 0, s, p, n, nodes,
                           sc.partials, sc.done, scene, flags, indices, vertex_count);
    else if (type == NXB_PRIM_TRIANGLE)
        hipLaunchKernelGGL(k_scene_bounds<9>, dim3(grid), dim3(kBoundsThreads), 0, s, p, n, nodes, sc.partials,
                           sc.done, scene, flags, nullptr, 0u);
    else
        hipLaunchKernelGGL(k_scene_bounds<6>, dim3(grid), dim3(kBoundsThreads), 0, s, p, n, nodes, sc.partials,
                           sc.done, scene, flags, nullptr, 0u);
    return hipGetLastError();
}

hipError_t launch_morton(const nxb_bvh2_node* nodes, uint32_t n, const nxb_aabb* scene, uint32_t bits,
                         void* keys, uint32_t* values, hipStream_t s) {
    uint32_t grid = min((n + 255u) / 256u, 4096u);
    if (grid == 0) grid = 1;
    if (bits == 32)
        hipLaunchKernelGGL(k_morton<uint32_t>, dim3(grid), dim3(256), 0, s, nodes, n, scene,
                           static_cast<uint32_t*>(keys), values);
    else
        hipLaunchKernelGGL(k_morton<uint64_t>, dim3(grid), dim3(256), 0, s, nodes, n, scene,
                           static_cast<uint64_t*>(keys), values);
    return hipGetLastError();
}

template <typename K>
static hipError_t morton_hist_bits(const nxb_bvh2_node* nodes, uint32_t n, const nxb_aabb* scene, K* keys,
                                   const SortPlan& pl, uint32_t* hist, hipStream_t s) {
    // one block per 3072 keys, 2048 to 8192 blocks: more blocks were faster
    // at 28.1M keys, fewer at 1.1M, where the per-block histogram flushes
    // weigh more
    uint32_t grid = min((n + 255u) / 256u, max(2048u, min(n / 3072u, 8192u)));
    if (grid == 0) grid = 1;
    switch (pl.bits) {
        case 6: hipLaunchKernelGGL((k_morton_hist<K, 6>), dim3(grid), dim3(256), 0, s, nodes, n, scene, keys, pl.begin, pl.end, pl.passes, hist); break;
        case 7: hipLaunchKernelGGL((k_morton_hist<K, 7>), dim3(grid), dim3(256), 0, s, nodes, n, scene, keys, pl.begin, pl.end, pl.passes, hist); break;
        case 8: hipLaunchKernelGGL((k_morton_hist<K, 8>), dim3(grid), dim3(256), 0, s, nodes, n, scene, keys, pl.begin, pl.end, pl.passes, hist); break;
        case 9: hipLaunchKernelGGL((k_morton_hist<K, 9>), dim3(grid), dim3(256), 0, s, nodes, n, scene, keys, pl.begin, pl.end, pl.passes, hist); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_morton_hist(const nxb_bvh2_node* nodes, uint32_t n, const nxb_aabb* scene, void* keys,
                              const SortPlan& pl, uint32_t* hist, hipStream_t s) {
    if (pl.passes > (uint32_t)kHistMaxPasses) return hipErrorInvalidValue;
    if (pl.key_bits == 32) return morton_hist_bits(nodes, n, scene, static_cast<uint32_t*>(keys), pl, hist, s);
    return morton_hist_bits(nodes, n, scene, static_cast<uint64_t*>(keys), pl, hist, s);
}

hipError_t launch_build_prep(uint32_t* hdr, uint32_t hdr_words, uint32_t* parent, uint32_t n, uint32_t* hist,
                             uint32_t hist_words, hipStream_t s) {
    if (hist_words > kHistCopyWords) return hipErrorInvalidValue;
    size_t most = n;
    if (hdr_words > most) most = hdr_words;
    if ((size_t)kHistCopies * hist_words > most) most = (size_t)kHistCopies * hist_words;
    const size_t blocks = (most + 255) / 256;
    const uint32_t grid = (uint32_t)(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(k_build_prep, dim3(grid), dim3(256), 0, s, hdr, hdr_words, parent, n, hist, hist_words);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    uint32_t grid = (uint32_t)(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, p, v, count);
    return hipGetLastError();
}

}  // namespace nxb
