// sah.hip - exact SAH evaluator.
//
// ComputeBVH2CostKernel / ComputeBVH8CostKernel (NexusBVH/src/Cuda/Eval.cu:
// 12-80). Each per-node (BVH2) or per-child (BVH8) term is the reference's
// FP32 value C * (Area / sceneArea); the terms are summed in FP64 (block
// reduction, one FP64 atomic per block), so the result differs from a
// sequential FP64 sum only by FP64 reassociation.
//   BVH2: internal nodes weigh C_T = 3, leaves C_I = 2 (Eval.cu:7-8,24-27).
//   BVH8: inner children weigh C_I = 2, leaf children C_T = 3 (Eval.cu:70-73);
//         child boxes are dequantized as p + 2^(e-127) * q (Eval.cu:55-63).
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

constexpr int kSahThreads = 256;

__device__ __forceinline__ void block_add_double(double v, double* acc) {
    __shared__ double part[kSahThreads / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kSahThreads / 64; ++w) s += part[w];
        atomicAdd(acc, s);
    }
}

__global__ __launch_bounds__(kSahThreads) void k_sah_bvh2(const nxb_bvh2_node* __restrict__ nodes, uint32_t count,
                                                          const nxb_aabb* __restrict__ scene, double* acc) {
    Box sb = {scene->bmin.x, scene->bmin.y, scene->bmin.z, scene->bmax.x, scene->bmax.y, scene->bmax.z};
    const float sa = box_area(sb);
    double sum = 0.0;
    const uint32_t stride = gridDim.x * kSahThreads;
    for (uint32_t i = blockIdx.x * kSahThreads + threadIdx.x; i < count; i += stride) {
        Box b;
        uint32_t l, r;
        load_node(nodes, i, b, l, r);
        float m = box_area(b) / sa;
        float t = (l != kInvalid) ? 3.0f * m : 2.0f * m;
        sum += (double)t;
    }
    block_add_double(sum, acc);
}

__global__ __launch_bounds__(kSahThreads) void k_sah_bvh8(const nxb_bvh8_node* __restrict__ nodes, uint32_t count,
                                                          const nxb_aabb* __restrict__ scene, double* acc) {
    Box sb = {scene->bmin.x, scene->bmin.y, scene->bmin.z, scene->bmax.x, scene->bmax.y, scene->bmax.z};
    const float sa = box_area(sb);
    double sum = 0.0;
    const uint32_t stride = gridDim.x * kSahThreads;
    for (uint32_t k = blockIdx.x * kSahThreads + threadIdx.x; k < count; k += stride) {
        const uint4* p = reinterpret_cast<const uint4*>(nodes + k);
        uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
        float px = __uint_as_float(w0.x), py = __uint_as_float(w0.y), pz = __uint_as_float(w0.z);
        float sx = __uint_as_float((w0.w & 0xFFu) << 23);
        float sy = __uint_as_float(((w0.w >> 8) & 0xFFu) << 23);
        float sz = __uint_as_float(((w0.w >> 16) & 0xFFu) << 23);
        uint32_t meta[2] = {w1.z, w1.w};
        uint32_t qlx[2] = {w2.x, w2.y}, qly[2] = {w2.z, w2.w}, qlz[2] = {w3.x, w3.y};
        uint32_t qhx[2] = {w3.z, w3.w}, qhy[2] = {w4.x, w4.y}, qhz[2] = {w4.z, w4.w};
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i) {
            const uint32_t w = i >> 2, sh = 8 * (i & 3);
            uint32_t m8 = (meta[w] >> sh) & 0xFFu;
            if (!m8) continue;
            bool internal = (m8 & 31u) >= 24u;
            Box c;
            c.lx = px + sx * (float)((qlx[w] >> sh) & 0xFFu);
            c.ly = py + sy * (float)((qly[w] >> sh) & 0xFFu);
            c.lz = pz + sz * (float)((qlz[w] >> sh) & 0xFFu);
            c.hx = px + sx * (float)((qhx[w] >> sh) & 0xFFu);
            c.hy = py + sy * (float)((qhy[w] >> sh) & 0xFFu);
            c.hz = pz + sz * (float)((qhz[w] >> sh) & 0xFFu);
            float m = box_area(c) / sa;
            sum += (double)(internal ? 2.0f * m : 3.0f * m);
        }
    }
    block_add_double(sum, acc);
}

// The BVH2 SAH of a BVH8 build from its wide records (record i: both
// children of internal node n + i, 64 bytes; n >= 2): each record adds its
// two children's terms (internal iff id >= n), and the root (n - 2's
// node, the union of its children) adds its own. Every node but the root is
// one record's child, so the terms are exactly k_sah_bvh2's.
__global__ __launch_bounds__(kSahThreads) void k_sah_rec8(const uint4* __restrict__ rec, uint32_t n,
                                                          const nxb_aabb* __restrict__ scene, double* acc) {
    Box sb = {scene->bmin.x, scene->bmin.y, scene->bmin.z, scene->bmax.x, scene->bmax.y, scene->bmax.z};
    const float sa = box_area(sb);
    double sum = 0.0;
    const uint32_t stride = gridDim.x * kSahThreads;
    for (uint32_t i = blockIdx.x * kSahThreads + threadIdx.x; i < n - 1; i += stride) {
        const uint4* p = rec + 4ull * i;
        const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
        Box l = {__uint_as_float(q0.x), __uint_as_float(q0.y), __uint_as_float(q0.z),
                 __uint_as_float(q0.w), __uint_as_float(q1.x), __uint_as_float(q1.y)};
        Box r = {__uint_as_float(q1.z), __uint_as_float(q1.w), __uint_as_float(q2.x),
                 __uint_as_float(q2.y), __uint_as_float(q2.z), __uint_as_float(q2.w)};
        const float ml = box_area(l) / sa, mr = box_area(r) / sa;
        sum += (double)(q3.x >= n ? 3.0f * ml : 2.0f * ml);
        sum += (double)(q3.y >= n ? 3.0f * mr : 2.0f * mr);
        if (i == n - 2) sum += (double)(3.0f * (box_area(box_union(l, r)) / sa));
    }
    block_add_double(sum, acc);
}

hipError_t launch_sah_rec8(const uint4* rec, uint32_t n, const nxb_aabb* scene, double* acc, hipStream_t s) {
    if (n < 2) return hipErrorInvalidValue;
    uint32_t grid = min((n - 1 + kSahThreads - 1) / kSahThreads, 2048u);
    hipLaunchKernelGGL(k_sah_rec8, dim3(grid), dim3(kSahThreads), 0, s, rec, n, scene, acc);
    return hipGetLastError();
}

hipError_t launch_sah_bvh2(const nxb_bvh2_node* nodes, uint32_t count, const nxb_aabb* scene, double* acc,
                           hipStream_t s) {
    uint32_t grid = min((count + kSahThreads - 1) / kSahThreads, 2048u);
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_sah_bvh2, dim3(grid), dim3(kSahThreads), 0, s, nodes, count, scene, acc);
    return hipGetLastError();
}

hipError_t launch_sah_bvh8(const nxb_bvh8_node* nodes, uint32_t count, const nxb_aabb* scene, double* acc,
                           hipStream_t s) {
    uint32_t grid = min((count + kSahThreads - 1) / kSahThreads, 2048u);
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_sah_bvh8, dim3(grid), dim3(kSahThreads), 0, s, nodes, count, scene, acc);
    return hipGetLastError();
}

}  // namespace nxb
