// trace.hip - closest-hit ray traversal of a BVH8 (the compressed wide node
// of BVH.h:39-96, Ylitie, Karras and Laine, "Efficient Incoherent Ray
// Traversal on GPUs Through Compressed Wide BVHs", HPG 2017). The reference
// builds BVH8s but has no traversal; this is the consumer that checks the
// 80-byte format end to end and measures traced cost.
//
// One lane per ray. A node's eight child boxes are decoded from the 8-bit
// grid (lo = p + qlo * 2^(e-127), hi = p + qhi * 2^(e-127)) and slab-tested;
// hit leaves (one triangle each) are intersected at once, and the hit inner
// children are pushed as one group entry (their parent's child base, inner
// mask and hit mask), so the stack holds at most one entry per level. Within
// a group the children go in the order slot ^ octant: the builder's greedy
// slot assignment puts the child on the -x side into slots with bit 2 clear
// (likewise y, bit 1, and z, bit 0), so this order is front to back along
// the ray's octant.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

namespace {

constexpr int kTraceThreads = 128;
constexpr uint32_t kTraceStack = kBvh8LevelWords;  // deeper than any BVH8 the builder makes

struct Vec3 {
    float x, y, z;
};

__device__ __forceinline__ Vec3 sub3(Vec3 a, Vec3 b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ Vec3 cross3(Vec3 a, Vec3 b) {
    return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Moller-Trumbore ("Fast, Minimum Storage Ray/Triangle Intersection", 1997)
// in FP32, evaluated in this exact order (no contraction). Updates the
// closest hit for tmin <= t <= best (equal t: smaller index).
// Triangle prim's corners: nine floats per triangle, or (indexed mesh) three
// vertex indices per triangle into vertices of three floats. A corner index
// past vertex_count drops the triangle.
struct TriSource {
    const float* __restrict__ v;
    const uint32_t* __restrict__ idx;  // null: triangle soup
    uint32_t vertex_count;
};

template <bool INDEXED>
__device__ __forceinline__ bool fetch_triangle(const TriSource& ts, uint32_t prim, Vec3& v0, Vec3& v1, Vec3& v2) {
    if constexpr (INDEXED) {
        const uint32_t a = ts.idx[(size_t)prim * 3], b = ts.idx[(size_t)prim * 3 + 1],
                       c = ts.idx[(size_t)prim * 3 + 2];
        if (a >= ts.vertex_count || b >= ts.vertex_count || c >= ts.vertex_count) return false;
        v0 = Vec3{ts.v[(size_t)a * 3], ts.v[(size_t)a * 3 + 1], ts.v[(size_t)a * 3 + 2]};
        v1 = Vec3{ts.v[(size_t)b * 3], ts.v[(size_t)b * 3 + 1], ts.v[(size_t)b * 3 + 2]};
        v2 = Vec3{ts.v[(size_t)c * 3], ts.v[(size_t)c * 3 + 1], ts.v[(size_t)c * 3 + 2]};
    } else {
        const float* v = ts.v + (size_t)prim * 9;
        v0 = Vec3{v[0], v[1], v[2]};
        v1 = Vec3{v[3], v[4], v[5]};
        v2 = Vec3{v[6], v[7], v[8]};
    }
    return true;
}

template <bool INDEXED>
__device__ __forceinline__ void hit_triangle(const TriSource& ts, uint32_t prim, Vec3 o, Vec3 d, float tmin,
                                             float& best_t, uint32_t& best_prim, float& bu, float& bv) {
    Vec3 v0, v1, v2;
    if (!fetch_triangle<INDEXED>(ts, prim, v0, v1, v2)) return;
    const Vec3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    const Vec3 pv = cross3(d, e2);
    const float det = dot3(e1, pv);
    if (det == 0.0f) return;
    const float inv = 1.0f / det;
    const Vec3 tv = sub3(o, v0);
    const float u = dot3(tv, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return;
    const Vec3 qv = cross3(tv, e1);
    const float w = dot3(d, qv) * inv;
    if (!(w >= 0.0f && u + w <= 1.0f)) return;
    const float t = dot3(e2, qv) * inv;
    if (!(t >= tmin && t <= best_t)) return;
    if (t < best_t || prim < best_prim) {
        best_t = t;
        best_prim = prim;
        bu = u;
        bv = w;
    }
}

__device__ __forceinline__ float exp_scale(uint32_t e) { return __uint_as_float(e << 23); }  // 2^(e-127)

template <bool INDEXED>
__global__ __launch_bounds__(kTraceThreads) void k_trace(const nxb_bvh8_node* __restrict__ nodes, uint32_t node_count,
                                                         const uint32_t* __restrict__ prim_idx, uint32_t prim_count,
                                                         TriSource tris,
                                                         const nxb_ray* __restrict__ rays, uint32_t ray_count,
                                                         nxb_hit* __restrict__ hits, uint32_t* __restrict__ stats) {
    const uint32_t r = blockIdx.x * kTraceThreads + threadIdx.x;
    if (r >= ray_count) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + r);
    const float4 ra = rp[0], rb = rp[1];
    const Vec3 o{ra.x, ra.y, ra.z}, d{rb.x, rb.y, rb.z};
    const float tmin = ra.w;
    float best_t = rb.w, bu = 0.0f, bv = 0.0f;
    uint32_t best_prim = NXB_INVALID_IDX;
    // reciprocal direction; a zero component becomes a tiny one of the same sign
    auto safe_inv = [](float x) { return 1.0f / (fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x)); };
    const float ix = safe_inv(d.x), iy = safe_inv(d.y), iz = safe_inv(d.z);
    const uint32_t oct = (d.x < 0.0f ? 4u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 1u : 0u);

    uint2 stack[kTraceStack];
    uint32_t sp = 0, visited = 0, tested = 0;
    uint32_t node = 0;
    bool have = node_count > 0;
    for (;;) {
        if (have) {
            have = false;
            // a tree visits each node at most once: the bound only ends a
            // walk over a malformed node array
            if (node >= node_count || visited >= node_count) break;
            ++visited;
            const float4* np = reinterpret_cast<const float4*>(nodes + node);
            const float4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
            // p.xyz | e[3], imask | child_base | prim_base | meta[8] | q*[8] x 6
            const float px = w0.x, py = w0.y, pz = w0.z;
            const uint32_t em = __float_as_uint(w0.w);
            const float sx = exp_scale(em & 0xFFu), sy = exp_scale((em >> 8) & 0xFFu), sz = exp_scale((em >> 16) & 0xFFu);
            const uint32_t imask = em >> 24;
            const uint32_t child_base = __float_as_uint(w1.x), prim_base = __float_as_uint(w1.y);
            const uint32_t meta_lo = __float_as_uint(w1.z), meta_hi = __float_as_uint(w1.w);
            const uint32_t qlx[2] = {__float_as_uint(w2.x), __float_as_uint(w2.y)};
            const uint32_t qly[2] = {__float_as_uint(w2.z), __float_as_uint(w2.w)};
            const uint32_t qlz[2] = {__float_as_uint(w3.x), __float_as_uint(w3.y)};
            const uint32_t qhx[2] = {__float_as_uint(w3.z), __float_as_uint(w3.w)};
            const uint32_t qhy[2] = {__float_as_uint(w4.x), __float_as_uint(w4.y)};
            const uint32_t qhz[2] = {__float_as_uint(w4.z), __float_as_uint(w4.w)};
            uint32_t inner_hits = 0;
#pragma unroll
            for (uint32_t s = 0; s < 8; ++s) {
                const uint32_t h = s >> 2, sh = (s & 3) * 8;
                const uint32_t meta = ((s < 4 ? meta_lo : meta_hi) >> sh) & 0xFFu;
                if (meta == 0) continue;
                // Decoded child box, widened outward. The grid bounds the
                // child exactly, but two roundings move a decoded face
                // inward: the builder's difference fl(c - p), off by up to
                // |c - p| * 2^-24 <= (|c| + |p|) * 2^-24, which can land
                // exactly on a grid point, and the decode p + q * s, off by
                // |l| * 2^-24. (|l| + |p|) * 2^-21 covers both with margin.
                float lx = px + (float)((qlx[h] >> sh) & 0xFFu) * sx;
                float ly = py + (float)((qly[h] >> sh) & 0xFFu) * sy;
                float lz = pz + (float)((qlz[h] >> sh) & 0xFFu) * sz;
                float hx = px + (float)((qhx[h] >> sh) & 0xFFu) * sx;
                float hy = py + (float)((qhy[h] >> sh) & 0xFFu) * sy;
                float hz = pz + (float)((qhz[h] >> sh) & 0xFFu) * sz;
                const float apx = fabsf(px), apy = fabsf(py), apz = fabsf(pz);
                lx -= (fabsf(lx) + apx) * 0x1p-21f; ly -= (fabsf(ly) + apy) * 0x1p-21f; lz -= (fabsf(lz) + apz) * 0x1p-21f;
                hx += (fabsf(hx) + apx) * 0x1p-21f; hy += (fabsf(hy) + apy) * 0x1p-21f; hz += (fabsf(hz) + apz) * 0x1p-21f;
                const float tx0 = (lx - o.x) * ix, tx1 = (hx - o.x) * ix;
                const float ty0 = (ly - o.y) * iy, ty1 = (hy - o.y) * iy;
                const float tz0 = (lz - o.z) * iz, tz1 = (hz - o.z) * iz;
                const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
                // the far distance is widened by 2^-20 against the slab test's
                // own rounding (Ize, "Robust BVH Ray Traversal", 2013)
                const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * (1.0f + 0x1p-20f);
                if (!(tn <= tf && tn <= best_t)) continue;
                if (imask & (1u << s)) {
                    inner_hits |= 1u << s;
                } else {
                    const uint32_t k = prim_base + (meta & 0x1Fu);
                    if (k < prim_count) {
                        ++tested;
                        const uint32_t prim = prim_idx[k];
                        if (prim < prim_count) hit_triangle<INDEXED>(tris, prim, o, d, tmin, best_t, best_prim, bu, bv);
                    }
                }
            }
            if (inner_hits && sp < kTraceStack) stack[sp++] = make_uint2(child_base, imask | (inner_hits << 8));
        }
        if (sp == 0) break;
        uint2& top = stack[sp - 1];
        const uint32_t im = top.y & 0xFFu;
        uint32_t hm = top.y >> 8;
        uint32_t s = 0;
#pragma unroll
        for (int k = 7; k >= 0; --k)  // the first slot in the order slot ^ oct
            if (hm & (1u << ((uint32_t)k ^ oct))) s = (uint32_t)k ^ oct;
        hm &= ~(1u << s);
        node = top.x + (uint32_t)__popc(im & ((1u << s) - 1u));
        if (hm) top.y = im | (hm << 8);
        else --sp;
        have = true;
    }
    hits[r] = nxb_hit{best_prim == NXB_INVALID_IDX ? rb.w : best_t, best_prim, bu, bv};
    if (stats) {
        stats[2 * (size_t)r] = visited;
        stats[2 * (size_t)r + 1] = tested;
    }
}

}  // namespace

hipError_t launch_trace(const nxb_bvh8_node* nodes, uint32_t node_count, const uint32_t* prim_idx,
                        uint32_t prim_count, const float* tris, const nxb_ray* rays, uint32_t ray_count,
                        nxb_hit* hits, uint32_t* stats, hipStream_t s, const uint32_t* indices,
                        uint32_t vertex_count) {
    if (ray_count == 0) return hipSuccess;
    const uint32_t grid = (ray_count + kTraceThreads - 1) / kTraceThreads;
    const TriSource ts{tris, indices, vertex_count};
    hipLaunchKernelGGL(indices ? k_trace<true> : k_trace<false>, dim3(grid), dim3(kTraceThreads), 0, s, nodes,
                       node_count, prim_idx, prim_count, ts, rays, ray_count, hits, stats);
    return hipGetLastError();
}

}  // namespace nxb
