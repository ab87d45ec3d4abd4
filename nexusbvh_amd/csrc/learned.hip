// learned.hip - learned per-segment H-PLOC search radius.
//
// The sorted primitive array is cut into segments of 2^segment_log2
// positions. H-PLOC takes the radius of a merge task from the segment of the
// task's first sorted position (hploc.hip radius_for). This file provides
//  * segment features (16 floats per segment, from the sorted keys, the
//    permutation and the leaf boxes);
//  * per-segment SAH attribution of a built BVH2 (training labels): internal
//    node id = n - 1 + rep, where rep is a sorted position inside the node's
//    range, so the node's term goes to segment rep >> segment_log2;
//  * a 16 -> 64 -> 64 -> 4 MLP (ReLU) on MFMA: BF16 operands, FP32
//    accumulation, FP32 biases. Inference writes a radius table (classes
//    2, 4, 8, 16) with a confidence fallback to radius 8; training computes
//    the mean cross-entropy and its gradient in one fused forward/backward
//    kernel, followed by an Adam step.
//
// MFMA operand layout (v_mfma_f32_16x16x16_bf16, wave64), lane l, g = l / 16:
//   A (16 x 16, rows i, k): lane holds A[l % 16][4g + j], j = 0..3
//   B (16 x 16, k, cols n): lane holds B[4g + j][l % 16]
//   D (16 x 16):            lane holds D[4g + j][l % 16]
// Parameters are FP32, row-major [in][out]:
//   W1 16x64 | b1 64 | W2 64x64 | b2 64 | W3 64x4 | b3 4  (5508 floats).
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFeat = NXB_MLP_FEATURES;    // 16
constexpr int kHid = NXB_MLP_HIDDEN;       // 64
constexpr int kCls = NXB_MLP_CLASSES;      // 4
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kFeat * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kCls, kParams = kOffB3 + kCls;
static_assert(kParams == NXB_MLP_PARAMS, "parameter layout");
constexpr int kMlpThreads = 256;
constexpr int kMlpWaves = kMlpThreads / 64;

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Operand fetches from row-major BF16 LDS matrices (ld = row length).
// A[i][k] = M[r0 + i][k0 + k]
__device__ __forceinline__ bf16x4 lds_a(const uint16_t* M, int ld, int r0, int k0) {
    const int l = __lane_id(), g = l >> 4;
    const uint16_t* p = M + (r0 + (l & 15)) * ld + k0 + 4 * g;
    return bf16x4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}
// A[i][k] = M[r0 + k][c0 + i]   (A = M^T)
__device__ __forceinline__ bf16x4 lds_at(const uint16_t* M, int ld, int r0, int c0) {
    const int l = __lane_id(), g = l >> 4;
    const uint16_t* p = M + (r0 + 4 * g) * ld + c0 + (l & 15);
    return bf16x4{(short)p[0], (short)p[ld], (short)p[2 * ld], (short)p[3 * ld]};
}
// B[k][n] = M[k0 + k][n0 + n]
__device__ __forceinline__ bf16x4 lds_b(const uint16_t* M, int ld, int k0, int n0) {
    const int l = __lane_id(), g = l >> 4;
    const uint16_t* p = M + (k0 + 4 * g) * ld + n0 + (l & 15);
    return bf16x4{(short)p[0], (short)p[ld], (short)p[2 * ld], (short)p[3 * ld]};
}
// B[k][n] = M[n0 + n][k0 + k]   (B = M^T)
__device__ __forceinline__ bf16x4 lds_bt(const uint16_t* M, int ld, int k0, int n0) {
    const int l = __lane_id(), g = l >> 4;
    const uint16_t* p = M + (n0 + (l & 15)) * ld + k0 + 4 * g;
    return bf16x4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}
// D -> M[r0 + 4g + j][c0 + l % 16] as BF16.
__device__ __forceinline__ void lds_store_d(uint16_t* M, int ld, int r0, int c0, f32x4 d) {
    const int l = __lane_id(), g = l >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) M[(r0 + 4 * g + j) * ld + c0 + (l & 15)] = to_bf16(d[j]);
}
// Column sums of a D tile: lanes 0..15 end with the sum of column l.
__device__ __forceinline__ float d_colsum(f32x4 d) {
    float s = d[0] + d[1] + d[2] + d[3];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    return s;
}

// Weights as BF16 in LDS plus FP32 biases; W3 padded to 16 classes.
struct MlpWeights {
    uint16_t w1[kFeat * kHid];
    uint16_t w2[kHid * kHid];
    uint16_t w3[kHid * 16];
    float b1[kHid], b2[kHid], b3[16];
};
// Per-wave activations of a tile of 16 samples.
struct MlpTile {
    uint16_t x[16 * kFeat];
    uint16_t h1[16 * kHid];
    uint16_t h2[16 * kHid];
    uint16_t dz3[16 * 16];
    uint16_t dz2[16 * kHid];
    uint16_t dz1[16 * kHid];
};

__device__ void load_weights(MlpWeights& W, const float* __restrict__ p) {
    for (int i = threadIdx.x; i < kFeat * kHid; i += kMlpThreads) W.w1[i] = to_bf16(p[kOffW1 + i]);
    for (int i = threadIdx.x; i < kHid * kHid; i += kMlpThreads) W.w2[i] = to_bf16(p[kOffW2 + i]);
    for (int i = threadIdx.x; i < kHid * 16; i += kMlpThreads) {
        int u = i >> 4, c = i & 15;
        W.w3[i] = c < kCls ? to_bf16(p[kOffW3 + u * kCls + c]) : (uint16_t)0;
    }
    for (int i = threadIdx.x; i < kHid; i += kMlpThreads) {
        W.b1[i] = p[kOffB1 + i];
        W.b2[i] = p[kOffB2 + i];
    }
    if (threadIdx.x < 16) W.b3[threadIdx.x] = threadIdx.x < kCls ? p[kOffB3 + threadIdx.x] : 0.0f;
    __syncthreads();
}

// Forward pass of the 16 samples [s0, s0+16) of one wave; rows past count
// are zero features. Returns the logits tile (D layout, columns 0..3 valid).
__device__ __forceinline__ f32x4 mlp_forward_tile(const MlpWeights& W, MlpTile& T, const float* __restrict__ feat,
                                                  uint32_t count, uint32_t s0) {
    const int l = __lane_id(), g = l >> 4, r = l & 15;
    // X tile, A operand directly: X[r][4g + j]
    bf16x4 ax;
    {
        const uint32_t s = s0 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v = s < count ? feat[(size_t)s * kFeat + 4 * g + j] : 0.0f;
            uint16_t h = to_bf16(v);
            ax[j] = (short)h;
            T.x[r * kFeat + 4 * g + j] = h;
        }
    }
#pragma unroll
    for (int t = 0; t < kHid / 16; ++t) {
        f32x4 acc = {0, 0, 0, 0};
        acc = mfma(ax, lds_b(W.w1, kHid, 0, 16 * t), acc);
        const float b = W.b1[16 * t + r];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = fmaxf(acc[j] + b, 0.0f);
        lds_store_d(T.h1, kHid, 0, 16 * t, acc);
    }
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < kHid / 16; ++t) {
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < kHid / 16; ++ks) acc = mfma(lds_a(T.h1, kHid, 0, 16 * ks), lds_b(W.w2, kHid, 16 * ks, 16 * t), acc);
        const float b = W.b2[16 * t + r];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = fmaxf(acc[j] + b, 0.0f);
        lds_store_d(T.h2, kHid, 0, 16 * t, acc);
    }
    wave_lds_sync();
    f32x4 z = {0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < kHid / 16; ++ks) z = mfma(lds_a(T.h2, kHid, 0, 16 * ks), lds_b(W.w3, 16, 16 * ks, 0), z);
    const float b3 = W.b3[r];
#pragma unroll
    for (int j = 0; j < 4; ++j) z[j] += b3;
    return z;
}

// Softmax over the 4 logits of sample 4g + j held by lanes 16g + 0..3
// (lanes 16g + 4..15 hold padding). Returns max and the sum of exp(z - max)
// (valid in lanes whose column is < 4).
__device__ __forceinline__ void softmax4(float z, float& m, float& s) {
    const int c = __lane_id() & 15;
    float v = c < kCls ? z : -INFINITY;
    m = fmaxf(v, __shfl_xor(v, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    float e = c < kCls ? expf(z - m) : 0.0f;
    s = e + __shfl_xor(e, 1);
    s = s + __shfl_xor(s, 2);
}

// ---- inference: radius table ------------------------------------------------

__global__ __launch_bounds__(kMlpThreads) void k_mlp_table(const float* __restrict__ params,
                                                           const float* __restrict__ feat, uint32_t count,
                                                           float min_confidence, uint8_t* __restrict__ table,
                                                           float* __restrict__ logits_out, uint32_t* fallbacks) {
    __shared__ MlpWeights W;
    __shared__ MlpTile tiles[kMlpWaves];
    load_weights(W, params);
    const int w = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, c = l & 15;
    MlpTile& T = tiles[w];
    uint32_t fb = 0;
    for (uint32_t s0 = (blockIdx.x * kMlpWaves + w) * 16u; s0 < count; s0 += gridDim.x * kMlpWaves * 16u) {
        f32x4 z = mlp_forward_tile(W, T, feat, count, s0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t s = s0 + 4 * g + j;
            float m, sum;
            softmax4(z[j], m, sum);
            // first class with the maximal logit
            uint32_t is_max = (c < kCls && z[j] == m) ? (1u << c) : 0u;
            is_max |= __shfl_xor(is_max, 1);
            is_max |= __shfl_xor(is_max, 2);
            if (c == 0 && s < count) {
                uint32_t cls = (uint32_t)__builtin_ctz(is_max);
                float conf = 1.0f / sum;  // probability of the arg-max class
                uint8_t rad = (uint8_t)(2u << cls);
                if (!(conf >= min_confidence)) {
                    rad = 8;
                    ++fb;
                }
                table[s] = rad;
            }
            if (logits_out && c < kCls && s < count) logits_out[(size_t)s * kCls + c] = z[j];
        }
    }
    if (fallbacks) {
        fb += __shfl_xor(fb, 16);
        fb += __shfl_xor(fb, 32);
        if (l == 0 && fb) atomic_add_agent(fallbacks, fb);
    }
}

// ---- training: fused forward + backward -------------------------------------

// Adds the mean cross-entropy over `count` samples to *loss and its gradient
// to grads (FP32, same layout as params). grads and loss must be zeroed by
// the caller; they are accumulated with atomics.
__global__ __launch_bounds__(kMlpThreads) void k_mlp_grad(const float* __restrict__ params,
                                                          const float* __restrict__ feat,
                                                          const uint8_t* __restrict__ labels, uint32_t count,
                                                          float* grads, float* loss) {
    __shared__ MlpWeights W;
    __shared__ MlpTile tiles[kMlpWaves];
    __shared__ float gsum[kParams];
    load_weights(W, params);
    for (int i = threadIdx.x; i < kParams; i += kMlpThreads) gsum[i] = 0.0f;
    __syncthreads();
    const int w = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, c = l & 15;
    MlpTile& T = tiles[w];
    const float inv = 1.0f / (float)count;

    f32x4 gw1[kHid / 16], gw2[kHid / 16][kHid / 16], gw3[kHid / 16];
    float gb1[kHid / 16], gb2[kHid / 16], gb3 = 0.0f, lsum = 0.0f;
#pragma unroll
    for (int a = 0; a < kHid / 16; ++a) {
        gw1[a] = f32x4{0, 0, 0, 0};
        gw3[a] = f32x4{0, 0, 0, 0};
        gb1[a] = 0.0f;
        gb2[a] = 0.0f;
#pragma unroll
        for (int b = 0; b < kHid / 16; ++b) gw2[a][b] = f32x4{0, 0, 0, 0};
    }

    for (uint32_t s0 = (blockIdx.x * kMlpWaves + w) * 16u; s0 < count; s0 += gridDim.x * kMlpWaves * 16u) {
        f32x4 z = mlp_forward_tile(W, T, feat, count, s0);
        // dZ3 = (softmax - onehot) / count, zero for padding rows/columns
        f32x4 d3;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t s = s0 + 4 * g + j;
            float m, sum;
            softmax4(z[j], m, sum);
            const bool valid = s < count && c < kCls;
            const uint32_t y = s < count ? labels[s] : 0u;
            float p = valid ? expf(z[j] - m) / sum : 0.0f;
            if (valid && (uint32_t)c == y) lsum += (m - z[j]) + logf(sum);
            d3[j] = valid ? (p - ((uint32_t)c == y ? 1.0f : 0.0f)) * inv : 0.0f;
        }
        lds_store_d(T.dz3, 16, 0, 0, d3);
        gb3 += d_colsum(d3);
        wave_lds_sync();
        // dW3 += H2^T dZ3
#pragma unroll
        for (int mt = 0; mt < kHid / 16; ++mt)
            gw3[mt] = mfma(lds_at(T.h2, kHid, 0, 16 * mt), lds_b(T.dz3, 16, 0, 0), gw3[mt]);
        // dZ2 = (dZ3 W3^T) * [H2 > 0]
#pragma unroll
        for (int t = 0; t < kHid / 16; ++t) {
            f32x4 acc = {0, 0, 0, 0};
            acc = mfma(lds_a(T.dz3, 16, 0, 0), lds_bt(W.w3, 16, 0, 16 * t), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (T.h2[(4 * g + j) * kHid + 16 * t + c] == 0) acc[j] = 0.0f;
            gb2[t] += d_colsum(acc);
            lds_store_d(T.dz2, kHid, 0, 16 * t, acc);
        }
        wave_lds_sync();
        // dW2 += H1^T dZ2
#pragma unroll
        for (int mt = 0; mt < kHid / 16; ++mt)
#pragma unroll
            for (int nt = 0; nt < kHid / 16; ++nt)
                gw2[mt][nt] = mfma(lds_at(T.h1, kHid, 0, 16 * mt), lds_b(T.dz2, kHid, 0, 16 * nt), gw2[mt][nt]);
        // dZ1 = (dZ2 W2^T) * [H1 > 0]
#pragma unroll
        for (int t = 0; t < kHid / 16; ++t) {
            f32x4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < kHid / 16; ++ks)
                acc = mfma(lds_a(T.dz2, kHid, 0, 16 * ks), lds_bt(W.w2, kHid, 16 * ks, 16 * t), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (T.h1[(4 * g + j) * kHid + 16 * t + c] == 0) acc[j] = 0.0f;
            gb1[t] += d_colsum(acc);
            lds_store_d(T.dz1, kHid, 0, 16 * t, acc);
        }
        wave_lds_sync();
        // dW1 += X^T dZ1
#pragma unroll
        for (int nt = 0; nt < kHid / 16; ++nt)
            gw1[nt] = mfma(lds_at(T.x, kFeat, 0, 0), lds_b(T.dz1, kHid, 0, 16 * nt), gw1[nt]);
        wave_lds_sync();
    }

    // Block reduction in LDS, then one global atomic per parameter.
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = 4 * g + j;
#pragma unroll
        for (int a = 0; a < kHid / 16; ++a) {
            atomicAdd(&gsum[kOffW1 + row * kHid + 16 * a + c], gw1[a][j]);
            if (c < kCls) atomicAdd(&gsum[kOffW3 + (16 * a + row) * kCls + c], gw3[a][j]);
#pragma unroll
            for (int b = 0; b < kHid / 16; ++b) atomicAdd(&gsum[kOffW2 + (16 * a + row) * kHid + 16 * b + c], gw2[a][b][j]);
        }
    }
    if (l < 16) {
#pragma unroll
        for (int a = 0; a < kHid / 16; ++a) {
            atomicAdd(&gsum[kOffB1 + 16 * a + l], gb1[a]);
            atomicAdd(&gsum[kOffB2 + 16 * a + l], gb2[a]);
        }
        if (l < kCls) atomicAdd(&gsum[kOffB3 + l], gb3);
    }
    lsum += __shfl_xor(lsum, 1);
    lsum += __shfl_xor(lsum, 2);
    lsum += __shfl_xor(lsum, 4);
    lsum += __shfl_xor(lsum, 8);
    lsum += __shfl_xor(lsum, 16);
    lsum += __shfl_xor(lsum, 32);
    __syncthreads();
    for (int i = threadIdx.x; i < kParams; i += kMlpThreads) {
        float v = gsum[i];
        if (v != 0.0f) atomicAdd(&grads[i], v);
    }
    if (l == 0 && lsum != 0.0f) atomicAdd(loss, lsum * inv);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_adam(float* params, const float* __restrict__ grads, float* m, float* v, uint32_t step, float lr,
                       float beta1, float beta2, float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kParams) return;
    const float gr = grads[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gr;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(beta1, (float)step));
    const float vh = vi / (1.0f - powf(beta2, (float)step));
    params[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---- segment features ---------------------------------------------------------

constexpr int kFeatThreads = 256;

__device__ __forceinline__ float median3(float a, float b, float c) {
    return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
}

// Octree depth (top Morton bit = depth 0) of the first bit where two keys
// differ; kEqualDepth for equal keys.
constexpr int kEqualDepth = 64;
template <typename K>
__device__ __forceinline__ int split_depth(K a, K b) {
    K x = a ^ b;
    if (!x) return kEqualDepth;
    const int top = sizeof(K) == 8 ? 62 : 29;  // highest Morton bit
    int hb = sizeof(K) == 8 ? 63 - __builtin_clzll((uint64_t)x) : 31 - __builtin_clz((uint32_t)x);
    return max(top - hb, 0) / 3;
}
// Histogram bin of a split depth relative to the segment's own common prefix
// depth d0: 0, 1, 2, 3, 4-5, 6-8, 9+, equal keys.
__device__ __forceinline__ uint32_t depth_bin(int d, int d0) {
    if (d == kEqualDepth) return 7;
    int r = max(d - d0, 0);
    return r < 4 ? (uint32_t)r : (r < 6 ? 4u : (r < 9 ? 5u : 6u));
}

constexpr uint32_t kFeatBatch = 8;

// One block per segment. Features depend only on the segment's own geometry
// (not on the scene size), all roughly in [-1, 1]:
//   0      log2(primitive count) / 16
//   1, 2   log2(middle / largest extent), log2(smallest / largest extent) of
//          the segment box, / 24
//   3      log2(segment half-area / sum of leaf half-areas) / 16
//   4      mean of log2(leaf half-area / segment half-area) / 48
//   5      standard deviation of log2(leaf half-area) / 16
//   6      fraction of sliver leaves (longest extent > 10 x middle extent)
//   7      mean of log2(longest / middle leaf extent) / 24
//   8..15  split depth between consecutive sorted keys relative to the depth
//          of the segment's common key prefix: 0, 1, 2, 3, 4-5, 6-8, 9+,
//          equal keys (fractions of the key pairs)
template <typename K>
__global__ __launch_bounds__(kFeatThreads) void k_segment_features(const K* __restrict__ keys,
                                                                   const uint32_t* __restrict__ perm,
                                                                   const nxb_bvh2_node* __restrict__ leaves,
                                                                   uint32_t n, uint32_t seg_log2,
                                                                   float* __restrict__ feat) {
    constexpr int NW = kFeatThreads / 64;
    __shared__ float red[10][NW];
    __shared__ uint32_t hist[8];
    const uint32_t seg = blockIdx.x, tid = threadIdx.x;
    const uint32_t lo = seg << seg_log2, hi = min(n, lo + (1u << seg_log2));
    if (tid < 8) hist[tid] = 0;
    const int d0 = split_depth<K>(keys[lo], keys[hi - 1]);
    // leaf log-areas are summed relative to the segment's first leaf, so the
    // variance of near-equal areas does not cancel in FP32
    float la_ref;
    {
        const uint32_t v0 = perm[lo];
        Box b = v0 < n ? load_box(leaves, v0) : Box{0, 0, 0, 0, 0, 0};
        float ex = b.hx - b.lx, ey = b.hy - b.ly, ez = b.hz - b.lz;
        la_ref = log2f(fmaxf(ex * ey + ey * ez + ez * ex, 1e-30f));
    }
    __syncthreads();
    Box sb = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    float sum_a = 0.0f, la = 0.0f, la2 = 0.0f, sliver = 0.0f, aniso = 0.0f;
    // kFeatBatch keys per thread at a time: their permutation entries, then
    // their leaf boxes (a gather by primitive index) are all requested before
    // the first is used, so a thread waits for one round trip per batch
    // instead of two per key
    for (uint32_t i0 = lo + tid; i0 < hi; i0 += kFeatBatch * kFeatThreads) {
        uint32_t v[kFeatBatch];
#pragma unroll
        for (uint32_t k = 0; k < kFeatBatch; ++k) {
            const uint32_t i = i0 + k * kFeatThreads;
            v[k] = i < hi ? perm[i] : kInvalid;
        }
        Box bb[kFeatBatch];
#pragma unroll
        for (uint32_t k = 0; k < kFeatBatch; ++k) bb[k] = v[k] < n ? load_box(leaves, v[k]) : Box{0, 0, 0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < kFeatBatch; ++k) {
            const uint32_t i = i0 + k * kFeatThreads;
            if (i >= hi) break;
            const Box& b = bb[k];
            sb = box_union(sb, b);
            float ex = b.hx - b.lx, ey = b.hy - b.ly, ez = b.hz - b.lz;
            float a = ex * ey + ey * ez + ez * ex;
            sum_a += a;
            float l2 = log2f(fmaxf(a, 1e-30f)) - la_ref;
            la += l2;
            la2 += l2 * l2;
            float e0 = fmaxf(ex, fmaxf(ey, ez)), e1 = median3(ex, ey, ez);
            sliver += e0 > 10.0f * e1 ? 1.0f : 0.0f;
            aniso += log2f(fmaxf(e0, 1e-30f) / fmaxf(e1, 1e-30f));
            if (i + 1 < hi) atomicAdd(&hist[depth_bin(split_depth<K>(keys[i], keys[i + 1]), d0)], 1u);
        }
    }
    float v[10] = {sb.lx, sb.ly, sb.lz, -sb.hx, -sb.hy, -sb.hz, sum_a, la, la2, sliver};
#pragma unroll
    for (int q = 0; q < 6; ++q)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[q] = fminf(v[q], __shfl_xor(v[q], off));
#pragma unroll
    for (int q = 6; q < 10; ++q)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[q] += __shfl_xor(v[q], off);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) aniso += __shfl_xor(aniso, off);
    __shared__ float red_aniso[NW];
    if ((tid & 63) == 0) {
#pragma unroll
        for (int q = 0; q < 10; ++q) red[q][tid >> 6] = v[q];
        red_aniso[tid >> 6] = aniso;
    }
    __syncthreads();
    if (tid == 0) {
        float r[10];
        for (int q = 0; q < 10; ++q) {
            float x = red[q][0];
            for (int k = 1; k < NW; ++k) x = q < 6 ? fminf(x, red[q][k]) : x + red[q][k];
            r[q] = x;
        }
        float an = 0.0f;
        for (int k = 0; k < NW; ++k) an += red_aniso[k];
        const float cnt = (float)(hi - lo);
        const float ex = -r[3] - r[0], ey = -r[4] - r[1], ez = -r[5] - r[2];
        const float e0 = fmaxf(ex, fmaxf(ey, ez)), e2 = fminf(ex, fminf(ey, ez)), e1 = median3(ex, ey, ez);
        const float seg_a = fmaxf(ex * ey + ey * ez + ez * ex, 1e-30f);
        const float mean_d = r[7] / cnt;
        const float var_la = fmaxf(r[8] / cnt - mean_d * mean_d, 0.0f);
        const float mean_la = la_ref + mean_d;
        float* f = feat + (size_t)seg * kFeat;
        f[0] = log2f(cnt) / 16.0f;
        f[1] = log2f(fmaxf(e1, 1e-30f) / fmaxf(e0, 1e-30f)) / 24.0f;
        f[2] = log2f(fmaxf(e2, 1e-30f) / fmaxf(e0, 1e-30f)) / 24.0f;
        f[3] = log2f(seg_a / fmaxf(r[6], 1e-30f)) / 16.0f;
        f[4] = (mean_la - log2f(seg_a)) / 48.0f;
        f[5] = sqrtf(var_la) / 16.0f;
        f[6] = r[9] / cnt;
        f[7] = an / cnt / 24.0f;
        const float pairs = cnt > 1.0f ? cnt - 1.0f : 1.0f;
        for (int q = 0; q < 8; ++q) f[8 + q] = (float)hist[q] / pairs;
    }
}

// ---- per-segment SAH of a BVH2 (training labels) -----------------------------

constexpr int kSegSahThreads = 256;

// Internal node id (n..2n-2) contributes its SAH term 3 * area / scene area
// (Eval.cu:24-27) to segment (id - n + 1) >> seg_log2. A block covers 256
// consecutive ids, i.e. at most two segments when seg_log2 >= 8.
__global__ __launch_bounds__(kSegSahThreads) void k_segment_sah(const nxb_bvh2_node* __restrict__ nodes, uint32_t n,
                                                                uint32_t seg_log2,
                                                                const nxb_aabb* __restrict__ scene,
                                                                double* seg_cost) {
    __shared__ double part[2][kSegSahThreads / 64];
    Box sbx = {scene->bmin.x, scene->bmin.y, scene->bmin.z, scene->bmax.x, scene->bmax.y, scene->bmax.z};
    const float sa = box_area(sbx);
    const uint32_t first = n + blockIdx.x * kSegSahThreads;
    const uint32_t id = first + threadIdx.x;
    const uint32_t seg0 = (first - n + 1) >> seg_log2;
    double t0 = 0.0, t1 = 0.0;
    if (id < 2 * n - 1) {
        Box b = load_box(nodes, id);
        double t = (double)(3.0f * (box_area(b) / sa));
        if (((id - n + 1) >> seg_log2) == seg0) t0 = t;
        else t1 = t;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        t0 += __shfl_xor(t0, off);
        t1 += __shfl_xor(t1, off);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = t0;
        part[1][threadIdx.x >> 6] = t1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < kSegSahThreads / 64; ++k) {
            a += part[0][k];
            b += part[1][k];
        }
        if (a != 0.0) atomicAdd(&seg_cost[seg0], a);
        if (b != 0.0) atomicAdd(&seg_cost[seg0 + 1], b);
    }
}

// ---- launchers -----------------------------------------------------------------

hipError_t launch_segment_features(const void* keys, uint32_t key_bits, const uint32_t* perm,
                                   const nxb_bvh2_node* leaves, uint32_t n, uint32_t seg_log2, float* feat,
                                   hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t segs = ((n - 1) >> seg_log2) + 1;
    if (key_bits == 32)
        hipLaunchKernelGGL(k_segment_features<uint32_t>, dim3(segs), dim3(kFeatThreads), 0, s,
                           static_cast<const uint32_t*>(keys), perm, leaves, n, seg_log2, feat);
    else
        hipLaunchKernelGGL(k_segment_features<uint64_t>, dim3(segs), dim3(kFeatThreads), 0, s,
                           static_cast<const uint64_t*>(keys), perm, leaves, n, seg_log2, feat);
    return hipGetLastError();
}

hipError_t launch_segment_sah(const nxb_bvh2_node* nodes, uint32_t n, uint32_t seg_log2, const nxb_aabb* scene,
                              double* seg_cost, hipStream_t s) {
    if (n < 2) return hipSuccess;
    if (seg_log2 < 8) return hipErrorInvalidValue;
    const uint32_t blocks = (n - 1 + kSegSahThreads - 1) / kSegSahThreads;
    hipLaunchKernelGGL(k_segment_sah, dim3(blocks), dim3(kSegSahThreads), 0, s, nodes, n, seg_log2, scene, seg_cost);
    return hipGetLastError();
}

static uint32_t mlp_grid(uint32_t count) {
    uint32_t tiles = (count + 15) / 16;
    uint32_t blocks = (tiles + kMlpWaves - 1) / kMlpWaves;
    return blocks == 0 ? 1u : min(blocks, 256u);
}

hipError_t launch_mlp_table(const float* params, const float* feat, uint32_t count, float min_confidence,
                            uint8_t* table, float* logits, uint32_t* fallbacks, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_table, dim3(mlp_grid(count)), dim3(kMlpThreads), 0, s, params, feat, count,
                       min_confidence, table, logits, fallbacks);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* params, const float* feat, const uint8_t* labels, uint32_t count,
                           float* grads, float* loss, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_grad, dim3(mlp_grid(count)), dim3(kMlpThreads), 0, s, params, feat, labels, count,
                       grads, loss);
    return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grads, float* m, float* v, uint32_t step, float lr, float beta1,
                       float beta2, float eps, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, params, grads, m, v, step, lr, beta1,
                       beta2, eps);
    return hipGetLastError();
}

}  // namespace nxb
