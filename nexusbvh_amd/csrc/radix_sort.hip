// radix_sort.hip - stable LSD key/value radix sort (onesweep, decoupled look-back).
//
// Replaces the CUB DeviceRadixSort::SortPairs call of RadixSort
// (NexusBVH/src/Cuda/Setup.cu:62-112). Semantics: stable sort of
// (key, value) pairs on key bits [begin_bit, end_bit); keys are carried whole.
//
// Design (gfx950, wave64):
//  * one histogram pass computes the digit counts of every pass at once;
//  * each onesweep pass: 256-thread blocks take tiles in order from an
//    atomic tile counter (so a tile only ever waits on tiles that are
//    already running), load keys warp-striped (row k of wave w holds 64
//    consecutive keys), rank stably with per-row 64-lane ballot matching and
//    a per-wave LDS histogram, publish per-digit tile counts, and resolve
//    their global offsets by decoupled look-back over epoch-tagged 64-bit
//    status words (no clearing between passes or builds);
//  * keys/values are scattered into LDS in digit order and written out
//    from LDS so that runs of one digit go to consecutive addresses.
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kMaxPasses = 10;

// Look-back status word: [63:34] epoch, [33:32] flag, [31:0] count.
constexpr uint64_t kFlagAggregate = 1;
constexpr uint64_t kFlagInclusive = 2;
__device__ __forceinline__ uint64_t pack_status(uint32_t epoch, uint64_t flag, uint32_t count) {
    return ((uint64_t)epoch << 34) | (flag << 32) | count;
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kSortWaves; ++k) {
        uint32_t t = tmp[k];
        if ((uint32_t)k < w) before += t;
        all += t;
    }
    __syncthreads();
    if (total) *total = all;
    return before + x - v;
}

__global__ void k_sort_init(uint32_t* hist, uint32_t hist_words, uint32_t* tile_counter, uint32_t passes) {
    for (uint32_t i = threadIdx.x; i < hist_words; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x < passes) tile_counter[threadIdx.x] = 0;
}

template <typename K, int BITS>
__global__ __launch_bounds__(kSortThreads) void k_sort_histogram(const K* __restrict__ keys, uint32_t n,
                                                                 uint32_t begin_bit, uint32_t end_bit,
                                                                 uint32_t passes, uint32_t* hist) {
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t sh[kMaxPasses * RADIX];
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += kSortThreads) sh[i] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * kSortThreads;
    for (uint32_t i = blockIdx.x * kSortThreads + threadIdx.x; i < n; i += stride) {
        K key = keys[i];
        for (uint32_t p = 0; p < passes; ++p) {
            uint32_t shift = begin_bit + p * BITS;
            uint32_t w = min((uint32_t)BITS, end_bit - shift);
            uint32_t d = digit_of(key, shift, (1u << w) - 1u);
            atomicAdd(&sh[p * RADIX + d], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += kSortThreads) {
        uint32_t c = sh[i];
        if (c) atomic_add_agent(&hist[i], c);
    }
}

template <typename K, int BITS, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift, uint32_t width,
    const uint32_t* __restrict__ hist, uint32_t* tile_counter, uint64_t* status, uint32_t epoch) {
    constexpr int RADIX = 1 << BITS;
    constexpr int TILE = kSortThreads * ITEMS;
    static_assert(RADIX <= kSortThreads, "one digit per thread");
    __shared__ K skeys[TILE];
    __shared__ uint32_t svals[TILE];
    __shared__ uint32_t wave_hist[kSortWaves][RADIX];
    __shared__ uint32_t local_start[RADIX];
    __shared__ uint32_t global_base[RADIX];
    __shared__ uint32_t scan_tmp[kSortWaves];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (uint32_t i = tid; i < (uint32_t)(kSortWaves * RADIX); i += kSortThreads) (&wave_hist[0][0])[i] = 0;
    if (tid == 0) s_tile = atomic_add_agent(tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * TILE;
    const uint32_t valid_count = min((uint32_t)TILE, n - base);
    const uint32_t mask = (1u << width) - 1u;

    K key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t dig[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        uint32_t idx = base + w * 64 * ITEMS + k * 64 + lane;
        if (idx < n) {
            key[k] = keys_in[idx];
            val[k] = vals_in ? vals_in[idx] : idx;
            dig[k] = digit_of(key[k], shift, mask);
        } else {
            key[k] = 0;
            val[k] = 0;
            dig[k] = RADIX - 1;  // ranks after every valid item of the tile
        }
    }

    // Stable rank within the wave: rows in order, lanes in order.
    uint32_t rank[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        uint32_t d = dig[k];
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t m = ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rr = mask_rank(peers);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t before = wave_hist[w][d];
        rank[k] = before + rr;
        wave_lds_sync();
        if (rr == 0) wave_hist[w][d] = before + cnt;
        wave_lds_sync();
    }
    __syncthreads();

    // Per digit: wave prefix (in place) and tile total.
    uint32_t cnt = 0;
    if (tid < (uint32_t)RADIX) {
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ++ww) {
            uint32_t c = wave_hist[ww][tid];
            wave_hist[ww][tid] = cnt;
            cnt += c;
        }
    }
    uint32_t lstart = block_exclusive_scan(tid < (uint32_t)RADIX ? cnt : 0u, scan_tmp, nullptr);
    uint32_t hbase = block_exclusive_scan(tid < (uint32_t)RADIX ? hist[tid] : 0u, scan_tmp, nullptr);

    if (tid < (uint32_t)RADIX) {
        uint32_t excl = 0;
        if (tile == 0) {
            store_agent64(&status[tid], pack_status(epoch, kFlagInclusive, cnt));
        } else {
            store_agent64(&status[(size_t)tile * RADIX + tid], pack_status(epoch, kFlagAggregate, cnt));
            uint32_t j = tile - 1;
            for (;;) {
                uint64_t s = load_agent64(&status[(size_t)j * RADIX + tid]);
                uint32_t e = (uint32_t)(s >> 34);
                uint32_t f = (uint32_t)(s >> 32) & 3u;
                if (e != epoch || f == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (uint32_t)s;
                if (f == kFlagInclusive) break;
                --j;
            }
            store_agent64(&status[(size_t)tile * RADIX + tid], pack_status(epoch, kFlagInclusive, excl + cnt));
        }
        local_start[tid] = lstart;
        global_base[tid] = hbase + excl;
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        uint32_t d = dig[k];
        uint32_t pos = local_start[d] + wave_hist[w][d] + rank[k];
        skeys[pos] = key[k];
        svals[pos] = val[k];
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        uint32_t i = k * kSortThreads + tid;
        if (i < valid_count) {
            K kk = skeys[i];
            uint32_t d = digit_of(kk, shift, mask);
            uint32_t o = global_base[d] + (i - local_start[d]);
            if (o < n) {  // always true when every tile is consistent
                keys_out[o] = kk;
                vals_out[o] = svals[i];
            }
        }
    }
}

template <typename K>
__global__ void k_copy_pairs(const K* __restrict__ ki, const uint32_t* __restrict__ vi, K* __restrict__ ko,
                             uint32_t* __restrict__ vo, uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        ko[i] = ki[i];
        vo[i] = vi ? vi[i] : i;
    }
}

namespace {

template <typename K>
constexpr int items_for() { return sizeof(K) == 4 ? 16 : 12; }

template <typename K, int BITS>
hipError_t run_sort(const K* keys_in, const uint32_t* values_in, K* keys_out, uint32_t* values_out, K* keys_tmp,
                    uint32_t* values_tmp, uint32_t n, uint32_t begin_bit, uint32_t end_bit, SortScratch sc,
                    hipStream_t s) {
    constexpr int RADIX = 1 << BITS;
    constexpr int ITEMS = items_for<K>();
    constexpr uint32_t TILE = kSortThreads * ITEMS;
    uint32_t passes = (end_bit - begin_bit + BITS - 1) / BITS;
    uint32_t grid_blocks = (n + 255) / 256;
    if (passes == 0 || n <= 1) {
        hipLaunchKernelGGL(k_copy_pairs<K>, dim3(grid_blocks ? min(grid_blocks, 4096u) : 1), dim3(256), 0, s,
                           keys_in, values_in, keys_out, values_out, n);
        return hipGetLastError();
    }
    uint32_t tiles = (n + TILE - 1) / TILE;
    if ((size_t)tiles * RADIX > sc.status_words || passes > (uint32_t)kMaxPasses) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(256), 0, s, sc.hist, passes * RADIX, sc.tile_counter, passes);
    uint32_t hgrid = min((n + kSortThreads * 16 - 1) / (kSortThreads * 16), 2048u);
    hipLaunchKernelGGL((k_sort_histogram<K, BITS>), dim3(hgrid), dim3(kSortThreads), 0, s, keys_in, n, begin_bit,
                       end_bit, passes, sc.hist);
    const K* src_k = keys_in;
    const uint32_t* src_v = values_in;
    for (uint32_t p = 0; p < passes; ++p) {
        bool to_out = ((passes - 1 - p) % 2) == 0;
        K* dk = to_out ? keys_out : keys_tmp;
        uint32_t* dv = to_out ? values_out : values_tmp;
        uint32_t shift = begin_bit + p * BITS;
        uint32_t width = min((uint32_t)BITS, end_bit - shift);
        uint32_t epoch = ++(*sc.epoch);
        if (epoch >= (1u << 30)) { *sc.epoch = 1; epoch = 1; }
        hipLaunchKernelGGL((k_onesweep<K, BITS, ITEMS>), dim3(tiles), dim3(kSortThreads), 0, s, src_k, src_v, dk,
                           dv, n, shift, width, sc.hist + p * RADIX, sc.tile_counter + p, sc.status, epoch);
        src_k = dk;
        src_v = dv;
    }
    return hipGetLastError();
}

template <typename K>
hipError_t dispatch_bits(const void* ki, const uint32_t* vi, void* ko, uint32_t* vo, void* kt, uint32_t* vt,
                         uint32_t n, uint32_t b, uint32_t e, uint32_t digit_bits, SortScratch sc, hipStream_t s) {
    const K* a = static_cast<const K*>(ki);
    K* o = static_cast<K*>(ko);
    K* t = static_cast<K*>(kt);
    switch (digit_bits) {
        case 6: return run_sort<K, 6>(a, vi, o, vo, t, vt, n, b, e, sc, s);
        case 7: return run_sort<K, 7>(a, vi, o, vo, t, vt, n, b, e, sc, s);
        case 8: return run_sort<K, 8>(a, vi, o, vo, t, vt, n, b, e, sc, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

size_t sort_scratch_bytes(uint32_t n, uint32_t key_bits, uint32_t digit_bits) {
    (void)digit_bits;
    uint32_t items = key_bits == 32 ? items_for<uint32_t>() : items_for<uint64_t>();
    size_t tiles = (n + kSortThreads * items - 1) / (kSortThreads * items);
    if (tiles == 0) tiles = 1;
    return tiles * 256 * sizeof(uint64_t);
}

hipError_t launch_radix_sort(const void* keys_in, const uint32_t* values_in, void* keys_out,
                             uint32_t* values_out, void* keys_tmp, uint32_t* values_tmp, uint32_t n,
                             uint32_t key_bits, uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits,
                             SortScratch sc, hipStream_t s) {
    if (end_bit < begin_bit || end_bit > key_bits) return hipErrorInvalidValue;
    if (key_bits == 32)
        return dispatch_bits<uint32_t>(keys_in, values_in, keys_out, values_out, keys_tmp, values_tmp, n,
                                       begin_bit, end_bit, digit_bits, sc, s);
    if (key_bits == 64)
        return dispatch_bits<uint64_t>(keys_in, values_in, keys_out, values_out, keys_tmp, values_tmp, n,
                                       begin_bit, end_bit, digit_bits, sc, s);
    return hipErrorInvalidValue;
}

}  // namespace nxb
