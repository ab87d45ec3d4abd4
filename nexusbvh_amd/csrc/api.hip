// api.hip - host orchestration and the C ABI of libnxb.so.
//
// Replaces the host side of the reference (NexusBVH/src/BVHBuilder.cpp:13-294,
// NexusBVH/src/Cuda/CudaUtils.h): allocation, initialisation, launches,
// per-stage timing and SAH metrics. Differences by design:
//   * errors are returned as nxb_status (the reference prints and exit(99)s);
//   * scratch comes from a per-device workspace that is grown on demand and
//     reused across builds (the reference cudaMallocAsync's it per build);
//     output blocks released by nxb_free_device_* are kept and reused;
//   * all initialisation is done by small fill kernels (graph-capturable);
//   * per-stage times come from events recorded once per build and read
//     after a single synchronisation.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/nxb.h"
#include "device_common.h"
#include "nxb_internal.h"

namespace nxb {

namespace {

constexpr uint32_t kDefaultSpinLimit = 1u << 22;
constexpr size_t kAlign = 256;

inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

nxb_status from_hip(hipError_t e) {
    if (e == hipSuccess) return NXB_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return NXB_ERROR_OUT_OF_MEMORY;
    return NXB_ERROR_HIP;
}

#define NXB_HIP(expr)                                       \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return from_hip(e_);          \
    } while (0)

#define NXB_TRY(expr)                                       \
    do {                                                    \
        nxb_status st_ = (expr);                            \
        if (st_ != NXB_SUCCESS) return st_;                 \
    } while (0)

bool fail_alloc_injected() {
    const char* v = std::getenv("NXB_FAIL_ALLOC");
    return v && v[0] == '1';
}

// BVH8 levels enqueued before the first counter readback. NXB_BVH8_FIRST_LEVELS
// (1..96) fixes the count; tests use it to reach the path that adds levels
// after the readback. Otherwise: one more than the depth of the device's last
// tree when that had within 1/8 as many primitives (rebuilds of one scene),
// else kBvh8FirstLevels. A level launch past the end of the tree costs ~4.5 us
// on the GPU, one too few a readback round trip.
uint32_t bvh8_first_levels(uint32_t n, uint32_t last_n, uint32_t last_depth) {
    if (const char* v = std::getenv("NXB_BVH8_FIRST_LEVELS")) {
        const long x = std::strtol(v, nullptr, 10);
        return x >= 1 && x < (long)kBvh8LevelWords ? (uint32_t)x : kBvh8FirstLevels;
    }
    const uint32_t tol = n / 8;
    if (last_depth && last_n + tol >= n && last_n <= n + tol)
        return last_depth + 1 < kBvh8LevelWords ? last_depth + 1 : kBvh8LevelWords - 1;
    return kBvh8FirstLevels;
}

// Scratch carved from one device buffer. Sizes depend on (n, key bytes).
struct Layout {
    size_t header, partials, hist, counts, keys_a, keys_b, keys_tmp, vals_b, vals_tmp, parent, recs, csizes, nsize,
        parent2, work, work2, feat, rtable, rec8;
    size_t counts_bytes;
    size_t total;
};

constexpr uint32_t kMaxBoundsBlocks = 2048;

// Header words (uint32 unless noted):
//   [0] bounds done counter   [1] flags   [2] H-PLOC moved-node parent
//   [3] H-PLOC moved-node size   [4] learned-radius fallbacks
//   [16..22) scene box (6 floats)
//   [32..36) two FP64 SAH accumulators   [48..56) H-PLOC spare node slot
//   [64..80) H-PLOC spare wide record (BVH8 builds; 64-byte aligned)
//   [128..130) BVH8 node counter (finished << 32 | allocated), own cache line
//   [160] BVH8 work grab counter, own cache line
//   [256..512) H-PLOC chunk counters
//   [512..609) BVH8 level counters (level-synchronous collapse)
constexpr size_t kHdrDone = 0, kHdrFlags = 1, kHdrMovedParent = 2, kHdrMovedSize = 3, kHdrFallbacks = 4,
                 kHdrScene = 16, kHdrHplocChunks = 256,
                 kHdrSah = 32, kHdrMoved = 48, kHdrMovedRec = 64, kHdrBvh8Ctr = 128, kHdrBvh8Grab = 160,
                 kHdrBvh8Levels = 512;
constexpr size_t kHeaderBytes = 4096;
// header words read back after a build: flags, scene box, BVH8 level counters
constexpr size_t kReadbackWords = kHdrBvh8Levels + kBvh8LevelWords;
static_assert(kHdrHplocChunks + kHplocCounterWords <= kHdrBvh8Levels, "header words overlap");

// full: the layout of a whole build (key and value double buffers, no
// second parent table); bvh8: with the BVH8 collapse's work arrays; rec8:
// with the wide records of the internal nodes (a BVH8 build's H-PLOC writes
// them for its collapse).
Layout make_layout(uint32_t n, uint32_t key_bytes, bool full, bool bvh8 = true, bool rec8 = false) {
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align_up(bytes ? bytes : 1); return o; };
    L.header = take(kHeaderBytes);
    L.partials = take(kMaxBoundsBlocks * 32);
    L.hist = take(kHistCopies * kHistCopyWords * 4);
    L.counts_bytes = sort_scratch_bytes(n, key_bytes * 8, 8);
    L.counts = take(L.counts_bytes);
    L.keys_a = full ? take((size_t)n * key_bytes) : 0;
    L.keys_b = full ? take((size_t)n * key_bytes) : 0;
    L.keys_tmp = take((size_t)n * key_bytes);
    L.vals_b = full ? take((size_t)n * 4) : 0;
    L.vals_tmp = take((size_t)n * 4);
    L.parent = take((size_t)n * 4);
    L.recs = take((size_t)n * 32);
    L.csizes = take((size_t)n * 4);
    L.nsize = take((size_t)n * 4);
    L.parent2 = full ? 0 : take((size_t)2 * n * 4);
    L.work = bvh8 ? take((size_t)nxb_bvh8_capacity(n) * 8) : 0;
    L.work2 = bvh8 ? take((size_t)nxb_bvh8_capacity(n) * 8) : 0;
    // learned radius: features and radius table for the smallest segments (2^8)
    const size_t segs = ((size_t)n + 255) / 256;
    L.feat = take(segs * NXB_MLP_FEATURES * sizeof(float));
    L.rtable = take(segs);
    L.rec8 = rec8 && n > 1 ? take((size_t)(n - 1) * 64) : 0;
    L.total = off;
    return L;
}

struct Carved {
    uint8_t* base;
    Layout L;
    uint32_t* hdr() const { return reinterpret_cast<uint32_t*>(base + L.header); }
    uint32_t* done() const { return hdr() + kHdrDone; }
    uint32_t* flags() const { return hdr() + kHdrFlags; }
    uint32_t* moved_parent() const { return hdr() + kHdrMovedParent; }
    nxb_bvh2_node* moved() const { return reinterpret_cast<nxb_bvh2_node*>(hdr() + kHdrMoved); }
    uint32_t* recs() const { return reinterpret_cast<uint32_t*>(base + L.recs); }
    unsigned long long* bvh8_ctr() const { return reinterpret_cast<unsigned long long*>(hdr() + kHdrBvh8Ctr); }
    uint32_t* bvh8_grab() const { return hdr() + kHdrBvh8Grab; }
    uint32_t* moved_size() const { return hdr() + kHdrMovedSize; }
    uint32_t* hploc_chunks() const { return hdr() + kHdrHplocChunks; }
    uint32_t* csizes() const { return reinterpret_cast<uint32_t*>(base + L.csizes); }
    uint32_t* nsize() const { return reinterpret_cast<uint32_t*>(base + L.nsize); }
    uint32_t* parent2() const { return reinterpret_cast<uint32_t*>(base + L.parent2); }
    nxb_aabb* scene() const { return reinterpret_cast<nxb_aabb*>(hdr() + kHdrScene); }
    double* sah() const { return reinterpret_cast<double*>(hdr() + kHdrSah); }
    Box* partials() const { return reinterpret_cast<Box*>(base + L.partials); }
    uint32_t* hist() const { return reinterpret_cast<uint32_t*>(base + L.hist); }
    SortScratch sort() const {
        return SortScratch{hist(), reinterpret_cast<uint32_t*>(base + L.counts), L.counts_bytes / sizeof(uint32_t)};
    }
    void* keys_a() const { return base + L.keys_a; }
    void* keys_b() const { return base + L.keys_b; }
    void* keys_tmp() const { return base + L.keys_tmp; }
    uint32_t* vals_b() const { return reinterpret_cast<uint32_t*>(base + L.vals_b); }
    uint32_t* vals_tmp() const { return reinterpret_cast<uint32_t*>(base + L.vals_tmp); }
    uint32_t* parent() const { return reinterpret_cast<uint32_t*>(base + L.parent); }
    uint64_t* work() const { return reinterpret_cast<uint64_t*>(base + L.work); }
    uint64_t* work2() const { return reinterpret_cast<uint64_t*>(base + L.work2); }
    float* feat() const { return reinterpret_cast<float*>(base + L.feat); }
    uint8_t* rtable() const { return base + L.rtable; }
    uint4* rec8() const { return L.rec8 ? reinterpret_cast<uint4*>(base + L.rec8) : nullptr; }
    uint4* moved_rec() const { return reinterpret_cast<uint4*>(hdr() + kHdrMovedRec); }
};

// Per-device cached workspace for the builder entry points.
struct DeviceState {
    uint8_t* ws = nullptr;
    size_t ws_bytes = 0;
    uint32_t* host = nullptr;  // pinned readback buffer (kReadbackWords words)
    uint32_t last_n = 0, last_depth = 0;  // primitives and BVH8 levels of the last build
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // stage timing (builds hold g_mutex)
    // Stage entry points that return before their kernels finish
    // (nxb_compute_scene_bounds, nxb_radix_sort_pairs without a workspace of
    // the caller's) record ws_done on their stream; the next user of ws, on
    // any stream, waits for it. Calls that synchronize their stream after
    // that wait clear ws_pending.
    hipEvent_t ws_done = nullptr;
    bool ws_pending = false;
};

std::mutex g_mutex;
std::vector<DeviceState> g_dev;
uint32_t g_learned_segments = 0, g_learned_fallbacks = 0;  // last learned-policy build (under g_mutex)

DeviceState& dev_state(int dev) {
    if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
    return g_dev[dev];
}

nxb_status ensure_workspace(DeviceState& d, size_t bytes) {
    if (!d.host) NXB_HIP(hipHostMalloc((void**)&d.host, kReadbackWords * sizeof(uint32_t), hipHostMallocDefault));
    if (d.ws_bytes >= bytes) return NXB_SUCCESS;
    if (d.ws) {
        NXB_HIP(hipDeviceSynchronize());
        NXB_HIP(hipFree(d.ws));
        d.ws = nullptr;
        d.ws_bytes = 0;
        d.ws_pending = false;
    }
    size_t want = bytes + bytes / 8;
    NXB_HIP(hipMalloc((void**)&d.ws, want));
    d.ws_bytes = want;
    return NXB_SUCCESS;
}

// Orders work enqueued next on s after the asynchronous stage calls that may
// still be using d.ws.
nxb_status ws_wait(DeviceState& d, hipStream_t s) {
    if (d.ws_pending) NXB_HIP(hipStreamWaitEvent(s, d.ws_done, 0));
    return NXB_SUCCESS;
}

// Called after enqueueing asynchronous work that uses d.ws on s.
nxb_status ws_mark(DeviceState& d, hipStream_t s) {
    if (!d.ws_done) NXB_HIP(hipEventCreateWithFlags(&d.ws_done, hipEventDisableTiming));
    NXB_HIP(hipEventRecord(d.ws_done, s));
    d.ws_pending = true;
    return NXB_SUCCESS;
}

nxb_status check_config(const nxb_build_config* c) {
    if (!c) return NXB_SUCCESS;
    if (c->morton_bits != 0 && c->morton_bits != 32 && c->morton_bits != 64) return NXB_ERROR_INVALID_VALUE;
    if (c->search_radius < 1 || c->search_radius > 16) return NXB_ERROR_INVALID_VALUE;
    if (c->merge_mode != NXB_MERGE_WAVE32 && c->merge_mode != NXB_MERGE_WAVE64) return NXB_ERROR_INVALID_VALUE;
    if (c->radius_policy != NXB_RADIUS_FIXED && c->radius_policy != NXB_RADIUS_TABLE &&
        c->radius_policy != NXB_RADIUS_LEARNED)
        return NXB_ERROR_INVALID_VALUE;
    if (c->radius_policy == NXB_RADIUS_TABLE && (!c->d_radius_table || c->segment_log2 > 30))
        return NXB_ERROR_INVALID_VALUE;
    if (c->radius_policy == NXB_RADIUS_LEARNED &&
        (!c->d_mlp_params || c->segment_log2 < 8 || c->segment_log2 > 20))
        return NXB_ERROR_INVALID_VALUE;
    if (c->sort_digit_bits != 0 && (c->sort_digit_bits < 6 || c->sort_digit_bits > 9)) return NXB_ERROR_INVALID_VALUE;
    if (c->slot_assignment != NXB_SLOT_GREEDY && c->slot_assignment != NXB_SLOT_AUCTION) return NXB_ERROR_INVALID_VALUE;
    if (c->hploc_blocks_per_cu > kHplocMaxBlocksPerCu) return NXB_ERROR_INVALID_VALUE;
    return NXB_SUCCESS;
}

uint32_t morton_bits_of(const nxb_build_config& c) {
    if (c.morton_bits) return c.morton_bits;
    return c.prioritize_speed ? 32u : 64u;  // BVHBuilder.cpp:160-163
}

RadiusParams radius_params(const nxb_build_config& c) {
    RadiusParams rp;
    rp.radius = c.search_radius;
    rp.table = c.radius_policy == NXB_RADIUS_TABLE ? c.d_radius_table : nullptr;
    rp.seg_log2 = c.segment_log2;
    rp.table_len = c.radius_table_len;
    return rp;
}

// Sort window of the pipeline: bits [2,32) / [1,64) (Setup.cu:74-78). The top
// bits 30-31 / 63 of a Morton code are always zero, so [2,30) / [1,63) give
// the same permutation with fewer digits.
void morton_sort_window(uint32_t bits, uint32_t& b, uint32_t& e) {
    if (bits == 32) { b = 2; e = 30; } else { b = 1; e = 63; }
}

uint32_t digit_bits_of(const nxb_build_config& c, uint32_t key_bits, uint32_t n) {
    if (c.sort_digit_bits) return c.sort_digit_bits;
    return default_digit_bits(n, key_bits);
}

// The stage-timing events of a build: the device's cached set, recorded only
// when metrics are requested.
struct Events {
    hipEvent_t* e = nullptr;
    bool on = false;
    nxb_status create(bool want, hipEvent_t* cached) {
        on = want;
        if (!on) return NXB_SUCCESS;
        e = cached;
        // Created once per device. Timing only, read after a stream
        // synchronize: no system-scope fence (with it, each record cost a
        // ~5.5 us bubble between kernels).
        for (int i = 0; i < 6; ++i)
            if (!e[i]) NXB_HIP(hipEventCreateWithFlags(&e[i], hipEventDisableSystemFence));
        return NXB_SUCCESS;
    }
    void record(int i, hipStream_t s) { if (on) (void)hipEventRecord(e[i], s); }
    float ms(int a, int b) {
        float t = 0.0f;
        if (on) (void)hipEventElapsedTime(&t, e[a], e[b]);
        return t;
    }
};

// Scratch initialisation of a BVH2 build (outside the timed stages, as the
// reference's memsets are, BVHBuilder.cpp:22-24).
nxb_status prep_bvh2(uint32_t n, const nxb_build_config& cfg, const Carved& w, hipStream_t s) {
    const uint32_t bits = morton_bits_of(cfg);
    uint32_t b, e;
    morton_sort_window(bits, b, e);
    const SortPlan pl = plan_sort(n, bits, b, e, digit_bits_of(cfg, bits, n));
    // header, parent table and the sort's digit histograms in one launch
    NXB_HIP(launch_build_prep(w.hdr(), kHdrHplocChunks + kHplocCounterWords, w.parent(), n, w.hist(),
                              pl.passes << pl.bits, s));
    return NXB_SUCCESS;
}

// The pipeline stages up to the BVH2. On success nodes hold the BVH2 and the
// carved workspace header holds scene box and flags (not yet read back).
nxb_status run_bvh2(const void* prims, nxb_prim_type type, uint32_t n, const nxb_build_config& cfg,
                    nxb_bvh2_node* nodes, const Carved& w, Events& ev, hipStream_t s, const uint32_t* indices,
                    uint32_t vertex_count) {
    const uint32_t bits = morton_bits_of(cfg);
    uint32_t b, e;
    morton_sort_window(bits, b, e);
    ev.record(0, s);
    BoundsScratch bs{w.partials(), w.done(), kMaxBoundsBlocks};
    NXB_HIP(launch_scene_bounds(prims, type, n, nodes, w.scene(), w.flags(), bs, s, indices, vertex_count));
    ev.record(1, s);
    // Morton codes with the sort's digit histograms (zeroed in prep_bvh2).
    SortPlan pl = plan_sort(n, bits, b, e, digit_bits_of(cfg, bits, n));
    NXB_HIP(launch_morton_hist(nodes, n, w.scene(), w.keys_a(), pl, w.hist(), s));
    ev.record(2, s);
    NXB_HIP(launch_sort_passes(w.keys_a(), nullptr, w.keys_b(), w.vals_b(), w.keys_tmp(), w.vals_tmp(), n, pl,
                               w.sort(), s));
    ev.record(3, s);
    RadiusParams rp = radius_params(cfg);
    if (cfg.radius_policy == NXB_RADIUS_LEARNED) {
        // segment features -> MLP -> radius table (timed with the BVH2 stage)
        const uint32_t segs = ((n - 1) >> cfg.segment_log2) + 1;
        NXB_HIP(launch_segment_features(w.keys_b(), bits, w.vals_b(), nodes, n, cfg.segment_log2, w.feat(), s));
        NXB_HIP(launch_mlp_table(cfg.d_mlp_params, w.feat(), segs, cfg.mlp_min_confidence, w.rtable(), nullptr,
                                 w.hdr() + kHdrFallbacks, s));
        rp.table = w.rtable();
        rp.table_len = segs;
    }
    // no subtree sizes: a BVH8 build's collapse takes them from the wide
    // records, and a BVH2 build returns nodes only
    HplocScratch hs{w.parent(), w.recs(), w.csizes(), w.moved(), w.moved_parent(), w.moved_size(), nullptr,
                    w.flags(), w.hploc_chunks()};
    hs.rec8 = w.rec8();  // a BVH8 build's layout: the collapse reads these records
    hs.moved_rec = w.moved_rec();
    NXB_HIP(launch_hploc(w.keys_b(), bits, w.vals_b(), n, nodes, cfg.merge_mode, rp, cfg.hploc_blocks_per_cu, hs, s));
    ev.record(4, s);
    return NXB_SUCCESS;
}

nxb_status flags_status(uint32_t f) {
    if (f & kFlagNonFinite) return NXB_ERROR_NONFINITE_INPUT;
    if (f & kFlagTimeout) return NXB_ERROR_TIMEOUT;
    if (f & kFlagBadInput) return NXB_ERROR_INVALID_VALUE;
    if (f & kFlagBadIndex) return NXB_ERROR_HIP;
    return NXB_SUCCESS;
}

nxb_status validate_build(const void* prims, nxb_prim_type type, uint32_t n, const nxb_build_config* cfg,
                          const void* out) {
    if (!out) return NXB_ERROR_INVALID_VALUE;
    if (type != NXB_PRIM_AABB && type != NXB_PRIM_TRIANGLE) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n > 0 && !prims) return NXB_ERROR_INVALID_VALUE;
    return check_config(cfg);
}

// Device memory of build outputs (and of the BVH2 a BVH8 build discards):
// released blocks are kept per device and handed out again. On MI355X,
// hipFree of the three output buffers of a build, or their release through
// the stream-ordered pool, costs ~0.3 ms per build; reusing a block costs
// nothing. Sizes are rounded up to 2 MiB; a request takes the smallest idle
// block of at least its size and at most twice it. Idle memory is capped.
struct OutCache {
    struct Block { size_t bytes; int device; };
    std::mutex mu;
    std::unordered_map<void*, Block> live;
    std::vector<std::multimap<size_t, void*>> idle;   // per device
    size_t idle_bytes = 0;
    static constexpr size_t kGranule = 2ull << 20;
    static constexpr size_t kIdleLimit = 8ull << 30;
};
OutCache g_out;

void flush_idle_locked(int dev) {
    if (dev < 0 || (size_t)dev >= g_out.idle.size()) return;
    for (auto& kv : g_out.idle[dev]) {
        (void)hipFree(kv.second);
        g_out.idle_bytes -= kv.first;
    }
    g_out.idle[dev].clear();
}

nxb_status alloc_out(void** p, size_t bytes, hipStream_t) {
    *p = nullptr;
    if (fail_alloc_injected()) return NXB_ERROR_OUT_OF_MEMORY;
    int dev = 0;
    NXB_HIP(hipGetDevice(&dev));
    const size_t want = (bytes + OutCache::kGranule - 1) / OutCache::kGranule * OutCache::kGranule;
    std::lock_guard<std::mutex> lock(g_out.mu);
    if ((size_t)dev >= g_out.idle.size()) g_out.idle.resize(dev + 1);
    auto& idle = g_out.idle[dev];
    auto it = idle.lower_bound(want);
    if (it != idle.end() && it->first <= 2 * want) {
        *p = it->second;
        g_out.live[it->second] = OutCache::Block{it->first, dev};
        g_out.idle_bytes -= it->first;
        idle.erase(it);
        return NXB_SUCCESS;
    }
    hipError_t e = hipMalloc(p, want);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        flush_idle_locked(dev);
        e = hipMalloc(p, want);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return from_hip(e);
    }
    g_out.live[*p] = OutCache::Block{want, dev};
    return NXB_SUCCESS;
}

// Returns a block from alloc_out; the caller guarantees no queued work uses it.
void release_out(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_out.mu);
    auto it = g_out.live.find(p);
    if (it == g_out.live.end()) {
        (void)hipFree(p);  // not ours
        return;
    }
    OutCache::Block b = it->second;
    g_out.live.erase(it);
    if ((size_t)b.device >= g_out.idle.size()) g_out.idle.resize(b.device + 1);
    g_out.idle[b.device].emplace(b.bytes, p);
    g_out.idle_bytes += b.bytes;
    while (g_out.idle_bytes > OutCache::kIdleLimit) {
        // drop the largest idle blocks first
        int dmax = -1;
        size_t smax = 0;
        for (size_t d = 0; d < g_out.idle.size(); ++d)
            if (!g_out.idle[d].empty() && g_out.idle[d].rbegin()->first > smax) {
                smax = g_out.idle[d].rbegin()->first;
                dmax = (int)d;
            }
        if (dmax < 0) break;
        auto last = std::prev(g_out.idle[dmax].end());
        (void)hipFree(last->second);
        g_out.idle_bytes -= last->first;
        g_out.idle[dmax].erase(last);
    }
}

// Output blocks of a build in flight: on every early return they go back to
// the cache after the stream has drained; a successful build hands them to
// the caller with keep().
struct OutBlocks {
    hipStream_t s;
    nxb_bvh2_node* nodes = nullptr;
    nxb_bvh8_node* n8 = nullptr;
    uint32_t* prim_idx = nullptr;
    bool kept = false;
    explicit OutBlocks(hipStream_t st) : s(st) {}
    OutBlocks(const OutBlocks&) = delete;
    OutBlocks& operator=(const OutBlocks&) = delete;
    void keep() { kept = true; }
    ~OutBlocks() {
        if (kept) return;
        (void)hipStreamSynchronize(s);
        release_out(prim_idx);
        release_out(n8);
        release_out(nodes);
    }
};

// indices (nullable): an indexed triangle mesh, prims then holding
// vertex_count vertices of 3 floats.
nxb_status build_common(const void* prims, nxb_prim_type type, uint32_t n, const nxb_build_config* cfgp,
                        nxb_bvh2* out2, nxb_bvh8* out8, nxb_metrics* metrics, hipStream_t s,
                        const uint32_t* indices = nullptr, uint32_t vertex_count = 0) {
    nxb_build_config cfg;
    nxb_default_config(&cfg);
    if (cfgp) cfg = *cfgp;
    const uint32_t bits = morton_bits_of(cfg);
    int dev = 0;
    NXB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState& d = dev_state(dev);
    Layout L = make_layout(n, bits / 8, true, true, out8 != nullptr);
    nxb_status st = ensure_workspace(d, L.total);
    if (st) return st;
    Carved w{d.ws, L};
    st = ws_wait(d, s);
    if (st) return st;

    OutBlocks ob(s);
    st = alloc_out((void**)&ob.nodes, (size_t)(2 * n - 1) * sizeof(nxb_bvh2_node), s);
    if (st) return st;
    Events ev;
    st = ev.create(metrics != nullptr, d.ev);
    if (st) return st;

    uint32_t cap = nxb_bvh8_capacity(n);
    Bvh8Scratch bs{w.work(), w.bvh8_ctr(), w.bvh8_grab(), w.flags(), w.nsize(), w.work2(), cfg.slot_assignment,
                   w.rec8()};
    st = prep_bvh2(n, cfg, w, s);
    if (!st && out8) {
        st = alloc_out((void**)&ob.n8, (size_t)cap * sizeof(nxb_bvh8_node), s);
        if (!st) st = alloc_out((void**)&ob.prim_idx, (size_t)n * sizeof(uint32_t), s);
    }
    if (!st) st = run_bvh2(prims, type, n, cfg, ob.nodes, w, ev, s, indices, vertex_count);
    // level-synchronous BVH8: the first levels here; their counters come back
    // with the header below, and deeper trees get more levels there
    uint32_t* level_cnt = w.hdr() + kHdrBvh8Levels;
    uint32_t launched = 0, node8 = 0;
    if (!st && out8) {
        launched = max(bvh8_first_levels(n, d.last_n, d.last_depth), kBvh8TopLevels);
        st = from_hip(launch_bvh8_levels(ob.nodes, n, ob.n8, ob.prim_idx, cap, bs, level_cnt, 0, launched, s));
    }
    if (st) return st;
    ev.record(5, s);
    // Read back flags, scene box and BVH8 level counters in one copy.
    NXB_HIP(hipMemcpyAsync(d.host, w.hdr(), kReadbackWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NXB_HIP(hipStreamSynchronize(s));
    d.ws_pending = false;  // s waited for the pending stage work and has drained
    while (out8 && !bvh8_levels_done(n, d.host + kHdrBvh8Levels, launched, &node8)) {
        if (launched >= kBvh8LevelWords - 1) {
            d.host[kHdrFlags] |= kFlagBadIndex;  // deeper than the level counters reach
            break;
        }
        NXB_HIP(launch_bvh8_levels(ob.nodes, n, ob.n8, ob.prim_idx, cap, bs, level_cnt, launched, 16, s));
        launched += 16;
        ev.record(5, s);
        NXB_HIP(hipMemcpyAsync(d.host, w.hdr(), kReadbackWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NXB_HIP(hipStreamSynchronize(s));
    }
    if (out8 && node8) {
        uint32_t depth = 0;
        while (depth < kBvh8LevelWords && d.host[kHdrBvh8Levels + depth] != 0) ++depth;
        d.last_n = n;
        d.last_depth = depth;
    }
    uint32_t flags = d.host[kHdrFlags];
    nxb_aabb scene;
    std::memcpy(&scene, d.host + kHdrScene, sizeof(scene));

    if (cfg.radius_policy == NXB_RADIUS_LEARNED) {
        g_learned_segments = ((n - 1) >> cfg.segment_log2) + 1;
        g_learned_fallbacks = d.host[kHdrFallbacks];
    }
    st = flags_status(flags);
    if (st) return st;

    if (metrics) {
        std::memset(metrics, 0, sizeof(*metrics));
        metrics->compute_scene_bounds_time = ev.ms(0, 1);
        metrics->compute_morton_codes_time = ev.ms(1, 2);
        metrics->radix_sort_time = ev.ms(2, 3);
        metrics->bvh_build_time = ev.ms(3, 4);
        metrics->total_time = metrics->compute_scene_bounds_time + metrics->compute_morton_codes_time +
                              metrics->radix_sort_time + metrics->bvh_build_time;
        if (out8) {
            metrics->bvh8_conversion_time = ev.ms(4, 5);
            metrics->total_time += metrics->bvh8_conversion_time;
            metrics->average_child_per_node = (float)(n + node8 - 1) / (float)node8;
        }
    }
    if (metrics && !(cfg.metrics_flags & NXB_METRICS_TIMES_ONLY)) {
        // SAH of the BVH2 (BVHBuilder.cpp:92-103) and BVH8 (237-255).
        NXB_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(w.sah()), 0u, 4, s));
        if (out8 && w.rec8())  // a BVH8 build keeps its internal nodes as wide records only
            NXB_HIP(launch_sah_rec8(w.rec8(), n, w.scene(), w.sah(), s));
        else
            NXB_HIP(launch_sah_bvh2(ob.nodes, 2 * n - 1, w.scene(), w.sah(), s));
        if (out8) NXB_HIP(launch_sah_bvh8(ob.n8, node8, w.scene(), w.sah() + 1, s));
        double costs[2];
        NXB_HIP(hipMemcpyAsync(costs, w.sah(), sizeof(costs), hipMemcpyDeviceToHost, s));
        NXB_HIP(hipStreamSynchronize(s));
        metrics->bvh2_cost = (float)costs[0];
        // averageChildPerNode assumes one primitive per leaf (BVHBuilder.cpp:252-254).
        if (out8) metrics->bvh8_cost = (float)costs[1];
    }

    ob.keep();
    if (out8) {
        release_out(ob.nodes);  // FreeDeviceBVH(bvh2) (BVHBuilder.cpp:257); the stream is synchronized
        out8->nodes = ob.n8;
        out8->node_count = node8;
        out8->prim_idx = ob.prim_idx;
        out8->prim_count = n;
        out8->bounds = scene;
    } else {
        out2->nodes = ob.nodes;
        out2->node_count = 2 * n - 1;
        out2->prim_count = n;
        out2->bounds = scene;
    }
    return NXB_SUCCESS;
}

// ---- allocation-free BVH2 build ------------------------------------------
// Scratch of nxb_build_bvh2_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends. Host code must not pass `scratch` itself
// to a kernel (in host code the name is a stand-in, not a device address):
// its device address is looked up once per device with hipGetSymbolAddress
// and only that address is used.
constexpr size_t kNoAllocScratchBytes = (size_t)NXB_NOALLOC_SCRATCH_MIB << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
uint8_t* g_noalloc_scratch[kNoAllocMaxDevices] = {};  // device address per device (under g_mutex)

// The calling thread's current device's copy of `scratch`.
nxb_status noalloc_scratch(uint8_t** out) {
    int dev = 0;
    NXB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return NXB_ERROR_NOT_SUPPORTED;
    if (!g_noalloc_scratch[dev]) {
        void* ptr = nullptr;
        NXB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return NXB_ERROR_HIP;
        g_noalloc_scratch[dev] = static_cast<uint8_t*>(ptr);
    }
    *out = g_noalloc_scratch[dev];
    return NXB_SUCCESS;
}

// nxb_build_bvh2 without any allocation: the caller's output array, the
// static scratch array, no timing events and a readback into host stack
// memory.
nxb_status build_bvh2_noalloc(const void* prims, nxb_prim_type type, uint32_t n, const nxb_build_config& cfg,
                              nxb_bvh2_node* nodes, nxb_bvh2* out, hipStream_t s) {
    const uint32_t bits = morton_bits_of(cfg);
    const Layout L = make_layout(n, bits / 8, true, false);
    if (L.total > kNoAllocScratchBytes) return NXB_ERROR_TOO_LARGE;
    std::lock_guard<std::mutex> lock(g_mutex);
    uint8_t* base = nullptr;
    nxb_status st = noalloc_scratch(&base);
    if (st) return st;
    const Carved w{base, L};
    Events ev;  // off: no events are created or recorded
    st = prep_bvh2(n, cfg, w, s);
    if (!st) st = run_bvh2(prims, type, n, cfg, nodes, w, ev, s, nullptr, 0);
    if (st) {
        (void)hipStreamSynchronize(s);
        return st;
    }
    uint32_t hdr[kHdrScene + 6];
    NXB_HIP(hipMemcpyAsync(hdr, w.hdr(), sizeof(hdr), hipMemcpyDeviceToHost, s));
    NXB_HIP(hipStreamSynchronize(s));
    if (cfg.radius_policy == NXB_RADIUS_LEARNED) {
        g_learned_segments = ((n - 1) >> cfg.segment_log2) + 1;
        g_learned_fallbacks = hdr[kHdrFallbacks];
    }
    st = flags_status(hdr[kHdrFlags]);
    if (st) return st;
    out->nodes = nodes;
    out->node_count = 2 * n - 1;
    out->prim_count = n;
    std::memcpy(&out->bounds, hdr + kHdrScene, sizeof(nxb_aabb));
    return NXB_SUCCESS;
}

hipStream_t as_hip(nxb_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Carves the workspace of a stage entry point: the caller's buffer, or else
// the device's cached one. In the second case *cached is set, and the entry
// point orders its work after pending asynchronous stage work (ws_wait) and
// then either marks its own (ws_mark) or synchronizes and clears ws_pending.
nxb_status stage_workspace(void* d_ws, size_t bytes, uint32_t n, uint32_t key_bits, Carved& w,
                           DeviceState** cached) {
    Layout L = make_layout(n, key_bits / 8, false);
    *cached = nullptr;
    if (d_ws) {
        if (bytes < L.total) return NXB_ERROR_INVALID_VALUE;
        w = Carved{static_cast<uint8_t*>(d_ws), L};
        return NXB_SUCCESS;
    }
    int dev = 0;
    NXB_HIP(hipGetDevice(&dev));
    DeviceState& d = dev_state(dev);
    nxb_status st = ensure_workspace(d, L.total);
    if (st) return st;
    w = Carved{d.ws, L};
    *cached = &d;
    return NXB_SUCCESS;
}

}  // namespace

uint32_t device_cu_count() {
    static std::atomic<int> cached{0};
    int v = cached.load();
    if (v > 0) return (uint32_t)v;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    cached.store(cus);
    return (uint32_t)cus;
}

}  // namespace nxb

using namespace nxb;

extern "C" {

void nxb_default_config(nxb_build_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->prioritize_speed = 0;  // BuildConfig.h:11
    cfg->morton_bits = 0;
    cfg->search_radius = 8;     // BinaryBuilder.cu:9
    cfg->merge_mode = NXB_MERGE_WAVE32;
    cfg->radius_policy = NXB_RADIUS_FIXED;
    cfg->segment_log2 = 12;
    cfg->mlp_min_confidence = 0.5f;
}

const char* nxb_status_string(nxb_status s) {
    switch (s) {
        case NXB_SUCCESS: return "success";
        case NXB_ERROR_INVALID_VALUE: return "invalid value";
        case NXB_ERROR_OUT_OF_MEMORY: return "out of memory";
        case NXB_ERROR_NONFINITE_INPUT: return "non-finite primitive coordinate";
        case NXB_ERROR_TOO_LARGE: return "too many primitives";
        case NXB_ERROR_TIMEOUT: return "device wait exceeded its spin limit";
        case NXB_ERROR_HIP: return "HIP runtime error";
        case NXB_ERROR_NOT_SUPPORTED: return "not supported";
    }
    return "unknown status";
}

const char* nxb_version(void) { return "nxb-mi355x 0.1.0 (gfx950)"; }

uint32_t nxb_bvh8_capacity(uint32_t n) {
    if (n == 0) return 0;
    return (uint32_t)(1 + (4ull * n - 2) / 7);  // DivideRoundUp(4n - 1, 7)
}

nxb_status nxb_build_bvh2(const void* d_prims, nxb_prim_type type, uint32_t prim_count,
                          const nxb_build_config* cfg, nxb_bvh2* out, nxb_metrics* metrics, nxb_stream_t stream) {
    nxb_status st = validate_build(d_prims, type, prim_count, cfg, out);
    if (st) return st;
    std::memset(out, 0, sizeof(*out));
    if (metrics) std::memset(metrics, 0, sizeof(*metrics));
    if (prim_count == 0) return NXB_SUCCESS;
    return build_common(d_prims, type, prim_count, cfg, out, nullptr, metrics, as_hip(stream));
}

nxb_status nxb_build_bvh8(const void* d_prims, nxb_prim_type type, uint32_t prim_count,
                          const nxb_build_config* cfg, nxb_bvh8* out, nxb_metrics* metrics, nxb_stream_t stream) {
    nxb_status st = validate_build(d_prims, type, prim_count, cfg, out);
    if (st) return st;
    std::memset(out, 0, sizeof(*out));
    if (metrics) std::memset(metrics, 0, sizeof(*metrics));
    if (prim_count == 0) return NXB_SUCCESS;
    return build_common(d_prims, type, prim_count, cfg, nullptr, out, metrics, as_hip(stream));
}

nxb_status nxb_build_bvh2_noalloc(const void* d_prims, nxb_prim_type type, uint32_t prim_count,
                                  const nxb_build_config* cfg, nxb_bvh2_node* d_nodes, nxb_bvh2* out,
                                  nxb_stream_t stream) {
    nxb_status st = validate_build(d_prims, type, prim_count, cfg, out);
    if (st) return st;
    std::memset(out, 0, sizeof(*out));
    if (prim_count > NXB_NOALLOC_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (prim_count == 0) return NXB_SUCCESS;
    if (!d_nodes) return NXB_ERROR_INVALID_VALUE;
    nxb_build_config c;
    nxb_default_config(&c);
    if (cfg) c = *cfg;
    return build_bvh2_noalloc(d_prims, type, prim_count, c, d_nodes, out, as_hip(stream));
}

nxb_status nxb_build_bvh2_indexed(const float* d_vertices, uint32_t vertex_count, const uint32_t* d_indices,
                                  uint32_t tri_count, const nxb_build_config* cfg, nxb_bvh2* out,
                                  nxb_metrics* metrics, nxb_stream_t stream) {
    nxb_status st = validate_build(d_vertices, NXB_PRIM_TRIANGLE, tri_count, cfg, out);
    if (st) return st;
    if (tri_count && (!d_indices || vertex_count == 0)) return NXB_ERROR_INVALID_VALUE;
    std::memset(out, 0, sizeof(*out));
    if (metrics) std::memset(metrics, 0, sizeof(*metrics));
    if (tri_count == 0) return NXB_SUCCESS;
    return build_common(d_vertices, NXB_PRIM_TRIANGLE, tri_count, cfg, out, nullptr, metrics, as_hip(stream),
                        d_indices, vertex_count);
}

nxb_status nxb_build_bvh8_indexed(const float* d_vertices, uint32_t vertex_count, const uint32_t* d_indices,
                                  uint32_t tri_count, const nxb_build_config* cfg, nxb_bvh8* out,
                                  nxb_metrics* metrics, nxb_stream_t stream) {
    nxb_status st = validate_build(d_vertices, NXB_PRIM_TRIANGLE, tri_count, cfg, out);
    if (st) return st;
    if (tri_count && (!d_indices || vertex_count == 0)) return NXB_ERROR_INVALID_VALUE;
    std::memset(out, 0, sizeof(*out));
    if (metrics) std::memset(metrics, 0, sizeof(*metrics));
    if (tri_count == 0) return NXB_SUCCESS;
    return build_common(d_vertices, NXB_PRIM_TRIANGLE, tri_count, cfg, nullptr, out, metrics, as_hip(stream),
                        d_indices, vertex_count);
}

nxb_status nxb_bvh2_to_host(const nxb_bvh2* device_bvh, nxb_bvh2* host_bvh) {
    if (!device_bvh || !host_bvh) return NXB_ERROR_INVALID_VALUE;
    *host_bvh = *device_bvh;
    host_bvh->nodes = nullptr;
    if (device_bvh->node_count == 0) return NXB_SUCCESS;
    if (!device_bvh->nodes) return NXB_ERROR_INVALID_VALUE;
    nxb_bvh2_node* h = static_cast<nxb_bvh2_node*>(std::malloc((size_t)device_bvh->node_count * sizeof(nxb_bvh2_node)));
    if (!h) return NXB_ERROR_OUT_OF_MEMORY;
    hipError_t e = hipMemcpy(h, device_bvh->nodes, (size_t)device_bvh->node_count * sizeof(nxb_bvh2_node),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) { std::free(h); return from_hip(e); }
    host_bvh->nodes = h;
    return NXB_SUCCESS;
}

void nxb_free_host_bvh2(nxb_bvh2* host_bvh) {
    if (!host_bvh) return;
    std::free(host_bvh->nodes);
    host_bvh->nodes = nullptr;
}

nxb_status nxb_bvh8_to_host(const nxb_bvh8* device_bvh, nxb_bvh8* host_bvh) {
    if (!device_bvh || !host_bvh) return NXB_ERROR_INVALID_VALUE;
    *host_bvh = *device_bvh;
    host_bvh->nodes = nullptr;
    host_bvh->prim_idx = nullptr;
    const size_t nb = (size_t)device_bvh->node_count * sizeof(nxb_bvh8_node);
    const size_t pb = (size_t)device_bvh->prim_count * sizeof(uint32_t);
    if ((nb && !device_bvh->nodes) || (pb && !device_bvh->prim_idx)) return NXB_ERROR_INVALID_VALUE;
    nxb_bvh8_node* h = nb ? static_cast<nxb_bvh8_node*>(std::malloc(nb)) : nullptr;
    uint32_t* q = pb ? static_cast<uint32_t*>(std::malloc(pb)) : nullptr;
    if ((nb && !h) || (pb && !q)) {
        std::free(h);
        std::free(q);
        return NXB_ERROR_OUT_OF_MEMORY;
    }
    hipError_t e = nb ? hipMemcpy(h, device_bvh->nodes, nb, hipMemcpyDeviceToHost) : hipSuccess;
    if (e == hipSuccess && pb) e = hipMemcpy(q, device_bvh->prim_idx, pb, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        std::free(h);
        std::free(q);
        return from_hip(e);
    }
    host_bvh->nodes = h;
    host_bvh->prim_idx = q;
    return NXB_SUCCESS;
}

void nxb_free_host_bvh8(nxb_bvh8* host_bvh) {
    if (!host_bvh) return;
    std::free(host_bvh->nodes);
    std::free(host_bvh->prim_idx);
    host_bvh->nodes = nullptr;
    host_bvh->prim_idx = nullptr;
}

nxb_status nxb_free_device_bvh2(nxb_bvh2* device_bvh) {
    if (!device_bvh) return NXB_ERROR_INVALID_VALUE;
    // FreeDeviceBVH is a synchronous cudaFree (BVHBuilder.cpp:285-288): wait
    // for the device, then keep the block for the next build.
    if (device_bvh->nodes) {
        NXB_HIP(hipDeviceSynchronize());
        release_out(device_bvh->nodes);
    }
    device_bvh->nodes = nullptr;
    return NXB_SUCCESS;
}

nxb_status nxb_free_device_bvh8(nxb_bvh8* device_bvh) {
    if (!device_bvh) return NXB_ERROR_INVALID_VALUE;
    if (device_bvh->nodes || device_bvh->prim_idx) NXB_HIP(hipDeviceSynchronize());
    release_out(device_bvh->nodes);
    release_out(device_bvh->prim_idx);
    device_bvh->nodes = nullptr;
    device_bvh->prim_idx = nullptr;
    return NXB_SUCCESS;
}

size_t nxb_stage_workspace_bytes(uint32_t prim_count, uint32_t morton_bits) {
    return make_layout(prim_count, morton_bits == 32 ? 4 : 8, false).total;
}

nxb_status nxb_compute_scene_bounds(const void* d_prims, nxb_prim_type type, uint32_t n, nxb_bvh2_node* d_nodes,
                                    nxb_aabb* d_scene, uint32_t* d_flags, nxb_stream_t stream) {
    if (type != NXB_PRIM_AABB && type != NXB_PRIM_TRIANGLE) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n == 0) return NXB_SUCCESS;
    if (!d_prims || !d_nodes || !d_scene || !d_flags) return NXB_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(g_mutex);
    Carved w;
    DeviceState* d = nullptr;
    nxb_status st = stage_workspace(nullptr, 0, 1, 64, w, &d);
    if (st) return st;
    hipStream_t s = as_hip(stream);
    NXB_TRY(ws_wait(*d, s));
    NXB_HIP(launch_fill_u32(w.done(), 0u, 1, s));
    BoundsScratch bs{w.partials(), w.done(), kMaxBoundsBlocks};
    NXB_HIP(launch_scene_bounds(d_prims, type, n, d_nodes, d_scene, d_flags, bs, s));
    return ws_mark(*d, s);
}

nxb_status nxb_compute_morton(const nxb_bvh2_node* d_nodes, uint32_t n, const nxb_aabb* d_scene,
                              uint32_t morton_bits, void* d_keys, uint32_t* d_values, nxb_stream_t stream) {
    if (morton_bits != 32 && morton_bits != 64) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n == 0) return NXB_SUCCESS;
    if (!d_nodes || !d_scene || !d_keys) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_morton(d_nodes, n, d_scene, morton_bits, d_keys, d_values, as_hip(stream)));
    return NXB_SUCCESS;
}

nxb_status nxb_radix_sort_pairs(const void* d_keys_in, const uint32_t* d_values_in, void* d_keys_out,
                                uint32_t* d_values_out, uint32_t n, uint32_t key_bits, uint32_t begin_bit,
                                uint32_t end_bit, uint32_t digit_bits, void* d_workspace, size_t workspace_bytes,
                                nxb_stream_t stream) {
    if (key_bits != 32 && key_bits != 64) return NXB_ERROR_INVALID_VALUE;
    if (begin_bit > end_bit || end_bit > key_bits) return NXB_ERROR_INVALID_VALUE;
    if (digit_bits == 0) digit_bits = default_digit_bits(n, key_bits);
    if (digit_bits < 6 || digit_bits > 9) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n == 0) return NXB_SUCCESS;
    if (!d_keys_in || !d_keys_out || !d_values_out) return NXB_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(g_mutex);
    Carved w;
    DeviceState* d = nullptr;
    nxb_status st = stage_workspace(d_workspace, workspace_bytes, n, key_bits, w, &d);
    if (st) return st;
    hipStream_t s = as_hip(stream);
    if (d) NXB_TRY(ws_wait(*d, s));
    NXB_HIP(launch_radix_sort(d_keys_in, d_values_in, d_keys_out, d_values_out, w.keys_tmp(), w.vals_tmp(), n,
                              key_bits, begin_bit, end_bit, digit_bits, w.sort(), s));
    return d ? ws_mark(*d, s) : NXB_SUCCESS;
}

nxb_status nxb_build_hploc(const void* d_sorted_keys, uint32_t* d_sorted_values, uint32_t n, uint32_t key_bits,
                           const nxb_build_config* cfgp, nxb_bvh2_node* d_nodes, void* d_workspace,
                           size_t workspace_bytes, nxb_stream_t stream) {
    if (key_bits != 32 && key_bits != 64) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    nxb_status st = check_config(cfgp);
    if (st) return st;
    if (n == 0) return NXB_SUCCESS;
    if (!d_sorted_keys || !d_sorted_values || !d_nodes) return NXB_ERROR_INVALID_VALUE;
    nxb_build_config cfg;
    nxb_default_config(&cfg);
    if (cfgp) cfg = *cfgp;
    std::lock_guard<std::mutex> lock(g_mutex);
    Carved w;
    DeviceState* d = nullptr;
    st = stage_workspace(d_workspace, workspace_bytes, n, key_bits, w, &d);
    if (st) return st;
    hipStream_t s = as_hip(stream);
    if (d) NXB_TRY(ws_wait(*d, s));
    NXB_HIP(launch_fill_u32(w.hdr(), 0u, kHdrHplocChunks + kHplocCounterWords, s));
    NXB_HIP(launch_fill_u32(w.parent(), kInvalid, n, s));
    HplocScratch hs{w.parent(), w.recs(), w.csizes(), w.moved(), w.moved_parent(), w.moved_size(), w.nsize(),
                    w.flags(), w.hploc_chunks()};
    NXB_HIP(launch_hploc(d_sorted_keys, key_bits, d_sorted_values, n, d_nodes, cfg.merge_mode, radius_params(cfg),
                         cfg.hploc_blocks_per_cu, hs, s));
    uint32_t flags = 0;
    NXB_HIP(hipMemcpyAsync(&flags, w.flags(), 4, hipMemcpyDeviceToHost, s));
    NXB_HIP(hipStreamSynchronize(s));
    if (d) d->ws_pending = false;
    return flags_status(flags);
}

nxb_status nxb_collapse_bvh8(const nxb_bvh2_node* d_bvh2, uint32_t n, nxb_bvh8_node* d_bvh8, uint32_t* d_prim_idx,
                             uint32_t* out_node_count, uint32_t spin_limit, void* d_workspace,
                             size_t workspace_bytes, nxb_stream_t stream) {
    if (!out_node_count) return NXB_ERROR_INVALID_VALUE;
    *out_node_count = 0;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n == 0) return NXB_SUCCESS;
    if (!d_bvh2 || !d_bvh8 || !d_prim_idx) return NXB_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(g_mutex);
    // The collapse uses no key arrays: it carves the 32-bit-key layout, so a
    // workspace sized for either key width serves.
    Carved w;
    DeviceState* d = nullptr;
    nxb_status st = stage_workspace(d_workspace, workspace_bytes, n, 32, w, &d);
    if (st) return st;
    hipStream_t s = as_hip(stream);
    if (d) NXB_TRY(ws_wait(*d, s));
    NXB_HIP(launch_fill_u32(w.hdr(), 0u, 16, s));
    // primitive counts of the internal nodes, from the BVH2 itself
    NXB_HIP(launch_subtree_sizes(d_bvh2, n, w.parent2(), w.parent(), w.nsize(), w.flags(), s));
    Bvh8Scratch bs{w.work(), w.bvh8_ctr(), w.bvh8_grab(), w.flags(), w.nsize(), w.work2(), NXB_SLOT_GREEDY};
    NXB_HIP(launch_bvh8(d_bvh2, n, d_bvh8, d_prim_idx, nxb_bvh8_capacity(n), spin_limit ? spin_limit : kDefaultSpinLimit,
                        bs, s));
    uint32_t hdr[256];
    NXB_HIP(hipMemcpyAsync(hdr, w.hdr(), sizeof(hdr), hipMemcpyDeviceToHost, s));
    NXB_HIP(hipStreamSynchronize(s));
    if (d) d->ws_pending = false;
    *out_node_count = n == 1 ? 1u : hdr[kHdrBvh8Ctr];
    return flags_status(hdr[kHdrFlags]);
}

nxb_status nxb_sah_bvh2(const nxb_bvh2* bvh, double* out_cost, nxb_stream_t stream) {
    if (!bvh || !out_cost) return NXB_ERROR_INVALID_VALUE;
    *out_cost = 0.0;
    if (bvh->node_count == 0) return NXB_SUCCESS;
    std::lock_guard<std::mutex> lock(g_mutex);
    Carved w;
    DeviceState* d = nullptr;
    nxb_status st = stage_workspace(nullptr, 0, 1, 64, w, &d);
    if (st) return st;
    hipStream_t s = as_hip(stream);
    NXB_TRY(ws_wait(*d, s));
    NXB_HIP(hipMemcpyAsync(w.scene(), &bvh->bounds, sizeof(nxb_aabb), hipMemcpyHostToDevice, s));
    NXB_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(w.sah()), 0u, 2, s));
    NXB_HIP(launch_sah_bvh2(bvh->nodes, bvh->node_count, w.scene(), w.sah(), s));
    NXB_HIP(hipMemcpyAsync(out_cost, w.sah(), sizeof(double), hipMemcpyDeviceToHost, s));
    NXB_HIP(hipStreamSynchronize(s));
    d->ws_pending = false;
    return NXB_SUCCESS;
}

nxb_status nxb_sah_bvh8(const nxb_bvh8* bvh, double* out_cost, nxb_stream_t stream) {
    if (!bvh || !out_cost) return NXB_ERROR_INVALID_VALUE;
    *out_cost = 0.0;
    if (bvh->node_count == 0) return NXB_SUCCESS;
    std::lock_guard<std::mutex> lock(g_mutex);
    Carved w;
    DeviceState* d = nullptr;
    nxb_status st = stage_workspace(nullptr, 0, 1, 64, w, &d);
    if (st) return st;
    hipStream_t s = as_hip(stream);
    NXB_TRY(ws_wait(*d, s));
    NXB_HIP(hipMemcpyAsync(w.scene(), &bvh->bounds, sizeof(nxb_aabb), hipMemcpyHostToDevice, s));
    NXB_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(w.sah()), 0u, 2, s));
    NXB_HIP(launch_sah_bvh8(bvh->nodes, bvh->node_count, w.scene(), w.sah(), s));
    NXB_HIP(hipMemcpyAsync(out_cost, w.sah(), sizeof(double), hipMemcpyDeviceToHost, s));
    NXB_HIP(hipStreamSynchronize(s));
    d->ws_pending = false;
    return NXB_SUCCESS;
}

nxb_status nxb_trace_closest(const nxb_bvh8* bvh, const float* d_tris, const nxb_ray* d_rays, uint32_t ray_count,
                             nxb_hit* d_hits, uint32_t* d_stats, nxb_stream_t stream) {
    if (!bvh || (ray_count && (!d_rays || !d_hits))) return NXB_ERROR_INVALID_VALUE;
    if (bvh->node_count && (!bvh->nodes || !bvh->prim_idx || !d_tris)) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_trace(bvh->nodes, bvh->node_count, bvh->prim_idx, bvh->prim_count, d_tris, d_rays, ray_count,
                         d_hits, d_stats, as_hip(stream)));
    return NXB_SUCCESS;
}

nxb_status nxb_trace_closest_indexed(const nxb_bvh8* bvh, const float* d_vertices, uint32_t vertex_count,
                                     const uint32_t* d_indices, const nxb_ray* d_rays, uint32_t ray_count,
                                     nxb_hit* d_hits, uint32_t* d_stats, nxb_stream_t stream) {
    if (!bvh || (ray_count && (!d_rays || !d_hits))) return NXB_ERROR_INVALID_VALUE;
    if (bvh->node_count && (!bvh->nodes || !bvh->prim_idx || !d_vertices || !d_indices))
        return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_trace(bvh->nodes, bvh->node_count, bvh->prim_idx, bvh->prim_count, d_vertices, d_rays, ray_count,
                         d_hits, d_stats, as_hip(stream), d_indices, vertex_count));
    return NXB_SUCCESS;
}

nxb_status nxb_segment_features(const void* d_sorted_keys, const uint32_t* d_sorted_values,
                                const nxb_bvh2_node* d_leaves, uint32_t n, uint32_t key_bits, uint32_t segment_log2,
                                float* d_features, nxb_stream_t stream) {
    if (key_bits != 32 && key_bits != 64) return NXB_ERROR_INVALID_VALUE;
    if (segment_log2 < 8 || segment_log2 > 20) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n == 0) return NXB_SUCCESS;
    if (!d_sorted_keys || !d_sorted_values || !d_leaves || !d_features) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_segment_features(d_sorted_keys, key_bits, d_sorted_values, d_leaves, n, segment_log2, d_features,
                                    as_hip(stream)));
    return NXB_SUCCESS;
}

nxb_status nxb_segment_sah(const nxb_bvh2_node* d_nodes, uint32_t n, const nxb_aabb* d_scene, uint32_t segment_log2,
                           double* d_seg_cost, nxb_stream_t stream) {
    if (segment_log2 < 8 || segment_log2 > 20) return NXB_ERROR_INVALID_VALUE;
    if (n > NXB_MAX_PRIMS) return NXB_ERROR_TOO_LARGE;
    if (n < 2) return NXB_SUCCESS;
    if (!d_nodes || !d_scene || !d_seg_cost) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_segment_sah(d_nodes, n, segment_log2, d_scene, d_seg_cost, as_hip(stream)));
    return NXB_SUCCESS;
}

nxb_status nxb_mlp_radius_table(const float* d_params, const float* d_features, uint32_t count, float min_confidence,
                                uint8_t* d_table, float* d_logits, uint32_t* d_fallbacks, nxb_stream_t stream) {
    if (count == 0) return NXB_SUCCESS;
    if (!d_params || !d_features || !d_table) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_mlp_table(d_params, d_features, count, min_confidence, d_table, d_logits, d_fallbacks,
                             as_hip(stream)));
    return NXB_SUCCESS;
}

nxb_status nxb_mlp_gradients(const float* d_params, const float* d_features, const uint8_t* d_labels, uint32_t count,
                             float* d_grads, float* d_loss, nxb_stream_t stream) {
    if (count == 0) return NXB_SUCCESS;
    if (!d_params || !d_features || !d_labels || !d_grads || !d_loss) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_mlp_grad(d_params, d_features, d_labels, count, d_grads, d_loss, as_hip(stream)));
    return NXB_SUCCESS;
}

nxb_status nxb_mlp_adam_step(float* d_params, const float* d_grads, float* d_m, float* d_v, uint32_t step, float lr,
                             float beta1, float beta2, float eps, nxb_stream_t stream) {
    if (!d_params || !d_grads || !d_m || !d_v || step == 0) return NXB_ERROR_INVALID_VALUE;
    NXB_HIP(launch_adam(d_params, d_grads, d_m, d_v, step, lr, beta1, beta2, eps, as_hip(stream)));
    return NXB_SUCCESS;
}

void nxb_last_learned_stats(uint32_t* segments, uint32_t* fallbacks) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (segments) *segments = g_learned_segments;
    if (fallbacks) *fallbacks = g_learned_fallbacks;
}

}  // extern "C"
