// libtdp.so: the device layer of the join path behind include/tdp.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../../include/tdp.h"
#include "tdp_kernels.hpp"
#include "tdp_mixed.hpp"
#include "tdp_cand.hpp"

using namespace tdp;

struct tdp_ctx {
    int device = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    tdp_tuning tune{(int)tdp::kMxSmallPairs, 64, 0, 0, 4096, 0, 1};
    std::string err;
    // grow-only scratch for scans
    std::vector<void *> scan_buf;
    std::vector<size_t> scan_cap;
    uint64_t *d_total = nullptr;
    // block cache for join buffers: released blocks are kept and reused by
    // later joins (no hipMalloc / hipFree per join); freed at tdp_close, or
    // when an allocation would otherwise run out of memory
    std::multimap<size_t, void *> pool_free;
    std::unordered_map<void *, size_t> pool_size;
    // every device allocation: its size, the bytes held, the cap on them and
    // the most held (a tile may be freed from another thread: mem_mu)
    std::mutex mem_mu;
    std::unordered_map<void *, size_t> dsize;
    uint64_t mem_limit = 0, mem_used = 0, mem_peak = 0;
    // timing events of finished joins, reused by later ones
    std::vector<hipEvent_t> ev_pool;
    // pinned host blocks given back through tdp_host_free, kept for later
    // tdp_host_alloc calls (pinning costs milliseconds per call), at most
    // kHostCache bytes; freed at tdp_close (under mem_mu)
    static constexpr size_t kHostCache = size_t(1) << 30;
    std::multimap<size_t, void *> host_free;
    std::unordered_map<void *, size_t> host_size;
    size_t host_cached = 0;
    // points in the stream's queue the host can wait for (tdp_mark)
    static constexpr int kMarks = 16;
    hipEvent_t marks[kMarks] = {};
    uint64_t n_marks = 0;
};

struct tdp_tile {
    tdp_ctx *ctx = nullptr;
    size_t n_obj = 0, n_vox = 0, n_tri = 0;
    uint32_t *obj_vox_off = nullptr;
    float *obj_hd = nullptr;
    float *vox_box = nullptr;
    uint32_t *vox_obj = nullptr;
    uint64_t *vox_tri_off = nullptr;
    uint32_t *vox_tri_cnt = nullptr;
    float4 *tri = nullptr;
    float4 *bnd = nullptr;  // kBnd per triangle: sphere, plane, hausdorff pair
    // kSph per voxel and LOD slot: the voxel's sphere (k_vox_sph), read by
    // the mixed mark pass. Formed at upload, by the record puts of tiles
    // filled without records, and before the first join after a fill with
    // records (sph_stale)
    float4 *sph = nullptr;
    mutable bool sph_stale = false;
    float *obj_box = nullptr, *obj_tr = nullptr;  // for candidate queries (optional)
    size_t bytes = 0;
    size_t cap_obj = 0, cap_vox = 0, cap_tri = 0;  // reserved tiles: capacity
    // filled without records (tdp_tile_fill_meta_async): have[v] bit s = the
    // records of voxel v at LOD slot s are held; the rest is the scratch of
    // one round's list (tdp_join_lazy_needs / tdp_join_lazy_put)
    bool lazy = false;
    uint32_t *have = nullptr, *lz_flag = nullptr, *lz_list = nullptr, *lz_cnt = nullptr;
    uint64_t *lz_pos = nullptr, *lz_off = nullptr;
    size_t lz_cap = 0;
    // listed by tdp_join_lazy_needs and not yet put: the tile is unusable
    // until they are (or until it is filled again)
    size_t lz_pending = 0;
    uint64_t lz_records = 0;  // their records
    int lz_slot = -1;
    uint64_t put_records = 0;  // records put since the last fill
};

struct tdp_join {
    tdp_ctx *ctx = nullptr;
    const tdp_tile *t1 = nullptr, *t2 = nullptr;
    tdp_join_params prm{};
    size_t N = 0, P = 0;
    size_t rsz = 4;  // sizeof range value (4 REF32, 8 otherwise)
    uint32_t *prefix = nullptr, *pair_obj2 = nullptr, *pair_target = nullptr;
    int32_t *k_t = nullptr;  // per target: k (kNN)
    float *tau_t = nullptr;  // per target: within distance
    uint8_t *within_t = nullptr;  // kNN and within targets in one join: 1 for within
    void *rmin = nullptr, *rmax = nullptr;
    uint8_t *state = nullptr, *round = nullptr, *scratch = nullptr;
    int32_t *confirmed = nullptr;
    void *acc_min = nullptr, *acc_max = nullptr;
    uint32_t *count = nullptr;
    uint64_t *offs = nullptr;
    // voxel-pair lists, double-buffered for compaction
    uint32_t *v1[2] = {nullptr, nullptr}, *v2[2] = {nullptr, nullptr}, *vpp[2] = {nullptr, nullptr};
    void *vmin[2] = {nullptr, nullptr}, *vmax[2] = {nullptr, nullptr};
    uint32_t *keep = nullptr;
    // mixed-precision pipeline state
    float *mx_U = nullptr;
    uint32_t *mx_probe = nullptr, *mx_flags = nullptr;
    unsigned long long *mx_key = nullptr;
    VpPlan *mx_plan = nullptr;  // per voxel pair: triangles and tiling (k_mx_plan)
    unsigned long long *mx_pwb = nullptr;  // per voxel pair: its probes' mask word and bit
    uint32_t *mx_tested = nullptr;  // per voxel pair: pairs the mark pass tested one by one
    MarkPrep *mx_prep = nullptr;     // per voxel pair: the mark pass's thresholds (k_mx_prep)
    uint32_t *mx_nw = nullptr;   // candidate-mask words per voxel pair
    uint64_t *mx_woff = nullptr; // their exclusive scan
    size_t cap_mx = 0;
    unsigned long long *mx_mask = nullptr;
    size_t cap_mask = 0;
    uint32_t *item_vp = nullptr, *item_q = nullptr;
    size_t cap_items = 0;
    uint64_t *pos = nullptr;
    size_t cap_vp = 0;
    int cur = 0;
    uint64_t n_vp = 0;
    int rounds = 0;
    bool filtered = false;
    unsigned long long *stats = nullptr;
    uint64_t filter_vp = 0;
    tdp_tuning tune{};
    uint32_t *work = nullptr;  // work counter of the persistent stage
    uint32_t *obj_u = nullptr;  // MIXED: per object pair, the last round's cap on U_d (keys)
    // graph-captured rounds: device totals, pinned read-back, the executable
    unsigned long long *words5 = nullptr;  // per-slot mask-word bound
    uint64_t *d_words = nullptr, *d_kept = nullptr;
    uint64_t *h_kept = nullptr;             // pinned
    hipGraphExec_t gexec = nullptr;
    bool planned = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double refine_ms = 0, filter_ms = 0;
    // per-stage timing: events recorded between the stages of a round
    hipEvent_t sev[12] = {};
    int sev_stage[12] = {};
    int n_sev = 0;
    double stage_ms[TDP_STAGE_COUNT] = {};
    uint64_t stage_n[TDP_STAGE_COUNT] = {};
    // tiles filled without records: per side, the LOD whose records were
    // put (refine checks it), and the pinned copy of the last list
    int lz_ready[2] = {-1, -1};
    uint32_t *lz_host[2] = {nullptr, nullptr};
    size_t lz_host_cap[2] = {0, 0};
    float4 *lz_stage = nullptr;  // device copy of the records being put
    size_t lz_stage_cap = 0;     // records
};

// ------------------------------------------------------------ helpers

#define TDP_CHECK(ctx, call)                                                         \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return e_ == hipErrorOutOfMemory ? TDP_ERR_NOMEM : TDP_ERR_HIP;          \
        }                                                                            \
    } while (0)

#define TDP_LAUNCHED(ctx)                                                            \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess) {                                                      \
            (ctx)->err = std::string("kernel launch: ") + hipGetErrorString(e_);     \
            return TDP_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

// Synchronous copies and fills on the context's own stream (never the legacy
// stream: another thread's context may be capturing a graph meanwhile, and
// legacy-stream work would have to wait for a capturing stream).
static hipError_t copy_sync(tdp_ctx *ctx, void *dst, const void *src, size_t bytes,
                            hipMemcpyKind kind) {
    if (!bytes) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e;
}
static hipError_t set_sync(tdp_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!bytes) return hipSuccess;
    hipError_t e = hipMemsetAsync(dst, value, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e;
}

static inline unsigned blocks_for(size_t n, size_t per_block) {
    return (unsigned)((n + per_block - 1) / per_block);
}

// grid for one-wave-per-item kernels (wave_global): 1-D up to 2^20 blocks,
// 2-D beyond, so any 32-bit item count fits
static inline dim3 wave_grid(size_t n) {
    const size_t b = (n + kWavesPerBlock - 1) / kWavesPerBlock;
    const size_t x_max = (size_t)1 << 20;
    if (b <= x_max) return dim3((unsigned)(b ? b : 1), 1, 1);
    return dim3((unsigned)x_max, (unsigned)((b + x_max - 1) / x_max), 1);
}

// grid for grid-stride kernels
static inline unsigned stride_grid(size_t n, int cap = 4096) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)std::min<size_t>(b ? b : 1, (size_t)std::max(cap, 1));
}

static void pool_release_all(tdp_ctx *ctx);

// Held while a round is captured into a graph and around every device or
// pinned allocation and free: an allocation call may synchronise the
// device, which another thread's capture does not allow.
static std::recursive_mutex g_capture_mu;

// Every device allocation of the library: the context's cap on the bytes it
// holds, its high-water mark, and one release of the block cache before an
// allocation is refused.
static int dev_alloc(tdp_ctx *ctx, void **p, size_t bytes) {
    *p = nullptr;
    for (int attempt = 0; attempt < 2; attempt++) {
        bool over;
        {
            std::lock_guard<std::mutex> g(ctx->mem_mu);
            over = ctx->mem_limit && ctx->mem_used + bytes > ctx->mem_limit;
        }
        hipError_t e = hipSuccess;
        if (!over) {
            {
                std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
                e = hipMalloc(p, bytes);
            }
            if (e == hipSuccess) {
                std::lock_guard<std::mutex> g(ctx->mem_mu);
                ctx->dsize[*p] = bytes;
                ctx->mem_used += bytes;
                ctx->mem_peak = std::max(ctx->mem_peak, ctx->mem_used);
                return TDP_OK;
            }
            (void)hipGetLastError();
        }
        if (attempt == 0 && !ctx->pool_free.empty() && (over || e == hipErrorOutOfMemory)) {
            pool_release_all(ctx);
            continue;
        }
        if (over)
            ctx->err = "device memory cap: " + std::to_string(bytes) + " more bytes over " +
                       std::to_string(ctx->mem_limit) + " (" + std::to_string(ctx->mem_used) +
                       " held)";
        else
            ctx->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
        return over || e == hipErrorOutOfMemory ? TDP_ERR_NOMEM : TDP_ERR_HIP;
    }
    return TDP_ERR_NOMEM;
}

static void dev_free(tdp_ctx *ctx, void *p) {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
    (void)hipFree(p);
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    auto it = ctx->dsize.find(p);
    if (it != ctx->dsize.end()) {
        ctx->mem_used -= it->second;
        ctx->dsize.erase(it);
    }
}

template <typename T>
static int dalloc(tdp_ctx *ctx, T **p, size_t count) {
    if (count == 0) count = 1;
    return dev_alloc(ctx, (void **)p, count * sizeof(T));
}
static void dfree(tdp_ctx *ctx, void *p) { dev_free(ctx, p); }

static size_t size_class(size_t b) {
    if (b <= ((size_t)1 << 16)) return (size_t)1 << 16;
    size_t p = 1;
    while (p < b) p <<= 1;
    const size_t step = p / 8;
    return (b + step - 1) / step * step;
}

static void pool_release_all(tdp_ctx *ctx) {
    for (auto &kv : ctx->pool_free) {
        dev_free(ctx, kv.second);
        ctx->pool_size.erase(kv.second);
    }
    ctx->pool_free.clear();
}

// join buffers come from the context's block cache (callers release them only
// when no queued kernel uses them: after a stream synchronisation)
// zero count float4 records at p (tile padding), synchronously
static int zero_pad(tdp_ctx *ctx, float4 *p, size_t count) {
    if (!count) return TDP_OK;
    TDP_CHECK(ctx, hipMemsetAsync(p, 0, count * sizeof(float4), ctx->stream));
    TDP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TDP_OK;
}

template <typename T>
static int palloc(tdp_ctx *ctx, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    const size_t bytes = size_class(count * sizeof(T));
    auto it = ctx->pool_free.lower_bound(bytes);
    if (it != ctx->pool_free.end() && it->first <= bytes + bytes / 2) {
        *p = (T *)it->second;
        ctx->pool_free.erase(it);
        return TDP_OK;
    }
    void *q = nullptr;
    const int rc = dev_alloc(ctx, &q, bytes);
    if (rc != TDP_OK) return rc;
    ctx->pool_size[q] = bytes;
    *p = (T *)q;
    return TDP_OK;
}

static void pfree(tdp_ctx *ctx, void *p) {
    if (!p) return;
    auto it = ctx->pool_size.find(p);
    if (it == ctx->pool_size.end()) {
        dev_free(ctx, p);
        return;
    }
    ctx->pool_free.emplace(it->second, p);
}

// block-sum buffers of a scan of n elements (grow-only; allocating here, not
// during a graph capture)
static int scan_reserve(tdp_ctx *ctx, size_t n, std::vector<size_t> *levels = nullptr) {
    std::vector<size_t> sizes;
    size_t m = n ? n : 1;
    while (true) {
        size_t b = (m + 1023) / 1024;
        sizes.push_back(b);
        if (b == 1) break;
        m = b;
    }
    if (ctx->scan_buf.size() < 2 * sizes.size()) {
        ctx->scan_buf.resize(2 * sizes.size(), nullptr);
        ctx->scan_cap.resize(2 * sizes.size(), 0);
    }
    for (size_t l = 0; l < sizes.size(); l++)
        for (int k = 0; k < 2; k++) {
            size_t idx = 2 * l + k;
            if (ctx->scan_cap[idx] < sizes[l]) {
                dfree(ctx, ctx->scan_buf[idx]);
                ctx->scan_buf[idx] = nullptr;
                ctx->scan_cap[idx] = 0;
                const int rc = dev_alloc(ctx, &ctx->scan_buf[idx], sizes[l] * sizeof(uint64_t));
                if (rc != TDP_OK) return rc;
                ctx->scan_cap[idx] = sizes[l];
            }
        }
    if (!ctx->d_total) {
        const int rc = dalloc(ctx, &ctx->d_total, 1);
        if (rc != TDP_OK) return rc;
    }
    if (levels) *levels = sizes;
    return TDP_OK;
}

// exclusive scan of n counters (u32 or u64) into out (u64) and the total into
// *total_dev, kernels only (no host synchronisation: usable in a graph
// capture once scan_reserve has run for n)
template <typename In>
static int scan_launch(tdp_ctx *ctx, const In *in, size_t n, uint64_t *out, uint64_t *total_dev) {
    std::vector<size_t> sizes;
    int rc = scan_reserve(ctx, n, &sizes);
    if (rc != TDP_OK) return rc;
    if (n == 0) {
        k_scan_total<In><<<1, 1, 0, ctx->stream>>>(in, out, 0, total_dev);
        TDP_LAUNCHED(ctx);
        return TDP_OK;
    }
    uint64_t *sums0 = (uint64_t *)ctx->scan_buf[0];
    k_scan_local<In><<<(unsigned)sizes[0], kBlock, 0, ctx->stream>>>(in, n, out, sums0);
    TDP_LAUNCHED(ctx);
    // upper levels: scan the block sums in place into offsets
    for (size_t l = 0; l + 1 < sizes.size(); l++) {
        uint64_t *sums = (uint64_t *)ctx->scan_buf[2 * l];
        uint64_t *offs = (uint64_t *)ctx->scan_buf[2 * l + 1];
        uint64_t *nsums = (uint64_t *)ctx->scan_buf[2 * (l + 1)];
        k_scan_local<uint64_t><<<(unsigned)sizes[l + 1], kBlock, 0, ctx->stream>>>(sums, sizes[l],
                                                                                   offs, nsums);
        TDP_LAUNCHED(ctx);
    }
    // top level has one block: its offsets are the scan itself; propagate down
    for (size_t l = sizes.size() - 1; l-- > 0;) {
        uint64_t *offs = (uint64_t *)ctx->scan_buf[2 * l + 1];
        uint64_t *upper = (uint64_t *)ctx->scan_buf[2 * (l + 1) + 1];
        if (l + 1 < sizes.size() - 1)
            k_scan_add<<<blocks_for(sizes[l], kBlock), kBlock, 0, ctx->stream>>>(offs, sizes[l], upper);
        TDP_LAUNCHED(ctx);
    }
    if (sizes.size() > 1) {
        uint64_t *offs0 = (uint64_t *)ctx->scan_buf[1];
        k_scan_add<<<blocks_for(n, kBlock), kBlock, 0, ctx->stream>>>(out, n, offs0);
        TDP_LAUNCHED(ctx);
    }
    k_scan_total<In><<<1, 1, 0, ctx->stream>>>(in, out, n, total_dev);
    TDP_LAUNCHED(ctx);
    return TDP_OK;
}

// exclusive scan of n counters (u32 or u64) into out (u64); *total set on host
template <typename In>
static int scan_exclusive(tdp_ctx *ctx, const In *in, size_t n, uint64_t *out, uint64_t *total) {
    if (n == 0) {
        *total = 0;
        return TDP_OK;
    }
    int rc = scan_reserve(ctx, n);
    if (rc != TDP_OK) return rc;
    rc = scan_launch<In>(ctx, in, n, out, ctx->d_total);
    if (rc != TDP_OK) return rc;
    TDP_CHECK(ctx, hipMemcpyAsync(total, ctx->d_total, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
    TDP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TDP_OK;
}

// ------------------------------------------------------------ allocation-free batch

// Scratch of tdp_meshdist_batch_noalloc: the triangle bound records of up to
// kNaMaxTri triangles (kBnd float4 each), then its counters (the mixed form's
// statistics and the argument-check flag). One array per GPU, held from the
// library's load to the process's end. Host code never uses the name itself
// (on the host it is a stand-in, not the device address): tdp_open looks up
// each device's address once, with that device current.
constexpr size_t kNaMaxTri = (size_t)1 << 18;
constexpr size_t kNaBndBytes = kNaMaxTri * kBnd * sizeof(float4);  // 12 MiB
constexpr size_t kNaScratchBytes = kNaBndBytes + 256;
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu[kNaMaxDevices];  // one call at a time per device in a process
static void *g_na_ptr[kNaMaxDevices] = {};
static std::once_flag g_na_once[kNaMaxDevices];
static hipError_t g_na_err[kNaMaxDevices] = {};

// the device address of `scratch` on the current device (once per device)
static hipError_t na_scratch_setup(int device) {
    std::call_once(g_na_once[device], [device] {
        void *p = nullptr;
        g_na_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_na_err[device] == hipSuccess) g_na_ptr[device] = p;
    });
    return g_na_err[device];
}

// ------------------------------------------------------------ context

extern "C" int tdp_open(int device, tdp_ctx **out) {
    if (!out) return TDP_ERR_ARG;
    *out = nullptr;
    tdp_ctx *ctx = new tdp_ctx();
    ctx->device = device;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= device || device < 0) {
        ctx->err = std::string("no HIP device ") + std::to_string(device) +
                   (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : std::string());
        *out = ctx;  // returned so the caller can read the message
        return TDP_ERR_HIP;
    }
    e = hipSetDevice(device);
    // non-blocking: no implicit order with the legacy stream, which the
    // library never uses
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        ctx->err = std::string("hip init: ") + hipGetErrorString(e);
        *out = ctx;
        return TDP_ERR_HIP;
    }
    ctx->cu_count = prop.multiProcessorCount;
    if (device < kNaMaxDevices) {
        e = na_scratch_setup(device);
        if (e != hipSuccess) {
            ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            *out = ctx;
            return TDP_ERR_HIP;
        }
    }
    *out = ctx;
    return TDP_OK;
}

extern "C" int tdp_get_tuning(tdp_ctx *ctx, tdp_tuning *out) {
    if (!ctx || !out) return TDP_ERR_ARG;
    *out = ctx->tune;
    return TDP_OK;
}

extern "C" int tdp_set_tuning(tdp_ctx *ctx, const tdp_tuning *t) {
    if (!ctx || !t) return TDP_ERR_ARG;
    if (t->small_pairs < 0 || t->max_rows < 1 || t->max_rows > kWave || t->drain_blocks < 0 ||
        t->eval_blocks < 1) {
        ctx->err = "tdp_set_tuning: value out of range";
        return TDP_ERR_ARG;
    }
    ctx->tune = *t;
    return TDP_OK;
}

extern "C" int tdp_close(tdp_ctx *ctx) {
    if (!ctx) return TDP_ERR_ARG;
    std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
    for (auto &e : ctx->marks)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto &kv : ctx->host_free) (void)hipHostFree(kv.second);
    for (void *p : ctx->scan_buf) dfree(ctx, p);
    dfree(ctx, ctx->d_total);
    pool_release_all(ctx);
    for (auto &kv : ctx->pool_size) dev_free(ctx, kv.first);  // blocks still held by joins
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TDP_OK;
}

extern "C" const char *tdp_last_error(const tdp_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int tdp_device_memory(tdp_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!ctx) return TDP_ERR_ARG;
    TDP_CHECK(ctx, hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    TDP_CHECK(ctx, hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return TDP_OK;
}

extern "C" int tdp_device_info(tdp_ctx *ctx, char *name, size_t name_len, int *cu_count) {
    if (!ctx) return TDP_ERR_ARG;
    hipDeviceProp_t prop;
    TDP_CHECK(ctx, hipGetDeviceProperties(&prop, ctx->device));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    return TDP_OK;
}

// ------------------------------------------------------------ tiles

// What the kernels assume of a tile's voxels, checked before any upload:
// voxel ranges that do not decrease, every voxel's records inside the record
// array, and at most kMaxVoxTris triangles per voxel and LOD slot (the pair
// walks number a voxel pair's n1 x n2 triangle pairs in 32 bits). Returns
// nullptr or what is wrong.
static const char *check_voxels(const tdp_tile_desc *d) {
    for (size_t o = 0; o < d->n_obj; o++)
        if (d->obj_vox_off[o + 1] < d->obj_vox_off[o]) return "obj_vox_off decreases";
    for (size_t v = 0; v < 5 * d->n_vox; v++) {
        if (d->vox_tri_cnt[v] > kMaxVoxTris)
            return "a voxel holds more than 65535 triangles at one LOD";
        if (d->vox_tri_off[v] > d->n_tri || d->vox_tri_cnt[v] > d->n_tri - d->vox_tri_off[v])
            return "voxel triangles outside the record array";
    }
    return nullptr;
}

// tri_parts: the records in n_parts host buffers (concatenated in order), or
// nullptr for d->tri
static int tile_upload(tdp_ctx *ctx, const tdp_tile_desc *d, const float *const *tri_parts,
                       const size_t *part_n_tri, size_t n_parts, tdp_tile **out) {
    if (!ctx || !d || !out) return TDP_ERR_ARG;
    *out = nullptr;
    if (tri_parts) {
        size_t sum = 0;
        for (size_t k = 0; k < n_parts; k++) {
            if (part_n_tri[k] && !tri_parts[k]) {
                ctx->err = "tdp_tile_upload_parts: null part";
                return TDP_ERR_ARG;
            }
            sum += part_n_tri[k];
        }
        if (sum != d->n_tri) {
            ctx->err = "tdp_tile_upload_parts: parts do not add up to n_tri";
            return TDP_ERR_ARG;
        }
    }
    // arrays may be null where their count is zero (a tile without objects
    // is a valid tile: its joins have no candidates)
    if (!d->obj_vox_off || (d->n_vox && (!d->vox_box || !d->vox_tri_off || !d->vox_tri_cnt)) ||
        (d->n_tri && !d->tri && !tri_parts) || (d->n_obj && !d->obj_hd)) {
        ctx->err = "tdp_tile_upload: null array";
        return TDP_ERR_ARG;
    }
    if (d->obj_vox_off[0] != 0 || d->obj_vox_off[d->n_obj] != d->n_vox) {
        ctx->err = "tdp_tile_upload: obj_vox_off does not span the voxels";
        return TDP_ERR_ARG;
    }
    if (const char *bad = check_voxels(d))
        return ctx->err = std::string("tdp_tile_upload: ") + bad, TDP_ERR_ARG;
    tdp_tile *t = new tdp_tile();
    t->ctx = ctx;
    t->n_obj = d->n_obj;
    t->n_vox = d->n_vox;
    t->n_tri = d->n_tri;
    std::vector<uint32_t> vobj(d->n_vox);
    for (size_t o = 0; o < d->n_obj; o++)
        for (uint32_t v = d->obj_vox_off[o]; v < d->obj_vox_off[o + 1]; v++) vobj[v] = (uint32_t)o;
    int rc = TDP_OK;
    auto up = [&](auto **dst, const auto *src, size_t count) {
        if (rc != TDP_OK) return;
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
        rc = dalloc(ctx, dst, count);
        if (rc != TDP_OK) return;
        t->bytes += count * sizeof(T);
        if (count) {
            hipError_t e = copy_sync(ctx, *dst, src, count * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                ctx->err = std::string("tile upload: ") + hipGetErrorString(e);
                rc = TDP_ERR_HIP;
            }
        }
    };
    up(&t->obj_vox_off, d->obj_vox_off, d->n_obj + 1);
    up(&t->obj_hd, d->obj_hd, 10 * d->n_obj);
    up(&t->vox_box, d->vox_box, 9 * d->n_vox);
    up(&t->vox_obj, vobj.data(), d->n_vox);
    up(&t->vox_tri_off, d->vox_tri_off, 5 * d->n_vox);
    up(&t->vox_tri_cnt, d->vox_tri_cnt, 5 * d->n_vox);
    if ((d->obj_box && d->obj_transform) || d->n_obj == 0) {
        up(&t->obj_box, d->obj_box, 6 * d->n_obj);
        up(&t->obj_tr, d->obj_transform, 3 * d->n_obj);
    }
    // records and bounds carry kTriPad zeroed triangles past the last one
    if (rc == TDP_OK) rc = dalloc(ctx, &t->tri, 3 * (d->n_tri + kTriPad));
    if (rc == TDP_OK) {
        t->bytes += 48 * (d->n_tri + kTriPad);
        rc = zero_pad(ctx, t->tri + 3 * d->n_tri, 3 * kTriPad);
    }
    if (!tri_parts) {
        if (rc == TDP_OK && d->n_tri) {
            hipError_t e = copy_sync(ctx, t->tri, d->tri, 48 * d->n_tri, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                ctx->err = std::string("tile upload: ") + hipGetErrorString(e);
                rc = TDP_ERR_HIP;
            }
        }
    } else if (rc == TDP_OK) {
        size_t at = 0;
        for (size_t k = 0; rc == TDP_OK && k < n_parts; k++) {
            if (part_n_tri[k]) {
                hipError_t e = copy_sync(ctx, t->tri + 3 * at, tri_parts[k], 48 * part_n_tri[k],
                                         hipMemcpyHostToDevice);
                if (e != hipSuccess) {
                    ctx->err = std::string("tile upload: ") + hipGetErrorString(e);
                    rc = TDP_ERR_HIP;
                }
            }
            at += part_n_tri[k];
        }
    }
    if (rc == TDP_OK) rc = dalloc(ctx, &t->bnd, kBnd * (d->n_tri + kTriPad));
    if (rc == TDP_OK) {
        t->bytes += sizeof(float4) * kBnd * (d->n_tri + kTriPad);
        rc = zero_pad(ctx, t->bnd + kBnd * d->n_tri, kBnd * kTriPad);
    }
    if (rc == TDP_OK) rc = dalloc(ctx, &t->sph, kSph * 5 * std::max<size_t>(d->n_vox, 1));
    if (rc == TDP_OK) t->bytes += sizeof(float4) * kSph * 5 * d->n_vox;
    if (rc == TDP_OK && d->n_tri) {
        k_tri_bounds<<<blocks_for(d->n_tri, kBlock), kBlock, 0, ctx->stream>>>(d->n_tri, t->tri, t->bnd);
        k_vox_sph<<<stride_grid(5 * d->n_vox * kWave, 4096), kBlock, 0, ctx->stream>>>(
            d->n_vox, t->vox_tri_off, t->vox_tri_cnt, t->tri, t->bnd, t->sph);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            ctx->err = std::string("tile bounds: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    }
    if (rc != TDP_OK) {
        tdp_tile_free(t);
        return rc;
    }
    *out = t;
    return TDP_OK;
}

extern "C" int tdp_tile_upload(tdp_ctx *ctx, const tdp_tile_desc *d, tdp_tile **out) {
    return tile_upload(ctx, d, nullptr, nullptr, 0, out);
}

extern "C" int tdp_tile_upload_parts(tdp_ctx *ctx, const tdp_tile_desc *d,
                                     const float *const *tri_parts, const size_t *part_n_tri,
                                     size_t n_parts, tdp_tile **out) {
    if (!tri_parts || !part_n_tri) return TDP_ERR_ARG;
    return tile_upload(ctx, d, tri_parts, part_n_tri, n_parts, out);
}

// ------------------------------------------------------------ streaming fill

extern "C" int tdp_set_memory_limit(tdp_ctx *ctx, uint64_t bytes) {
    if (!ctx) return TDP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    ctx->mem_limit = bytes;
    return TDP_OK;
}

extern "C" int tdp_memory_usage(tdp_ctx *ctx, uint64_t *used, uint64_t *peak, int reset_peak) {
    if (!ctx) return TDP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mem_mu);
    if (used) *used = ctx->mem_used;
    if (peak) *peak = ctx->mem_peak;
    if (reset_peak) ctx->mem_peak = ctx->mem_used;
    return TDP_OK;
}

extern "C" int tdp_host_alloc(tdp_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return TDP_ERR_ARG;
    *out = nullptr;
    bytes = bytes ? bytes : 1;
    {
        // a cached block of at least the size and at most twice it
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        auto it = ctx->host_free.lower_bound(bytes);
        if (it != ctx->host_free.end() && it->first / 2 <= bytes) {
            *out = it->second;
            ctx->host_cached -= it->first;
            ctx->host_free.erase(it);
            return TDP_OK;
        }
    }
    void *p = nullptr;
    {
        std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
        TDP_CHECK(ctx, hipHostMalloc(&p, bytes, hipHostMallocDefault));
    }
    std::lock_guard<std::mutex> lk(ctx->mem_mu);
    ctx->host_size[p] = bytes;
    *out = p;
    return TDP_OK;
}

extern "C" int tdp_host_free(tdp_ctx *ctx, void *p) {
    if (!ctx) return TDP_ERR_ARG;
    if (!p) return TDP_OK;
    {
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        auto it = ctx->host_size.find(p);
        if (it == ctx->host_size.end()) {
            ctx->err = "tdp_host_free: not a block of this context";
            return TDP_ERR_ARG;
        }
        if (ctx->host_cached + it->second <= tdp_ctx::kHostCache) {
            ctx->host_free.emplace(it->second, p);
            ctx->host_cached += it->second;
            return TDP_OK;
        }
        ctx->host_size.erase(it);
    }
    std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
    TDP_CHECK(ctx, hipHostFree(p));
    return TDP_OK;
}

extern "C" int tdp_sync(tdp_ctx *ctx) {
    if (!ctx) return TDP_ERR_ARG;
    TDP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TDP_OK;
}

extern "C" int tdp_mark(tdp_ctx *ctx, uint64_t *ticket) {
    if (!ctx || !ticket) return TDP_ERR_ARG;
    hipEvent_t &e = ctx->marks[ctx->n_marks % tdp_ctx::kMarks];
    if (!e) TDP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    TDP_CHECK(ctx, hipEventRecord(e, ctx->stream));
    *ticket = ctx->n_marks++;
    return TDP_OK;
}

extern "C" int tdp_wait_mark(tdp_ctx *ctx, uint64_t ticket) {
    if (!ctx || ticket >= ctx->n_marks) return TDP_ERR_ARG;
    // a slot recorded again since marks a later point: waiting for it covers
    // this ticket too
    TDP_CHECK(ctx, hipEventSynchronize(ctx->marks[ticket % tdp_ctx::kMarks]));
    return TDP_OK;
}

extern "C" int tdp_tile_reserve(tdp_ctx *ctx, size_t n_obj, size_t n_vox, size_t n_tri,
                                tdp_tile **out) {
    if (!ctx || !out || !n_obj || !n_vox || !n_tri) return TDP_ERR_ARG;
    *out = nullptr;
    tdp_tile *t = new tdp_tile();
    t->ctx = ctx;
    t->cap_obj = n_obj;
    t->cap_vox = n_vox;
    t->cap_tri = n_tri;
    int rc = dalloc(ctx, &t->obj_vox_off, n_obj + 1);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->obj_hd, 10 * n_obj);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->vox_box, 9 * n_vox);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->vox_obj, n_vox);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->vox_tri_off, 5 * n_vox);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->vox_tri_cnt, 5 * n_vox);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->tri, 3 * (n_tri + kTriPad));
    if (rc == TDP_OK) rc = dalloc(ctx, &t->bnd, kBnd * (n_tri + kTriPad));
    if (rc == TDP_OK) rc = dalloc(ctx, &t->sph, kSph * 5 * n_vox);
    // zeroed whole: the padding past the capacity, and records a smaller fill
    // leaves untouched, are always readable values
    if (rc == TDP_OK) rc = zero_pad(ctx, t->tri, 3 * (n_tri + kTriPad));
    if (rc == TDP_OK) rc = zero_pad(ctx, t->bnd, kBnd * (n_tri + kTriPad));
    if (rc == TDP_OK) rc = dalloc(ctx, &t->obj_box, 6 * n_obj);
    if (rc == TDP_OK) rc = dalloc(ctx, &t->obj_tr, 3 * n_obj);
    if (rc != TDP_OK) {
        tdp_tile_free(t);
        return rc;
    }
    t->bytes = (n_obj + 1) * 4 + n_obj * 76 + n_vox * (36 + 4 + 40 + 20 + 16 * kSph * 5) +
               (n_tri + kTriPad) * 16 * (3 + kBnd);
    *out = t;
    return TDP_OK;
}

extern "C" int tdp_tile_fill_async(tdp_tile *t, const tdp_tile_desc *d) {
    if (!t || !d) return TDP_ERR_ARG;
    tdp_ctx *ctx = t->ctx;
    if (!t->cap_obj) {
        ctx->err = "tdp_tile_fill_async: the tile was not made by tdp_tile_reserve";
        return TDP_ERR_STATE;
    }
    if (d->n_obj > t->cap_obj || d->n_vox > t->cap_vox || d->n_tri > t->cap_tri) {
        ctx->err = "tdp_tile_fill_async: sizes beyond the reserved capacity";
        return TDP_ERR_ARG;
    }
    if (!d->obj_vox_off || (d->n_vox && (!d->vox_box || !d->vox_tri_off || !d->vox_tri_cnt)) ||
        (d->n_obj && !d->obj_hd)) {
        ctx->err = "tdp_tile_fill_async: null array";
        return TDP_ERR_ARG;
    }
    if (d->obj_vox_off[0] != 0 || d->obj_vox_off[d->n_obj] != d->n_vox) {
        ctx->err = "tdp_tile_fill_async: obj_vox_off does not span the voxels";
        return TDP_ERR_ARG;
    }
    if (const char *bad = check_voxels(d))
        return ctx->err = std::string("tdp_tile_fill_async: ") + bad, TDP_ERR_ARG;
    hipStream_t s = ctx->stream;
    auto cp = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    TDP_CHECK(ctx, cp(t->obj_vox_off, d->obj_vox_off, 4 * (d->n_obj + 1)));
    TDP_CHECK(ctx, cp(t->obj_hd, d->obj_hd, 40 * d->n_obj));
    TDP_CHECK(ctx, cp(t->vox_box, d->vox_box, 36 * d->n_vox));
    TDP_CHECK(ctx, cp(t->vox_tri_off, d->vox_tri_off, 40 * d->n_vox));
    TDP_CHECK(ctx, cp(t->vox_tri_cnt, d->vox_tri_cnt, 20 * d->n_vox));
    if (d->obj_box && d->obj_transform) {
        TDP_CHECK(ctx, cp(t->obj_box, d->obj_box, 24 * d->n_obj));
        TDP_CHECK(ctx, cp(t->obj_tr, d->obj_transform, 12 * d->n_obj));
    }
    if (d->n_obj) {
        k_vox_owner<<<blocks_for(d->n_obj, kBlock), kBlock, 0, s>>>(d->n_obj, t->obj_vox_off,
                                                                     t->vox_obj);
        TDP_LAUNCHED(ctx);
    }
    t->n_obj = d->n_obj;
    t->n_vox = d->n_vox;
    t->n_tri = d->n_tri;
    t->lazy = false;
    t->lz_pending = 0;
    t->sph_stale = true;
    if (d->tri) return tdp_tile_fill_tri_async(t, 0, d->n_tri, d->tri);
    return TDP_OK;
}

extern "C" int tdp_tile_fill_meta_async(tdp_tile *t, const tdp_tile_desc *d) {
    if (!t || !d) return TDP_ERR_ARG;
    tdp_ctx *ctx = t->ctx;
    if (!t->cap_obj) {
        ctx->err = "tdp_tile_fill_meta_async: the tile was not made by tdp_tile_reserve";
        return TDP_ERR_STATE;
    }
    if (t->lz_cap < t->cap_vox) {
        for (void **p : {(void **)&t->have, (void **)&t->lz_flag, (void **)&t->lz_list,
                         (void **)&t->lz_cnt, (void **)&t->lz_pos, (void **)&t->lz_off}) {
            dfree(ctx, *p);
            *p = nullptr;
        }
        t->lz_cap = 0;
        const size_t n = t->cap_vox;
        int rc = dalloc(ctx, &t->have, n);
        if (rc == TDP_OK) rc = dalloc(ctx, &t->lz_flag, n);
        if (rc == TDP_OK) rc = dalloc(ctx, &t->lz_list, n);
        if (rc == TDP_OK) rc = dalloc(ctx, &t->lz_cnt, n);
        if (rc == TDP_OK) rc = dalloc(ctx, &t->lz_pos, n);
        if (rc == TDP_OK) rc = dalloc(ctx, &t->lz_off, n);
        if (rc != TDP_OK) return rc;
        t->lz_cap = n;
        t->bytes += n * 32;
    }
    tdp_tile_desc m = *d;
    m.tri = nullptr;
    const int rc = tdp_tile_fill_async(t, &m);
    if (rc != TDP_OK) return rc;
    if (t->n_vox) {
        k_fill_u32<<<blocks_for(t->n_vox, kBlock), kBlock, 0, ctx->stream>>>(t->n_vox, t->have, 0u);
        TDP_LAUNCHED(ctx);
    }
    t->lazy = true;
    t->lz_pending = 0;
    t->lz_slot = -1;
    t->put_records = 0;
    return TDP_OK;
}

extern "C" int tdp_tile_put_records(const tdp_tile *t, uint64_t *records) {
    if (!t || !records) return TDP_ERR_ARG;
    *records = t->lazy ? t->put_records : t->n_tri;
    return TDP_OK;
}

extern "C" int tdp_tile_fill_tri_async(tdp_tile *t, size_t first, size_t n, const float *records) {
    if (!t || (n && !records)) return TDP_ERR_ARG;
    tdp_ctx *ctx = t->ctx;
    if (!t->cap_obj || first > t->n_tri || n > t->n_tri - first) {
        ctx->err = "tdp_tile_fill_tri_async: records outside the tile being filled";
        return TDP_ERR_ARG;
    }
    if (!n) return TDP_OK;
    hipStream_t s = ctx->stream;
    t->sph_stale = true;
    TDP_CHECK(ctx, hipMemcpyAsync(t->tri + 3 * first, records, 48 * n, hipMemcpyHostToDevice, s));
    k_tri_bounds<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, t->tri + 3 * first,
                                                          t->bnd + kBnd * first);
    TDP_LAUNCHED(ctx);
    return TDP_OK;
}

extern "C" int tdp_tile_free(tdp_tile *t) {
    if (!t) return TDP_ERR_ARG;
    tdp_ctx *ctx = t->ctx;
    for (void *p : {(void *)t->obj_vox_off, (void *)t->obj_hd, (void *)t->vox_box,
                    (void *)t->vox_obj, (void *)t->vox_tri_off, (void *)t->vox_tri_cnt,
                    (void *)t->tri, (void *)t->bnd, (void *)t->sph, (void *)t->obj_box,
                    (void *)t->obj_tr, (void *)t->have, (void *)t->lz_flag, (void *)t->lz_list, (void *)t->lz_cnt,
                    (void *)t->lz_pos, (void *)t->lz_off})
        dfree(ctx, p);
    delete t;
    return TDP_OK;
}

extern "C" size_t tdp_tile_bytes(const tdp_tile *t) { return t ? t->bytes : 0; }

// ------------------------------------------------------------ batch ops

extern "C" int tdp_tridist_batch(tdp_ctx *ctx, int precision, const float *S, const float *T,
                                 size_t n, void *out) {
    if (!ctx || (n && (!S || !T || !out))) return TDP_ERR_ARG;
    if (precision < 0 || precision > 2) {
        ctx->err = "tdp_tridist_batch: bad precision";
        return TDP_ERR_ARG;
    }
    if (n == 0) return TDP_OK;
    const size_t osz = precision == TDP_PREC_REF32 ? 4 : 8;
    float *dS = nullptr, *dT = nullptr;
    void *dO = nullptr;
    int rc = dalloc(ctx, &dS, 9 * n);
    if (rc == TDP_OK) rc = dalloc(ctx, &dT, 9 * n);
    if (rc == TDP_OK) rc = dalloc(ctx, (char **)&dO, n * osz);
    if (rc == TDP_OK) {
        hipError_t e = copy_sync(ctx, dS, S, 9 * n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = copy_sync(ctx, dT, T, 9 * n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            if (precision == TDP_PREC_REF32)
                k_tridist_batch<float><<<blocks_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                    n, dS, dT, (float *)dO);
            else
                k_tridist_batch<double><<<blocks_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                    n, dS, dT, (double *)dO);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = copy_sync(ctx, out, dO, n * osz, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_tridist_batch: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    }
    dfree(ctx, dS);
    dfree(ctx, dT);
    dfree(ctx, dO);
    return rc;
}

extern "C" int tdp_meshdist_batch(tdp_ctx *ctx, int precision, const float *tri, size_t n_tri,
                                  const uint64_t *os, size_t P, void *out) {
    if (!ctx || (P && (!os || !out)) || (n_tri && !tri)) return TDP_ERR_ARG;
    if (precision < 0 || precision > 2) {
        ctx->err = "tdp_meshdist_batch: bad precision";
        return TDP_ERR_ARG;
    }
    for (size_t p = 0; p < P; p++)
        if (os[4 * p] + os[4 * p + 1] > n_tri || os[4 * p + 2] + os[4 * p + 3] > n_tri) {
            ctx->err = "tdp_meshdist_batch: pair " + std::to_string(p) + " outside the records";
            return TDP_ERR_ARG;
        }
    if (P == 0) return TDP_OK;
    const size_t osz = precision == TDP_PREC_REF32 ? 4 : 8;
    float4 *dR = nullptr, *dB = nullptr;
    uint64_t *dOS = nullptr;
    void *dO = nullptr;
    unsigned long long *dst = nullptr;
    for (size_t p = 0; p < P; p++)
        if (os[4 * p + 1] * os[4 * p + 3] > 0x7fffffffull) {
            ctx->err = "tdp_meshdist_batch: pair " + std::to_string(p) + " has too many triangle pairs";
            return TDP_ERR_ARG;
        }
    int rc = dalloc(ctx, &dR, 3 * n_tri);
    if (rc == TDP_OK) rc = dalloc(ctx, &dB, kBnd * n_tri);
    if (rc == TDP_OK) rc = dalloc(ctx, &dOS, 4 * P);
    if (rc == TDP_OK) rc = dalloc(ctx, (char **)&dO, 3 * P * osz);
    if (rc == TDP_OK) rc = dalloc(ctx, &dst, 4);
    if (rc == TDP_OK) {
        hipError_t e = copy_sync(ctx, dR, tri, 48 * n_tri, hipMemcpyHostToDevice);
        if (e == hipSuccess && n_tri) {
            k_tri_bounds<<<blocks_for(n_tri, kBlock), kBlock, 0, ctx->stream>>>(n_tri, dR, dB);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = copy_sync(ctx, dOS, os, 32 * P, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = set_sync(ctx, dst, 0, 32);
        if (e == hipSuccess) {
            const dim3 g = wave_grid(P);
            if (precision == TDP_PREC_REF32)
                k_meshdist_batch<float, MODE_REF32><<<g, kBlock, 0, ctx->stream>>>(P, dOS, dR, dB, (float *)dO, dst);
            else if (precision == TDP_PREC_FP64)
                k_meshdist_batch<double, MODE_FP64><<<g, kBlock, 0, ctx->stream>>>(P, dOS, dR, dB, (double *)dO, dst);
            else
                k_meshdist_batch<double, MODE_MIXED><<<g, kBlock, 0, ctx->stream>>>(P, dOS, dR, dB, (double *)dO, dst);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = copy_sync(ctx, out, dO, 3 * P * osz, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_meshdist_batch: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    }
    dfree(ctx, dR);
    dfree(ctx, dB);
    dfree(ctx, dOS);
    dfree(ctx, dO);
    dfree(ctx, dst);
    return rc;
}

extern "C" int tdp_meshdist_batch_noalloc(tdp_ctx *ctx, int precision, const float *tri,
                                          size_t n_tri, const uint64_t *os, size_t P, void *out) {
    if (!ctx) return TDP_ERR_ARG;
    if ((P && (!os || !out)) || (n_tri && !tri)) {
        ctx->err = "tdp_meshdist_batch_noalloc: null argument";
        return TDP_ERR_ARG;
    }
    if (precision < 0 || precision > 2) {
        ctx->err = "tdp_meshdist_batch_noalloc: bad precision";
        return TDP_ERR_ARG;
    }
    if (n_tri > kNaMaxTri) {
        ctx->err = "tdp_meshdist_batch_noalloc: " + std::to_string(n_tri) +
                   " triangles, more than its scratch holds (" + std::to_string(kNaMaxTri) + ")";
        return TDP_ERR_ARG;
    }
    if (P == 0) return TDP_OK;
    if (ctx->device >= kNaMaxDevices || !g_na_ptr[ctx->device]) {
        ctx->err = "tdp_meshdist_batch_noalloc: no scratch address for this device";
        return TDP_ERR_STATE;
    }
    std::lock_guard<std::mutex> g(g_na_mu[ctx->device]);
    TDP_CHECK(ctx, hipSetDevice(ctx->device));
    unsigned char *base = (unsigned char *)g_na_ptr[ctx->device];
    float4 *bnd = (float4 *)base;
    unsigned long long *cnt = (unsigned long long *)(base + kNaBndBytes);  // [0..3] stats, [4] flag
    hipStream_t s = ctx->stream;
    // counters start at zero (the array keeps the last call's contents)
    k_fill_u64<<<1, 64, 0, s>>>(cnt, 5, 0ull);
    TDP_LAUNCHED(ctx);
    // the pairs' ranges are checked where they live, on the device
    k_check_pairs<<<blocks_for(P, kBlock), kBlock, 0, s>>>(P, os, (uint64_t)n_tri, cnt + 4);
    TDP_LAUNCHED(ctx);
    unsigned long long bad = 0;
    TDP_CHECK(ctx, hipMemcpyAsync(&bad, cnt + 4, sizeof bad, hipMemcpyDeviceToHost, s));
    TDP_CHECK(ctx, hipStreamSynchronize(s));
    if (bad) {
        ctx->err = "tdp_meshdist_batch_noalloc: a pair lies outside the records or has more "
                   "than 2^31 triangle pairs";
        return TDP_ERR_ARG;
    }
    const float4 *rec = (const float4 *)tri;
    if (n_tri) {
        k_tri_bounds<<<blocks_for(n_tri, kBlock), kBlock, 0, s>>>(n_tri, rec, bnd);
        TDP_LAUNCHED(ctx);
    }
    const dim3 gr = wave_grid(P);
    if (precision == TDP_PREC_REF32)
        k_meshdist_batch<float, MODE_REF32><<<gr, kBlock, 0, s>>>(P, os, rec, bnd, (float *)out, cnt);
    else if (precision == TDP_PREC_FP64)
        k_meshdist_batch<double, MODE_FP64><<<gr, kBlock, 0, s>>>(P, os, rec, bnd, (double *)out, cnt);
    else
        k_meshdist_batch<double, MODE_MIXED><<<gr, kBlock, 0, s>>>(P, os, rec, bnd, (double *)out, cnt);
    TDP_LAUNCHED(ctx);
    TDP_CHECK(ctx, hipStreamSynchronize(s));
    return TDP_OK;
}

extern "C" int tdp_knn_status(tdp_ctx *ctx, size_t N, const uint32_t *prefix,
                              const int32_t *confirmed, const double *rmin, const double *rmax,
                              int k, int32_t *status) {
    if (!ctx || (N && (!prefix || !confirmed || !status))) return TDP_ERR_ARG;
    if (N == 0) return TDP_OK;
    const size_t M = prefix[N];
    if (M && (!rmin || !rmax)) return TDP_ERR_ARG;
    uint32_t *dp = nullptr;
    int32_t *dc = nullptr, *ds = nullptr;
    double *dmn = nullptr, *dmx = nullptr;
    int rc = dalloc(ctx, &dp, N + 1);
    if (rc == TDP_OK) rc = dalloc(ctx, &dc, N);
    if (rc == TDP_OK) rc = dalloc(ctx, &ds, M);
    if (rc == TDP_OK) rc = dalloc(ctx, &dmn, M);
    if (rc == TDP_OK) rc = dalloc(ctx, &dmx, M);
    if (rc == TDP_OK) {
        hipError_t e = copy_sync(ctx, dp, prefix, 4 * (N + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = copy_sync(ctx, dc, confirmed, 4 * N, hipMemcpyHostToDevice);
        if (e == hipSuccess && M) e = copy_sync(ctx, dmn, rmin, 8 * M, hipMemcpyHostToDevice);
        if (e == hipSuccess && M) e = copy_sync(ctx, dmx, rmax, 8 * M, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            k_knn_status<<<blocks_for(N, kBlock), kBlock, 0, ctx->stream>>>(N, dp, dc, dmn, dmx, k, ds);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && M) e = copy_sync(ctx, status, ds, 4 * M, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_knn_status: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    }
    dfree(ctx, dp); dfree(ctx, dc); dfree(ctx, ds); dfree(ctx, dmn); dfree(ctx, dmx);
    return rc;
}

// ------------------------------------------------------------ join

// tile 2's objects binned into a grid of cubic cells of about two objects
// each, in cell order (k_grid_*): the grid, its cell offsets (doff) and the
// objects' records (gobj); device buffers from grab(), held by the caller
template <typename Grab>
static int build_cand_grid(tdp_ctx *ctx, const tdp_tile *t2, Grab &&grab, CandGrid &g,
                           uint64_t *&doff, float4 *&gobj) {
    const uint32_t n2 = (uint32_t)t2->n_obj;
    hipStream_t s = ctx->stream;
    auto grid_err = [&](hipError_t e) {
        ctx->err = std::string("candidate grid: ") + hipGetErrorString(e);
        return TDP_ERR_HIP;
    };
    // grid bounds and cell size
    float *dbnd = nullptr;
    int rc = grab(&dbnd, 8);
    if (rc != TDP_OK) return rc;
    k_grid_bounds<<<1, 1024, 0, s>>>(n2, t2->obj_box, t2->obj_tr, dbnd);
    float hb[7];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hb, dbnd, sizeof hb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return grid_err(e);
    double ext[3], emx = 0, amax = 0;
    bool finite = std::isfinite(hb[6]);
    for (int a = 0; a < 3; a++) {
        finite &= std::isfinite(hb[a]) && std::isfinite(hb[3 + a]);
        ext[a] = std::max(0.0, (double)hb[3 + a] - (double)hb[a]);
        emx = std::max(emx, ext[a]);
        amax = std::max(amax, std::max(std::fabs((double)hb[a]), std::fabs((double)hb[3 + a])));
    }
    if (!finite || emx == 0) {
        // one cell (every target then finds more than a wave holds: host lists)
        for (int a = 0; a < 3; a++) g.lo[a] = 0.f, g.n[a] = 1;
        g.inv_h = 0.f;
    } else {
        const double cells = std::max(1.0, n2 / 2.0);
        double vol = 1;
        for (int a = 0; a < 3; a++) vol *= std::max(ext[a], emx * 1e-3);
        double h = std::cbrt(vol / cells);
        for (;;) {
            double tot = 1;
            for (int a = 0; a < 3; a++) {
                g.n[a] = (int)std::min(std::max(1.0, std::ceil(ext[a] / h)), 1024.0);
                tot *= g.n[a];
            }
            if (tot <= 2.0 * n2 + 1024) break;
            h *= 1.25;
        }
        for (int a = 0; a < 3; a++) g.lo[a] = hb[a];
        g.inv_h = (float)(1.0 / h);
    }
    g.emax = finite ? hb[6] : INFINITY;
    g.tol = (float)(amax * 0x1p-18) + 1e-30f;
    const size_t ncell = (size_t)g.n[0] * g.n[1] * g.n[2];
    // objects in cell order
    uint32_t *dcell = nullptr, *dcnt = nullptr, *dcur = nullptr;
    if (rc == TDP_OK) rc = grab(&dcell, n2);
    if (rc == TDP_OK) rc = grab(&dcnt, ncell + 1);
    if (rc == TDP_OK) rc = grab(&dcur, ncell);
    if (rc == TDP_OK) rc = grab(&doff, ncell + 1);
    if (rc == TDP_OK) rc = grab(&gobj, 3 * (size_t)n2);
    if (rc != TDP_OK) return rc;
    e = hipMemsetAsync(dcnt, 0, 4 * (ncell + 1), s);
    if (e == hipSuccess) e = hipMemsetAsync(dcur, 0, 4 * ncell, s);
    if (e == hipSuccess) {
        k_grid_count<<<blocks_for(n2, kBlock), kBlock, 0, s>>>(n2, t2->obj_tr, g, dcell, dcnt);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return grid_err(e);
    uint64_t placed = 0;
    rc = scan_exclusive<uint32_t>(ctx, dcnt, ncell + 1, doff, &placed);
    if (rc != TDP_OK) return rc;
    if (placed != n2) {
        ctx->err = "candidate grid: binning lost objects";
        return TDP_ERR_HIP;
    }
    k_grid_fill<<<blocks_for(n2, kBlock), kBlock, 0, s>>>(n2, t2->obj_box, t2->obj_tr, dcell, doff,
                                                         dcur, gobj);
    if ((e = hipGetLastError()) != hipSuccess) return grid_err(e);
    return TDP_OK;
}

// kNN lists for tile 2 of more than kCandMaxObj objects: tile 2 binned into a
// grid of cubic cells of about two objects each (k_grid_*), then the targets
// in slices of kGridSlice, each slice's lists compacted on the device and
// kept until the total is known.
constexpr size_t kGridSlice = (size_t)1 << 18;

// Targets in slices of kGridSlice: launch(t0, sn, count, lists, flags) fills
// the slice's counts, lists (stride cap) and flags; each slice's lists are
// compacted on the device and kept until the total is known, then copied out
// as CSR (nothing is written to ids when they do not fit: *n_ids says how
// many there are).
template <typename Grab, typename Launch>
static int cand_slices(tdp_ctx *ctx, size_t N, uint32_t cap, Grab &&grab, Launch &&launch,
                       const char *who, uint32_t *prefix, uint32_t *ids, size_t ids_cap,
                       uint64_t *n_ids, uint8_t *fallback) {
    hipStream_t s = ctx->stream;
    auto hip_err = [&](hipError_t e) {
        ctx->err = std::string(who) + ": " + hipGetErrorString(e);
        return TDP_ERR_HIP;
    };
    const size_t S = std::min(N, kGridSlice);
    uint32_t *dc = nullptr, *dl = nullptr;
    uint8_t *df = nullptr;
    uint64_t *dso = nullptr;
    int rc = grab(&dc, S);
    if (rc == TDP_OK) rc = grab(&dl, S * cap);
    if (rc == TDP_OK) rc = grab(&df, S);
    if (rc == TDP_OK) rc = grab(&dso, S);
    if (rc != TDP_OK) return rc;
    hipError_t e = hipSuccess;
    std::vector<uint32_t> cnt(N);
    std::vector<std::pair<uint32_t *, uint64_t>> parts;
    uint64_t total = 0;
    for (size_t t0 = 0; t0 < N; t0 += S) {
        const size_t sn = std::min(S, N - t0);
        launch(t0, sn, dc, dl, df);
        if ((e = hipGetLastError()) != hipSuccess) return hip_err(e);
        uint64_t st = 0;
        rc = scan_exclusive<uint32_t>(ctx, dc, sn, dso, &st);
        if (rc != TDP_OK) return rc;
        uint32_t *dpart = nullptr;
        if ((rc = grab(&dpart, st)) != TDP_OK) return rc;
        k_cand_compact<<<wave_grid(sn), kBlock, 0, s>>>(sn, cap, dc, dso, dl, dpart);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(cnt.data() + t0, dc, 4 * sn, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(fallback + t0, df, sn, hipMemcpyDeviceToHost, s);
        // the slice buffers are reused by the next slice's kernel on this stream
        if (e != hipSuccess) return hip_err(e);
        parts.push_back({dpart, st});
        total += st;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_err(e);
    *n_ids = total;
    if (total > ids_cap) {
        ctx->err = std::string(who) + ": ids array too small";
        return TDP_ERR_ARG;
    }
    uint64_t at = 0;
    for (auto &pt : parts) {
        if (pt.second &&
            (e = hipMemcpyAsync(ids + at, pt.first, 4 * pt.second, hipMemcpyDeviceToHost, s)) !=
                hipSuccess)
            return hip_err(e);
        at += pt.second;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_err(e);
    for (size_t t = 0; t < N; t++) prefix[t + 1] = prefix[t] + cnt[t];
    return TDP_OK;
}

static int knn_candidates_grid(tdp_ctx *ctx, const tdp_tile *t1, size_t N, const tdp_tile *t2,
                               int self, const int32_t *k_target, uint32_t *prefix, uint32_t *ids,
                               size_t ids_cap, uint64_t *n_ids, uint8_t *fallback) {
    const uint32_t n2 = (uint32_t)t2->n_obj;
    const uint32_t cap = kCandMaxCap;
    hipStream_t s = ctx->stream;
    std::vector<void *> held;
    auto grab = [&](auto **p, size_t count) {
        const int rc = palloc(ctx, p, count);
        if (rc == TDP_OK) held.push_back((void *)*p);
        return rc;
    };
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(s);
        for (void *p : held) pfree(ctx, p);
        return rc;
    };
    auto hip_err = [&](hipError_t e) {
        ctx->err = std::string("tdp_knn_candidates: ") + hipGetErrorString(e);
        return done(TDP_ERR_HIP);
    };
    CandGrid g{};
    uint64_t *doff = nullptr;
    float4 *gobj = nullptr;
    int rc = build_cand_grid(ctx, t2, grab, g, doff, gobj);
    if (rc != TDP_OK) return done(rc);
    hipError_t e = hipSuccess;
    int32_t *dk = nullptr;
    if ((rc = grab(&dk, N)) != TDP_OK) return done(rc);
    if ((e = hipMemcpyAsync(dk, k_target, 4 * N, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_err(e);
    return done(cand_slices(
        ctx, N, cap, grab,
        [&](size_t t0, size_t sn, uint32_t *dc, uint32_t *dl, uint8_t *df) {
            k_knn_candidates_grid<<<wave_grid(sn), kBlock, 0, s>>>(
                t0 + sn, t0, t1->obj_box, t1->obj_tr, g, doff, gobj, self, dk, cap, dc, dl, df);
        },
        "tdp_knn_candidates", prefix, ids, ids_cap, n_ids, fallback));
}

// Within lists for tile 2 of more than kWithinMaxObj objects: the grid of
// tile 2, each object's leaves in walk order (depth-first rank, place in the
// leaf) from the caller's leaf arrays, then k_within_candidates_grid over the
// targets in slices.
static int within_candidates_grid(tdp_ctx *ctx, const tdp_tile *t1, size_t N, const tdp_tile *t2,
                                  int self, const float *tau_target, size_t n_leaves,
                                  const float *leaf_box, const uint32_t *leaf_off,
                                  const uint32_t *leaf_obj, uint32_t *prefix, uint32_t *ids,
                                  size_t ids_cap, uint64_t *n_ids, uint8_t *fallback) {
    const uint32_t n2 = (uint32_t)t2->n_obj;
    const size_t n_lobj = n_leaves ? leaf_off[n_leaves] : 0;
    if (!t2->obj_tr || n_lobj >= 0xffffffffull) {
        ctx->err = "tdp_within_candidates: tile 2 without proxy points, or too many leaf entries";
        return TDP_ERR_ARG;
    }
    const uint32_t cap = kCandMaxCap;
    hipStream_t s = ctx->stream;
    std::vector<void *> held;
    auto grab = [&](auto **p, size_t count) {
        const int rc = palloc(ctx, p, count);
        if (rc == TDP_OK) held.push_back((void *)*p);
        return rc;
    };
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(s);
        for (void *p : held) pfree(ctx, p);
        return rc;
    };
    // each object's leaves, in walk order
    std::vector<uint32_t> moff(n2 + 1, 0u);
    for (size_t k = 0; k < n_lobj; k++) moff[leaf_obj[k] + 1]++;
    for (uint32_t o = 0; o < n2; o++) moff[o + 1] += moff[o];
    std::vector<uint2> mem(std::max<size_t>(n_lobj, 1));
    {
        std::vector<uint32_t> cur(moff.begin(), moff.end() - 1);
        for (size_t l = 0; l < n_leaves; l++)
            for (uint32_t k = leaf_off[l]; k < leaf_off[l + 1]; k++)
                mem[cur[leaf_obj[k]]++] = make_uint2((uint32_t)l, k - leaf_off[l]);
    }
    uint32_t *dmoff = nullptr;
    uint2 *dmem = nullptr;
    float *dlbox = nullptr, *dtau = nullptr;
    int rc = grab(&dmoff, n2 + 1);
    if (rc == TDP_OK) rc = grab(&dmem, mem.size());
    if (rc == TDP_OK) rc = grab(&dlbox, 6 * std::max<size_t>(n_leaves, 1));
    if (rc == TDP_OK) rc = grab(&dtau, N);
    if (rc != TDP_OK) return done(rc);
    hipError_t e = hipMemcpyAsync(dmoff, moff.data(), 4 * (n2 + 1), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dmem, mem.data(), sizeof(uint2) * mem.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && n_leaves)
        e = hipMemcpyAsync(dlbox, leaf_box, 24 * n_leaves, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dtau, tau_target, 4 * N, hipMemcpyHostToDevice, s);
    // the host vectors must outlive the copies
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        ctx->err = std::string("tdp_within_candidates: ") + hipGetErrorString(e);
        return done(TDP_ERR_HIP);
    }
    CandGrid g{};
    uint64_t *doff = nullptr;
    float4 *gobj = nullptr;
    rc = build_cand_grid(ctx, t2, grab, g, doff, gobj);
    if (rc != TDP_OK) return done(rc);
    return done(cand_slices(
        ctx, N, cap, grab,
        [&](size_t t0, size_t sn, uint32_t *dc, uint32_t *dl, uint8_t *df) {
            k_within_candidates_grid<<<wave_grid(sn), kBlock, 0, s>>>(
                t0 + sn, t0, t1->obj_box, g, doff, gobj, self, dtau, dmoff, dmem, dlbox, cap, dc,
                dl, df);
        },
        "tdp_within_candidates", prefix, ids, ids_cap, n_ids, fallback));
}

extern "C" int tdp_knn_candidates(tdp_ctx *ctx, const tdp_tile *t1, size_t N, const tdp_tile *t2,
                                  int self, const int32_t *k_target, uint32_t *prefix,
                                  uint32_t *ids, size_t ids_cap, uint64_t *n_ids,
                                  uint8_t *fallback) {
    if (!ctx) return TDP_ERR_ARG;
    if (!t1 || !t2 || !n_ids || (N && (!k_target || !prefix || !fallback))) {
        ctx->err = "tdp_knn_candidates: null argument";
        return TDP_ERR_ARG;
    }
    *n_ids = 0;
    if (!t1->obj_box || !t2->obj_box || !t1->obj_tr || !t2->obj_tr) {
        ctx->err = "tdp_knn_candidates: tiles uploaded without object boxes";
        return TDP_ERR_ARG;
    }
    if (N > t1->n_obj || t2->n_obj >= 0xffffffffull) {
        ctx->err = "tdp_knn_candidates: more targets than tile 1 holds, or tile 2 too large";
        return TDP_ERR_ARG;
    }
    prefix[0] = 0;
    if (N == 0) return TDP_OK;
    if (t2->n_obj > (size_t)kCandMaxObj)
        return knn_candidates_grid(ctx, t1, N, t2, self, k_target, prefix, ids, ids_cap, n_ids,
                                   fallback);
    const uint32_t cap = kCandMaxCap;
    int32_t *dk = nullptr;
    uint32_t *dc = nullptr, *dl = nullptr, *dids = nullptr;
    uint64_t *doff = nullptr;
    uint8_t *df = nullptr;
    int rc = palloc(ctx, &dk, N);
    if (rc == TDP_OK) rc = palloc(ctx, &dc, N);
    if (rc == TDP_OK) rc = palloc(ctx, &dl, N * (size_t)cap);
    if (rc == TDP_OK) rc = palloc(ctx, &df, N);
    if (rc == TDP_OK) rc = palloc(ctx, &doff, N);
    hipError_t e = hipSuccess;
    uint64_t total = 0;
    if (rc == TDP_OK) {
        e = hipMemcpyAsync(dk, k_target, 4 * N, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) {
            k_knn_candidates<<<wave_grid(N), kBlock, 0, ctx->stream>>>(
                N, t1->obj_box, t1->obj_tr, (uint32_t)t2->n_obj, t2->obj_box, t2->obj_tr, self, dk,
                cap, dc, dl, df);
            e = hipGetLastError();
        }
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_knn_candidates: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    }
    if (rc == TDP_OK) rc = scan_exclusive<uint32_t>(ctx, dc, N, doff, &total);
    *n_ids = total;
    if (rc == TDP_OK && total > ids_cap) {
        ctx->err = "tdp_knn_candidates: ids array too small";
        rc = TDP_ERR_ARG;
    }
    if (rc == TDP_OK) rc = palloc(ctx, &dids, total);
    if (rc == TDP_OK) {
        k_cand_compact<<<wave_grid(N), kBlock, 0, ctx->stream>>>(N, cap, dc, doff, dl, dids);
        e = hipGetLastError();
        if (e == hipSuccess && total)
            e = hipMemcpyAsync(ids, dids, 4 * total, hipMemcpyDeviceToHost, ctx->stream);
        std::vector<uint32_t> cnt(N);
        if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), dc, 4 * N, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(fallback, df, N, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_knn_candidates: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        } else {
            for (size_t t = 0; t < N; t++) prefix[t + 1] = prefix[t] + cnt[t];
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    pfree(ctx, dk);
    pfree(ctx, dc);
    pfree(ctx, dl);
    pfree(ctx, df);
    pfree(ctx, doff);
    pfree(ctx, dids);
    return rc;
}

extern "C" int tdp_within_candidates(tdp_ctx *ctx, const tdp_tile *t1, size_t N,
                                     const tdp_tile *t2, int self, const float *tau_target,
                                     size_t n_leaves, const float *leaf_box,
                                     const uint32_t *leaf_off, const uint32_t *leaf_obj,
                                     uint32_t *prefix, uint32_t *ids, size_t ids_cap,
                                     uint64_t *n_ids, uint8_t *fallback) {
    if (!ctx) return TDP_ERR_ARG;
    // leaf_obj may be null when the leaves hold no objects (an empty tile 2)
    if (!t1 || !t2 || !n_ids || (N && (!tau_target || !prefix || !fallback)) ||
        (n_leaves && (!leaf_box || !leaf_off)) || (n_leaves && leaf_off[n_leaves] && !leaf_obj)) {
        ctx->err = "tdp_within_candidates: null argument";
        return TDP_ERR_ARG;
    }
    *n_ids = 0;
    if (!t1->obj_box || !t2->obj_box) {
        ctx->err = "tdp_within_candidates: tiles uploaded without object boxes";
        return TDP_ERR_ARG;
    }
    if (N > t1->n_obj || t2->n_obj >= 0xffffffffull || n_leaves > 0xffffffffu) {
        ctx->err = "tdp_within_candidates: more targets than tile 1 holds, or tile 2 too large";
        return TDP_ERR_ARG;
    }
    const size_t n_lobj = n_leaves ? leaf_off[n_leaves] : 0;
    for (size_t l = 0; l < n_leaves; l++)
        if (leaf_off[l] > leaf_off[l + 1] || (l == 0 && leaf_off[0] != 0)) {
            ctx->err = "tdp_within_candidates: leaf offsets must start at 0 and not decrease";
            return TDP_ERR_ARG;
        }
    for (size_t i = 0; i < n_lobj; i++)
        if (leaf_obj[i] >= t2->n_obj) {
            ctx->err = "tdp_within_candidates: a leaf names an object outside tile 2";
            return TDP_ERR_ARG;
        }
    prefix[0] = 0;
    if (N == 0) return TDP_OK;
    if (t2->n_obj > (size_t)kWithinMaxObj)
        return within_candidates_grid(ctx, t1, N, t2, self, tau_target, n_leaves, leaf_box,
                                      leaf_off, leaf_obj, prefix, ids, ids_cap, n_ids, fallback);
    const uint32_t cap = kCandMaxCap;
    float *dtau = nullptr, *dlbox = nullptr;
    uint32_t *dloff = nullptr, *dlobj = nullptr, *dc = nullptr, *dl = nullptr, *dids = nullptr;
    uint64_t *doff = nullptr;
    uint8_t *df = nullptr;
    int rc = palloc(ctx, &dtau, N);
    if (rc == TDP_OK) rc = palloc(ctx, &dlbox, 6 * n_leaves + 6);
    if (rc == TDP_OK) rc = palloc(ctx, &dloff, n_leaves + 1);
    if (rc == TDP_OK) rc = palloc(ctx, &dlobj, n_lobj + 1);
    if (rc == TDP_OK) rc = palloc(ctx, &dc, N);
    if (rc == TDP_OK) rc = palloc(ctx, &dl, N * (size_t)cap);
    if (rc == TDP_OK) rc = palloc(ctx, &df, N);
    if (rc == TDP_OK) rc = palloc(ctx, &doff, N);
    hipError_t e = hipSuccess;
    uint64_t total = 0;
    if (rc == TDP_OK) {
        e = hipMemcpyAsync(dtau, tau_target, 4 * N, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess && n_leaves) {
            e = hipMemcpyAsync(dlbox, leaf_box, 24 * n_leaves, hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dloff, leaf_off, 4 * (n_leaves + 1), hipMemcpyHostToDevice,
                                   ctx->stream);
            if (e == hipSuccess && n_lobj)
                e = hipMemcpyAsync(dlobj, leaf_obj, 4 * n_lobj, hipMemcpyHostToDevice, ctx->stream);
        }
        if (e == hipSuccess) {
            k_within_candidates<<<wave_grid(N), kBlock, 0, ctx->stream>>>(
                N, t1->obj_box, (uint32_t)t2->n_obj, t2->obj_box, self, dtau, (uint32_t)n_leaves,
                dlbox, dloff, dlobj, cap, dc, dl, df);
            e = hipGetLastError();
        }
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_within_candidates: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    }
    if (rc == TDP_OK) rc = scan_exclusive<uint32_t>(ctx, dc, N, doff, &total);
    *n_ids = total;
    if (rc == TDP_OK && total > ids_cap) {
        ctx->err = "tdp_within_candidates: ids array too small";
        rc = TDP_ERR_ARG;
    }
    if (rc == TDP_OK) rc = palloc(ctx, &dids, total);
    if (rc == TDP_OK) {
        k_cand_compact<<<wave_grid(N), kBlock, 0, ctx->stream>>>(N, cap, dc, doff, dl, dids);
        e = hipGetLastError();
        if (e == hipSuccess && total)
            e = hipMemcpyAsync(ids, dids, 4 * total, hipMemcpyDeviceToHost, ctx->stream);
        std::vector<uint32_t> cnt(N);
        if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), dc, 4 * N, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(fallback, df, N, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            ctx->err = std::string("tdp_within_candidates: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        } else {
            for (size_t t = 0; t < N; t++) prefix[t + 1] = prefix[t] + cnt[t];
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    pfree(ctx, dtau);
    pfree(ctx, dlbox);
    pfree(ctx, dloff);
    pfree(ctx, dlobj);
    pfree(ctx, dc);
    pfree(ctx, dl);
    pfree(ctx, df);
    pfree(ctx, doff);
    pfree(ctx, dids);
    return rc;
}

extern "C" int tdp_join_destroy(tdp_join *j) {
    if (!j) return TDP_ERR_ARG;
    tdp_ctx *ctx = j->ctx;
    void *ptrs[] = {j->within_t, j->work, j->obj_u, j->prefix, j->pair_obj2, j->pair_target, j->k_t, j->tau_t, j->rmin, j->rmax, j->state,
                    j->round, j->scratch, j->confirmed, j->acc_min, j->acc_max, j->count,
                    j->offs, j->keep, j->pos, j->stats, j->mx_U, j->mx_probe, j->mx_flags,
                    j->mx_key, j->mx_nw, j->mx_woff, j->mx_mask, j->item_vp, j->item_q,
                    j->mx_plan, j->mx_pwb, j->mx_tested, j->mx_prep};
    for (void *p : ptrs) pfree(ctx, p);
    for (int b = 0; b < 2; b++) {
        pfree(ctx, j->v1[b]); pfree(ctx, j->v2[b]); pfree(ctx, j->vpp[b]); pfree(ctx, j->vmin[b]); pfree(ctx, j->vmax[b]);
    }
    for (auto &e : j->ev)
        if (e) ctx->ev_pool.push_back(e);
    for (auto &e : j->sev)
        if (e) ctx->ev_pool.push_back(e);
    pfree(ctx, j->words5);
    pfree(ctx, j->d_words);
    pfree(ctx, j->d_kept);
    if (j->h_kept) {
        std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
        (void)hipHostFree(j->h_kept);
    }
    if (j->gexec) (void)hipGraphExecDestroy(j->gexec);
    for (int k = 0; k < 2; k++)
        if (j->lz_host[k]) (void)tdp_host_free(ctx, j->lz_host[k]);
    pfree(ctx, j->lz_stage);
    delete j;
    return TDP_OK;
}

// within_target: NULL, or per target 1 (within) / 0 (kNN) in a join whose
// params->query is kNN or within
static int join_create(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                       const tdp_join_params *prm, size_t N, const uint32_t *prefix,
                       const uint32_t *cand, const int32_t *k_target, const float *tau_target,
                       const uint8_t *within_target, tdp_join **out) {
    if (!ctx || !t1 || !t2 || !prm || !out || (N && !prefix)) return TDP_ERR_ARG;
    *out = nullptr;
    if (prm->query != TDP_QUERY_KNN && prm->query != TDP_QUERY_WITHIN &&
        prm->query != TDP_QUERY_INTERSECT) {
        ctx->err = "tdp_join_create: bad query";
        return TDP_ERR_ARG;
    }
    if (within_target && prm->query == TDP_QUERY_INTERSECT) {
        ctx->err = "tdp_join_create_kinds: kNN and within targets only";
        return TDP_ERR_ARG;
    }
    if (prm->query == TDP_QUERY_KNN || within_target) {
        bool ok = true;
        for (size_t t = 0; t < N; t++) {
            const bool knn = within_target ? within_target[t] == 0 : true;
            if (knn) ok = ok && (k_target ? k_target[t] : prm->k) >= 1;
        }
        if (!within_target && !k_target) ok = ok && prm->k >= 1;
        if (!ok) {
            ctx->err = "tdp_join_create: k must be >= 1";
            return TDP_ERR_ARG;
        }
    }
    if (prm->precision < 0 || prm->precision > 2 || prm->hausdorff_level < 0 ||
        prm->hausdorff_level > 2) {
        ctx->err = "tdp_join_create: bad precision or hausdorff level";
        return TDP_ERR_ARG;
    }
    if (N > t1->n_obj) {
        ctx->err = "tdp_join_create: more targets than tile-1 objects";
        return TDP_ERR_ARG;
    }
    const size_t P = N ? prefix[N] : 0;
    if (P > 0xffffffffull) {
        ctx->err = "tdp_join_create: too many candidate pairs for 32-bit pair ids";
        return TDP_ERR_ARG;
    }
    for (size_t t = 0; t < N; t++)
        if (prefix[t] > prefix[t + 1]) {
            ctx->err = "tdp_join_create: cand_prefix not monotone";
            return TDP_ERR_ARG;
        }
    for (size_t i = 0; i < P; i++)
        if (cand[i] >= t2->n_obj) {
            ctx->err = "tdp_join_create: candidate id outside tile 2";
            return TDP_ERR_ARG;
        }
    tdp_join *j = new tdp_join();
    j->ctx = ctx;
    j->t1 = t1;
    j->t2 = t2;
    j->prm = *prm;
    j->N = N;
    j->P = P;
    j->rsz = prm->precision == TDP_PREC_REF32 ? 4 : 8;
    j->tune = ctx->tune;
    int rc = TDP_OK;
    if (rc == TDP_OK) rc = palloc(ctx, &j->prefix, N + 1);
    if (rc == TDP_OK) rc = palloc(ctx, &j->pair_obj2, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->pair_target, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->k_t, N);
    if (rc == TDP_OK) rc = palloc(ctx, &j->tau_t, N);
    if (rc == TDP_OK) rc = palloc(ctx, (char **)&j->rmin, P * j->rsz);
    if (rc == TDP_OK) rc = palloc(ctx, (char **)&j->rmax, P * j->rsz);
    if (rc == TDP_OK) rc = palloc(ctx, &j->state, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->round, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->scratch, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->confirmed, N);
    if (rc == TDP_OK) rc = palloc(ctx, (char **)&j->acc_min, P * j->rsz);
    if (rc == TDP_OK) rc = palloc(ctx, (char **)&j->acc_max, P * j->rsz);
    if (rc == TDP_OK) rc = palloc(ctx, &j->count, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->offs, P);
    if (rc == TDP_OK) rc = palloc(ctx, &j->stats, (size_t)kStatN * kStatSlots);
    if (rc == TDP_OK) rc = palloc(ctx, &j->work, 4);
    if (rc == TDP_OK && prm->precision == TDP_PREC_MIXED) rc = palloc(ctx, &j->obj_u, P);
    if (rc != TDP_OK) {
        tdp_join_destroy(j);
        return rc;
    }
    hipError_t e = copy_sync(ctx, j->prefix, prefix, 4 * (N + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && P) e = copy_sync(ctx, j->pair_obj2, cand, 4 * P, hipMemcpyHostToDevice);
    if (e == hipSuccess && N) {
        k_pair_targets<<<blocks_for(N, kBlock), kBlock, 0, ctx->stream>>>(N, j->prefix,
                                                                        j->pair_target);
        e = hipGetLastError();
    }
    if (e == hipSuccess && N) {
        std::vector<int32_t> kv(N);
        std::vector<float> tv(N);
        for (size_t t = 0; t < N; t++) {
            kv[t] = k_target ? k_target[t] : prm->k;
            tv[t] = tau_target ? tau_target[t] : prm->within_dist;
        }
        e = copy_sync(ctx, j->k_t, kv.data(), 4 * N, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = copy_sync(ctx, j->tau_t, tv.data(), 4 * N, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && within_target && N) {
        if (palloc(ctx, &j->within_t, N) != TDP_OK) {
            tdp_join_destroy(j);
            return TDP_ERR_NOMEM;
        }
        std::vector<uint8_t> wv(within_target, within_target + N);
        for (auto &x : wv) x = x ? 1 : 0;
        e = copy_sync(ctx, j->within_t, wv.data(), N, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = set_sync(ctx, j->stats, 0, 8 * kStatN * kStatSlots);
    if (e == hipSuccess) e = set_sync(ctx, j->confirmed, 0, 4 * (N ? N : 1));
    if (e == hipSuccess) e = set_sync(ctx, j->round, 0xff, P ? P : 1);
    if (e == hipSuccess) e = set_sync(ctx, j->state, 0, P ? P : 1);
    auto take_event = [&](hipEvent_t &ev) {
        if (e != hipSuccess) return;
        if (!ctx->ev_pool.empty()) {
            ev = ctx->ev_pool.back();
            ctx->ev_pool.pop_back();
        } else {
            e = hipEventCreate(&ev);
        }
    };
    for (auto &ev : j->ev) take_event(ev);
    for (auto &ev : j->sev) take_event(ev);
    if (e != hipSuccess) {
        ctx->err = std::string("tdp_join_create: ") + hipGetErrorString(e);
        tdp_join_destroy(j);
        return TDP_ERR_HIP;
    }
    *out = j;
    return TDP_OK;
}

extern "C" int tdp_join_create_multi(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                                     const tdp_join_params *prm, size_t N, const uint32_t *prefix,
                                     const uint32_t *cand, const int32_t *k_target,
                                     const float *tau_target, tdp_join **out) {
    return join_create(ctx, t1, t2, prm, N, prefix, cand, k_target, tau_target, nullptr, out);
}

extern "C" int tdp_join_create_kinds(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                                     const tdp_join_params *prm, size_t N, const uint32_t *prefix,
                                     const uint32_t *cand, const uint8_t *within_target,
                                     const int32_t *k_target, const float *tau_target,
                                     tdp_join **out) {
    if (ctx && N && !within_target) {
        ctx->err = "tdp_join_create_kinds: null within_target";
        return TDP_ERR_ARG;
    }
    return join_create(ctx, t1, t2, prm, N, prefix, cand, k_target, tau_target, within_target,
                       out);
}

extern "C" int tdp_join_create(tdp_ctx *ctx, const tdp_tile *t1, const tdp_tile *t2,
                               const tdp_join_params *prm, size_t N, const uint32_t *prefix,
                               const uint32_t *cand, tdp_join **out) {
    return tdp_join_create_multi(ctx, t1, t2, prm, N, prefix, cand, nullptr, nullptr, out);
}

static int ensure_vp(tdp_join *j, size_t n) {
    if (n <= j->cap_vp) return TDP_OK;
    tdp_ctx *ctx = j->ctx;
    size_t cap = std::max(n, (size_t)1024);
    for (int b = 0; b < 2; b++) {
        pfree(ctx, j->v1[b]); pfree(ctx, j->v2[b]); pfree(ctx, j->vpp[b]); pfree(ctx, j->vmin[b]); pfree(ctx, j->vmax[b]);
        j->v1[b] = j->v2[b] = j->vpp[b] = nullptr;
        j->vmin[b] = j->vmax[b] = nullptr;
    }
    pfree(ctx, j->keep); pfree(ctx, j->pos);
    j->keep = nullptr; j->pos = nullptr;
    j->cap_vp = 0;
    int rc = TDP_OK;
    for (int b = 0; b < 2 && rc == TDP_OK; b++) {
        rc = palloc(ctx, &j->v1[b], cap);
        if (rc == TDP_OK) rc = palloc(ctx, &j->v2[b], cap);
        if (rc == TDP_OK) rc = palloc(ctx, &j->vpp[b], cap);
        if (rc == TDP_OK) rc = palloc(ctx, (char **)&j->vmin[b], cap * j->rsz);
        if (rc == TDP_OK) rc = palloc(ctx, (char **)&j->vmax[b], cap * j->rsz);
    }
    if (rc == TDP_OK) rc = palloc(ctx, &j->keep, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->pos, cap);
    if (rc == TDP_OK) j->cap_vp = cap;
    return rc;
}

// per-voxel-pair state of the mixed pipeline
static int ensure_mx(tdp_join *j, size_t n) {
    if (n <= j->cap_mx) return TDP_OK;
    tdp_ctx *ctx = j->ctx;
    const size_t cap = std::max(n, j->cap_vp);
    pfree(ctx, j->mx_U); pfree(ctx, j->mx_probe); pfree(ctx, j->mx_flags); pfree(ctx, j->mx_key);
    pfree(ctx, j->mx_nw); pfree(ctx, j->mx_woff); pfree(ctx, j->mx_plan); pfree(ctx, j->mx_pwb);
    pfree(ctx, j->mx_tested); pfree(ctx, j->mx_prep);
    j->mx_U = nullptr; j->mx_probe = nullptr; j->mx_flags = nullptr; j->mx_key = nullptr;
    j->mx_nw = nullptr; j->mx_woff = nullptr; j->mx_plan = nullptr; j->mx_pwb = nullptr;
    j->mx_tested = nullptr; j->mx_prep = nullptr;
    j->cap_mx = 0;
    int rc = palloc(ctx, &j->mx_U, 3 * cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_probe, 3 * cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_flags, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_key, 3 * cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_nw, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_woff, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_plan, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_pwb, 3 * cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_tested, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->mx_prep, cap);
    if (rc == TDP_OK) j->cap_mx = cap;
    return rc;
}

// candidate bit masks, grow-only
static int ensure_mask(tdp_join *j, size_t n) {
    if (n <= j->cap_mask) return TDP_OK;
    tdp_ctx *ctx = j->ctx;
    const size_t cap = std::max(n + n / 4, (size_t)1 << 16);
    pfree(ctx, j->mx_mask);
    j->mx_mask = nullptr;
    j->cap_mask = 0;
    int rc = palloc(j->ctx, &j->mx_mask, cap);
    if (rc == TDP_OK) j->cap_mask = cap;
    return rc;
}

// flat FP64 work list (voxel pair, triangle pair), grow-only
static int ensure_items(tdp_join *j, size_t n) {
    if (n <= j->cap_items) return TDP_OK;
    tdp_ctx *ctx = j->ctx;
    const size_t cap = std::max(n + n / 4, (size_t)1 << 16);
    pfree(ctx, j->item_vp); pfree(ctx, j->item_q);
    j->item_vp = nullptr; j->item_q = nullptr;
    j->cap_items = 0;
    int rc = palloc(ctx, &j->item_vp, cap);
    if (rc == TDP_OK) rc = palloc(ctx, &j->item_q, cap);
    if (rc == TDP_OK) j->cap_items = cap;
    return rc;
}

template <typename R>
static int join_filter_t(tdp_join *j, uint64_t *pairs_out) {
    tdp_ctx *ctx = j->ctx;
    const size_t P = j->P, N = j->N;
    const int mode = j->prm.query == TDP_QUERY_WITHIN      ? FILTER_WITHIN
                     : j->prm.query == TDP_QUERY_INTERSECT ? FILTER_INTERSECT
                                                           : FILTER_KNN;
    R *rmin = (R *)j->rmin, *rmax = (R *)j->rmax;
    hipStream_t s = ctx->stream;
    TDP_CHECK(ctx, hipEventRecord(j->ev[0], s));
    if (P && mode == FILTER_INTERSECT) {
        // no object-level ranges: every voxel pair whose boxes overlap
        k_state_alive<<<blocks_for(P, kBlock), kBlock, 0, s>>>(P, j->state, j->round);
        TDP_LAUNCHED(ctx);
    } else if (P) {
        k_aggregate<R><<<wave_grid(P), kBlock, 0, s>>>(
            P, j->pair_target, j->pair_obj2, j->t1->obj_vox_off, j->t1->vox_box,
            j->t2->obj_vox_off, j->t2->vox_box, rmin, rmax);
        TDP_LAUNCHED(ctx);
        // a join of kNN and within targets runs both rules, each on its own
        if ((mode == FILTER_KNN || j->within_t) && N) {
            k_filter_knn<R><<<blocks_for(N, kBlock), kBlock, 0, s>>>(
                N, j->prefix, rmin, rmax, j->k_t, j->state, j->round, j->confirmed, j->scratch,
                j->within_t);
            TDP_LAUNCHED(ctx);
        }
        if (mode == FILTER_WITHIN || j->within_t) {
            k_filter_within<R><<<blocks_for(P, kBlock), kBlock, 0, s>>>(
                P, rmin, rmax, j->pair_target, j->tau_t, j->state, j->round, j->within_t);
            TDP_LAUNCHED(ctx);
        }
    }
    if (P) {
        k_survivor_count<R><<<wave_grid(P), kBlock, 0, s>>>(
            P, j->pair_target, j->pair_obj2, j->t1->obj_vox_off, j->t1->vox_box,
            j->t2->obj_vox_off, j->t2->vox_box, rmax, mode, j->tau_t, j->within_t, j->state,
            j->round, j->count);
        TDP_LAUNCHED(ctx);
    }
    uint64_t total = 0;
    int rc = scan_exclusive<uint32_t>(ctx, j->count, P, j->offs, &total);
    if (rc != TDP_OK) return rc;
    rc = ensure_vp(j, total);
    if (rc != TDP_OK) return rc;
    j->cur = 0;
    if (total) {
        k_survivor_write<R><<<wave_grid(P), kBlock, 0, s>>>(
            P, j->pair_target, j->pair_obj2, j->t1->obj_vox_off, j->t1->vox_box,
            j->t2->obj_vox_off, j->t2->vox_box, rmax, mode, j->tau_t, j->within_t, j->state,
            j->offs, j->v1[0],
            j->v2[0], j->vpp[0], (R *)j->vmin[0], (R *)j->vmax[0]);
        TDP_LAUNCHED(ctx);
    }
    TDP_CHECK(ctx, hipEventRecord(j->ev[1], s));
    TDP_CHECK(ctx, hipStreamSynchronize(s));
    float ms = 0;
    TDP_CHECK(ctx, hipEventElapsedTime(&ms, j->ev[0], j->ev[1]));
    j->filter_ms += ms;
    j->n_vp = total;
    j->filter_vp = total;
    j->filtered = true;
    if (pairs_out) *pairs_out = total;
    return TDP_OK;
}

extern "C" int tdp_join_filter(tdp_join *j, uint64_t *pairs_out) {
    if (!j) return TDP_ERR_ARG;
    if (j->filtered) {
        j->ctx->err = "tdp_join_filter: already filtered";
        return TDP_ERR_STATE;
    }
    // the mixed mark pass reads the voxel spheres: a tile filled with its
    // records since its spheres were formed gets them now (tiles filled
    // without records get each voxel's with its records)
    if (j->prm.precision == TDP_PREC_MIXED)
        for (const tdp_tile *t : {j->t1, j->t2})
            if (t->sph_stale && !t->lazy) {
                if (t->n_vox)
                    k_vox_sph<<<stride_grid(5 * t->n_vox * kWave, 4096), kBlock, 0,
                                j->ctx->stream>>>(t->n_vox, t->vox_tri_off, t->vox_tri_cnt,
                                                  t->tri, t->bnd, t->sph);
                TDP_LAUNCHED(j->ctx);
                t->sph_stale = false;
            }
    return j->rsz == 4 ? join_filter_t<float>(j, pairs_out) : join_filter_t<double>(j, pairs_out);
}

// the level the folds take: the Hausdorff level, with the intersection
// join's rule or-ed in
static int fold_level(const tdp_join *j) {
    return j->prm.hausdorff_level | (j->prm.query == TDP_QUERY_INTERSECT ? kFoldIntersect : 0);
}

// quantities a mixed round folds: d (bit 0), lo (bit 1), hi (bit 2)
static uint32_t fold_need(const tdp_join *j, int is_last) {
    const int lv = j->prm.hausdorff_level;
    if (j->prm.query == TDP_QUERY_INTERSECT) return 4u | (lv == 2 ? 2u : 0u) | (lv == 1 ? 1u : 0u);
    return is_last || lv != 2 ? 1u : 6u;
}

// The bounds pass's row test (tuning row_cull) against the cap the mark pass
// will apply: 2 = the object pairs' range max so far, in the last round (after
// a first round). The kernel also takes the voxel pairs' ranges (caps bit 1)
// for rounds that fold lo and hi, but there it passed over few rows and the
// extra loads cost more than they saved (bounds 21.2 -> 23.2 ms over the four
// lo/hi rounds of the 10-NN 200k join), so it is not used. None for the
// intersection join, whose rounds fold their own values.
struct BoundCull {
    int caps = 0;
};
static BoundCull bound_cull(const tdp_join *j, int is_last) {
    BoundCull b;
    if (!j->tune.row_cull || j->prm.query == TDP_QUERY_INTERSECT) return b;
    if (is_last && j->rounds >= 1) b.caps = 2;
    return b;
}

// The mark pass's row test (tuning row_cull) where it pays: rounds that fold
// d alone. In rounds that fold lo and hi it rules out few rows (their
// thresholds carry the Hausdorff terms) and its registers cost occupancy.
static bool row_cull(const tdp_join *j, uint32_t need) { return j->tune.row_cull && need == 1u; }

// calls f(std::integral_constant<uint32_t, NEED>) for the round's quantity set
template <typename F>
static void with_need(uint32_t need, F &&f) {
    switch (need) {
        case 1: f(std::integral_constant<uint32_t, 1>{}); break;
        case 4: f(std::integral_constant<uint32_t, 4>{}); break;
        case 5: f(std::integral_constant<uint32_t, 5>{}); break;
        case 6: f(std::integral_constant<uint32_t, 6>{}); break;
        default: f(std::integral_constant<uint32_t, 7>{}); break;
    }
}

// the per-round decisions of the join's query
template <typename R>
static void launch_decide(tdp_join *j, uint8_t rnd) {
    hipStream_t s = j->ctx->stream;
    const size_t P = j->P, N = j->N;
    R *rmin = (R *)j->rmin, *rmax = (R *)j->rmax;
    if (j->within_t) {  // kNN and within targets: both rules, each on its own
        k_decide_knn<R><<<blocks_for(N, kBlock), kBlock, 0, s>>>(
            N, j->prefix, rmin, rmax, j->k_t, j->state, j->round, j->confirmed, j->scratch, rnd,
            j->within_t);
        k_decide_within<R><<<blocks_for(P, kBlock), kBlock, 0, s>>>(
            P, rmin, rmax, j->pair_target, j->tau_t, j->state, j->round, rnd, j->within_t);
    } else if (j->prm.query == TDP_QUERY_KNN)
        k_decide_knn<R><<<blocks_for(N, kBlock), kBlock, 0, s>>>(
            N, j->prefix, rmin, rmax, j->k_t, j->state, j->round, j->confirmed, j->scratch, rnd,
            nullptr);
    else if (j->prm.query == TDP_QUERY_WITHIN)
        k_decide_within<R><<<blocks_for(P, kBlock), kBlock, 0, s>>>(
            P, rmin, rmax, j->pair_target, j->tau_t, j->state, j->round, rnd, nullptr);
    else
        k_decide_intersect<R><<<blocks_for(P, kBlock), kBlock, 0, s>>>(
            P, rmin, rmax, j->prm.hausdorff_level, j->state, j->round, rnd);
}

// The mark pass's caps on U (see k_mx_mark): the voxel pairs' ranges in
// rounds that fold lo and hi into them, the object pairs' smallest U_d in the
// last round; neither for the intersection join, whose rounds fold the
// round's own values. Queues k_mx_obj_u when the last round's cap applies.
// The kernel reads the arrays whether or not a cap applies (its loads are
// issued together, before any branch); `caps` says which apply: 1 the range
// caps, 2 the object cap.
// The mark pass's thresholds per voxel pair (k_mx_prep) with the caps that
// apply: in the last round the object pairs' smallest U_d (k_mx_obj_u), in a
// round that folds lo and hi the voxel pairs' ranges; none in the
// intersection join, whose fold keeps no range.
static void mark_prep(tdp_join *j, int is_last, uint64_t n, int lo_free) {
    const int c = j->cur;
    hipStream_t s = j->ctx->stream;
    int caps = 0;
    if (j->prm.query != TDP_QUERY_INTERSECT) {
        if (is_last) {
            k_fill_u32<<<blocks_for(j->P, kBlock), kBlock, 0, s>>>(j->P, j->obj_u, 0xff800000u);
            k_mx_obj_u<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, j->mx_plan, j->mx_U,
                                                                 j->mx_key, j->vpp[c], j->obj_u);
            caps = 2;
        } else if (fold_need(j, is_last) == 6u) {
            caps = 1;
        }
    }
    k_mx_prep<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        n, j->mx_plan, j->mx_U, j->mx_key, j->mx_pwb, (const double *)j->vmin[c],
        (const double *)j->vmax[c], j->vpp[c], j->obj_u, caps, lo_free, j->mx_prep);
}

template <typename R, int MODE>
static int join_refine_t(tdp_join *j, int lod, int is_last, uint64_t *pairs_out) {
    tdp_ctx *ctx = j->ctx;
    hipStream_t s = ctx->stream;
    const size_t P = j->P, N = j->N;
    const int evict = j->prm.query != TDP_QUERY_INTERSECT;
    R *rmin = (R *)j->rmin, *rmax = (R *)j->rmax;
    using KT = typename Key<R>::type;
    int slot = -1;
    for (int k = 0; k < 5; k++)
        if (lod == 20 * (k + 1)) slot = k;
    const uint8_t rnd = (uint8_t)std::min(j->rounds + 1, 254);
    const int c = j->cur;
    const uint64_t n = j->n_vp;
    TDP_CHECK(ctx, hipEventRecord(j->ev[2], s));
    j->n_sev = 0;
    // closes the stage that ran since the previous mark
    auto mark = [&](int stage) -> int {
        if (j->n_sev >= 12) return TDP_OK;
        j->sev_stage[j->n_sev] = stage;
        j->stage_n[stage]++;
        TDP_CHECK(ctx, hipEventRecord(j->sev[j->n_sev++], s));
        return TDP_OK;
    };
    k_reset_acc<R, KT><<<blocks_for(P, kBlock), kBlock, 0, s>>>(P, j->state, (KT *)j->acc_min,
                                                                (KT *)j->acc_max);
    TDP_LAUNCHED(ctx);
    const tdp_tile *A = j->t1, *B = j->t2;
    if constexpr (MODE != MODE_MIXED) {  // the mixed form counts in k_mx_plan
        k_round_stats<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
            n, j->v1[c], j->v2[c], A->vox_tri_cnt, B->vox_tri_cnt, slot, j->stats);
        TDP_LAUNCHED(ctx);
    }
    if (mark(TDP_STAGE_RESET) != TDP_OK) return TDP_ERR_HIP;
    if constexpr (MODE != MODE_MIXED) {
        k_refine<R, MODE><<<wave_grid(n), kBlock, 0, s>>>(
            n, j->v1[c], j->v2[c], j->vpp[c], (R *)j->vmin[c], (R *)j->vmax[c], A->vox_tri_off,
            A->vox_tri_cnt, A->tri, B->vox_tri_off, B->vox_tri_cnt, B->tri, A->bnd, B->bnd,
            A->vox_obj, B->vox_obj, A->obj_hd, B->obj_hd, slot, is_last, fold_level(j),
            (KT *)j->acc_min, (KT *)j->acc_max, j->stats);
        TDP_LAUNCHED(ctx);
        if (mark(TDP_STAGE_DIRECT) != TDP_OK) return TDP_ERR_HIP;
    } else {
        int rc = ensure_mx(j, n);
        if (rc != TDP_OK) return rc;
        const dim3 gw = wave_grid(n);
        // quantities this round folds: d (bit 0), lo (bit 1), hi (bit 2)
        const uint32_t need = fold_need(j, is_last);
        k_mx_plan<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
            n, j->v1[c], j->v2[c], A->vox_tri_off, A->vox_tri_cnt, B->vox_tri_off, B->vox_tri_cnt,
            slot, A->sph, B->sph, (uint32_t)j->tune.small_pairs, (uint32_t)j->tune.max_rows,
            j->mx_plan, j->mx_flags, j->mx_nw, j->stats);
        TDP_LAUNCHED(ctx);
        // After the first round a lo/hi round bounds only hi: U_lo then comes
        // from the probes and the voxel pair's range min, which caps it in the
        // mark pass (the minima of lo rise with the LOD, so the range min is
        // the smaller one nearly always); the FP32 lo bound cost more than it
        // saved.
        const bool lo_free = need == 6u && j->rounds >= 1 && j->prm.query != TDP_QUERY_INTERSECT;
        const uint32_t bneed = lo_free ? 4u : need;
        const BoundCull bc = bound_cull(j, is_last);
        with_need(bneed, [&](auto nd) {
            auto go = [&](auto cull) {
                k_mx_bounds<decltype(nd)::value, decltype(cull)::value><<<gw, kBlock, 0, s>>>(
                    n, j->mx_plan, A->tri, A->bnd, B->tri, B->bnd, bneed, j->mx_U, j->mx_probe,
                    j->mx_key, j->v1[c], j->v2[c], A->sph, B->sph, slot,
                    (const double *)j->vmin[c], (const double *)j->vmax[c], j->vpp[c],
                    (const double *)j->rmax, bc.caps);
            };
            if (bc.caps) go(std::true_type{});
            else go(std::false_type{});
        });
        TDP_LAUNCHED(ctx);
        if (mark(TDP_STAGE_BOUNDS) != TDP_OK) return TDP_ERR_HIP;
        uint64_t words = 0;
        rc = scan_exclusive<uint32_t>(ctx, j->mx_nw, n, j->mx_woff, &words);
        if (rc != TDP_OK) return rc;
        if (mark(TDP_STAGE_SCAN) != TDP_OK) return TDP_ERR_HIP;
        rc = ensure_mask(j, words);
        if (rc != TDP_OK) return rc;
        with_need(need, [&](auto nd) {
            k_mx_eval<decltype(nd)::value>
                <<<stride_grid(3 * n, j->tune.eval_blocks), kBlock, 0, s>>>(
                    3 * n, nullptr, j->mx_probe, 1, j->mx_plan, A->tri, B->tri, j->mx_key,
                    j->stats, j->mx_pwb);
        });
        TDP_LAUNCHED(ctx);
        mark_prep(j, is_last, n, lo_free ? 1 : 0);
        TDP_LAUNCHED(ctx);
        if (mark(TDP_STAGE_PROBE) != TDP_OK) return TDP_ERR_HIP;
        with_need(need, [&](auto nd) {
            auto go = [&](auto cull) {
                k_mx_mark<decltype(nd)::value, decltype(cull)::value><<<gw, kBlock, 0, s>>>(
                    n, j->mx_plan, A->tri, A->bnd, B->tri, B->bnd, j->mx_prep, j->mx_woff,
                    j->v1[c], j->v2[c], A->sph, B->sph, slot, j->mx_mask, j->keep,
                    j->mx_tested);
            };
            if (row_cull(j, need)) go(std::true_type{});
            else go(std::false_type{});
        });
        TDP_LAUNCHED(ctx);
        if (mark(TDP_STAGE_MARK) != TDP_OK) return TDP_ERR_HIP;
        if (j->tune.drain) {
            // persistent FP64 stage: no work list, no host read-back
            const int nb = j->tune.drain_blocks ? j->tune.drain_blocks : 4 * ctx->cu_count;
            k_mx_drain<<<nb, kBlock, 0, s>>>(n, j->mx_plan, A->tri, B->tri, j->mx_woff,
                                             j->mx_mask, j->mx_key, j->stats);
            TDP_LAUNCHED(ctx);
            if (mark(TDP_STAGE_EVAL) != TDP_OK) return TDP_ERR_HIP;
        } else {
            uint64_t items = 0;
            rc = scan_exclusive<uint32_t>(ctx, j->keep, n, j->pos, &items);
            if (rc != TDP_OK) return rc;
            if (mark(TDP_STAGE_SCAN) != TDP_OK) return TDP_ERR_HIP;
            if (items) {
                rc = ensure_items(j, items);
                if (rc != TDP_OK) return rc;
                // one wave per kWave voxel pairs
                k_mx_expand<<<wave_grid((n + kWave - 1) / kWave), kBlock, 0, s>>>(
                    n, j->mx_plan, j->mx_woff, j->mx_mask, j->keep, j->pos, j->item_vp,
                    j->item_q);
                TDP_LAUNCHED(ctx);
                if (mark(TDP_STAGE_EXPAND) != TDP_OK) return TDP_ERR_HIP;
                with_need(need, [&](auto nd) {
                    k_mx_eval<decltype(nd)::value>
                        <<<stride_grid(items, j->tune.eval_blocks), kBlock, 0, s>>>(
                            items, j->item_vp, j->item_q, 0, j->mx_plan, A->tri, B->tri,
                            j->mx_key, j->stats, nullptr);
                });
                TDP_LAUNCHED(ctx);
                if (mark(TDP_STAGE_EVAL) != TDP_OK) return TDP_ERR_HIP;
            }
        }
        k_mx_fold<R><<<blocks_for(n, kBlock), kBlock, 0, s>>>(
            n, j->v1[c], j->v2[c], j->vpp[c], (R *)j->vmin[c], (R *)j->vmax[c], A->vox_tri_cnt,
            B->vox_tri_cnt, A->vox_obj, B->vox_obj, A->obj_hd, B->obj_hd, slot, is_last,
            fold_level(j), j->mx_flags, j->mx_key, (KT *)j->acc_min, (KT *)j->acc_max,
            j->stats, j->mx_tested);
        TDP_LAUNCHED(ctx);
        if (mark(TDP_STAGE_FOLD) != TDP_OK) return TDP_ERR_HIP;
    }
    TDP_CHECK(ctx, hipEventRecord(j->ev[3], s));
    k_pair_finish<R><<<blocks_for(P, kBlock), kBlock, 0, s>>>(P, j->state, (const KT *)j->acc_min,
                                                              (const KT *)j->acc_max, rmin, rmax);
    TDP_LAUNCHED(ctx);
    launch_decide<R>(j, rnd);
    TDP_LAUNCHED(ctx);
    k_vp_keep<R><<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, j->vpp[c], (const R *)j->vmin[c],
                                                          j->state, rmax, evict, j->keep);
    TDP_LAUNCHED(ctx);
    uint64_t kept = 0;
    int rc = scan_exclusive<uint32_t>(ctx, j->keep, n, j->pos, &kept);
    if (rc != TDP_OK) return rc;
    const int o = 1 - c;
    if (kept) {
        k_vp_scatter<R><<<blocks_for(n, kBlock), kBlock, 0, s>>>(
            n, j->keep, j->pos, j->v1[c], j->v2[c], j->vpp[c], (const R *)j->vmin[c],
            (const R *)j->vmax[c], j->v1[o], j->v2[o], j->vpp[o], (R *)j->vmin[o], (R *)j->vmax[o]);
        TDP_LAUNCHED(ctx);
    }
    TDP_CHECK(ctx, hipStreamSynchronize(s));
    float ms = 0;
    TDP_CHECK(ctx, hipEventElapsedTime(&ms, j->ev[2], j->ev[3]));
    j->refine_ms += ms;
    for (int k = 0; k < j->n_sev; k++) {
        float sm = 0;
        TDP_CHECK(ctx, hipEventElapsedTime(&sm, k ? j->sev[k - 1] : j->ev[2], j->sev[k]));
        j->stage_ms[j->sev_stage[k]] += sm;
    }
    j->cur = o;
    j->n_vp = kept;
    j->rounds++;
    if (pairs_out) *pairs_out = kept;
    return TDP_OK;
}

// The mixed round with no host synchronisation inside: mask capacity planned
// once from the filter's list (every later list is a subset), the FP64 stage
// persistent (no work-list size to read back), totals kept on the device --
// so the whole round is captured into a hipGraph and launched as one; the
// kept count comes back through pinned memory at the end.
static int plan_graph_rounds(tdp_join *j) {
    tdp_ctx *ctx = j->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t n = j->n_vp;
    int rc = palloc(ctx, &j->words5, 5);
    if (rc == TDP_OK) rc = palloc(ctx, &j->d_words, 1);
    if (rc == TDP_OK) rc = palloc(ctx, &j->d_kept, 1);
    if (rc != TDP_OK) return rc;
    if (!j->h_kept) {
        std::lock_guard<std::recursive_mutex> cg(g_capture_mu);
        TDP_CHECK(ctx, hipHostMalloc((void **)&j->h_kept, sizeof(uint64_t)));
    }
    k_zero_u32<<<1, 64, 0, s>>>((uint32_t *)j->words5, 10);
    TDP_LAUNCHED(ctx);
    const tdp_tile *A = j->t1, *B = j->t2;
    if (n) {
        k_words_bound<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
            n, j->v1[j->cur], j->v2[j->cur], A->vox_tri_off, A->vox_tri_cnt, B->vox_tri_off,
            B->vox_tri_cnt, (uint32_t)j->tune.max_rows, j->words5);
        TDP_LAUNCHED(ctx);
    }
    unsigned long long w5[5];
    TDP_CHECK(ctx, hipMemcpyAsync(w5, j->words5, sizeof w5, hipMemcpyDeviceToHost, s));
    TDP_CHECK(ctx, hipStreamSynchronize(s));
    uint64_t most = 0;
    for (int k = 0; k < 5; k++) most = std::max<uint64_t>(most, w5[k]);
    rc = ensure_mask(j, most);
    if (rc == TDP_OK) rc = ensure_mx(j, n);
    if (rc == TDP_OK) rc = scan_reserve(ctx, n);
    if (rc == TDP_OK) j->planned = true;
    return rc;
}

static int join_refine_graph(tdp_join *j, int lod, int is_last, uint64_t *pairs_out) {
    using R = double;
    using KT = Key<R>::type;
    if (j->n_vp == 0) return join_refine_t<double, MODE_MIXED>(j, lod, is_last, pairs_out);
    tdp_ctx *ctx = j->ctx;
    hipStream_t s = ctx->stream;
    if (!j->planned) {
        int rc = plan_graph_rounds(j);
        if (rc != TDP_OK) return rc;
    }
    const size_t P = j->P;
    const int evict = j->prm.query != TDP_QUERY_INTERSECT;
    R *rmin = (R *)j->rmin, *rmax = (R *)j->rmax;
    int slot = -1;
    for (int k = 0; k < 5; k++)
        if (lod == 20 * (k + 1)) slot = k;
    const uint8_t rnd = (uint8_t)std::min(j->rounds + 1, 254);
    const int c = j->cur, o = 1 - c;
    const uint64_t n = j->n_vp;
    const tdp_tile *A = j->t1, *B = j->t2;
    const dim3 gw = wave_grid(n);
    const uint32_t need = fold_need(j, is_last);
    const int nb = j->tune.drain_blocks ? j->tune.drain_blocks : 4 * ctx->cu_count;

    std::unique_lock<std::recursive_mutex> capture_lock(g_capture_mu);
    TDP_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = TDP_OK;
    auto launched = [&]() {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess && rc == TDP_OK) {
            ctx->err = std::string("captured launch: ") + hipGetErrorString(e);
            rc = TDP_ERR_HIP;
        }
    };
    k_reset_acc<R, KT><<<blocks_for(P, kBlock), kBlock, 0, s>>>(P, j->state, (KT *)j->acc_min,
                                                                (KT *)j->acc_max);
    launched();
    k_mx_plan<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        n, j->v1[c], j->v2[c], A->vox_tri_off, A->vox_tri_cnt, B->vox_tri_off, B->vox_tri_cnt,
        slot, A->sph, B->sph, (uint32_t)j->tune.small_pairs, (uint32_t)j->tune.max_rows,
        j->mx_plan, j->mx_flags, j->mx_nw, j->stats);
    launched();
    const bool lo_free = need == 6u && j->rounds >= 1 && j->prm.query != TDP_QUERY_INTERSECT;
    const uint32_t bneed = lo_free ? 4u : need;
    const BoundCull bc = bound_cull(j, is_last);
    with_need(bneed, [&](auto nd) {
        auto go = [&](auto cull) {
            k_mx_bounds<decltype(nd)::value, decltype(cull)::value><<<gw, kBlock, 0, s>>>(
                n, j->mx_plan, A->tri, A->bnd, B->tri, B->bnd, bneed, j->mx_U, j->mx_probe,
                j->mx_key, j->v1[c], j->v2[c], A->sph, B->sph, slot,
                (const double *)j->vmin[c], (const double *)j->vmax[c], j->vpp[c],
                (const double *)j->rmax, bc.caps);
        };
        if (bc.caps) go(std::true_type{});
        else go(std::false_type{});
    });
    launched();
    if (rc == TDP_OK) rc = scan_launch<uint32_t>(ctx, j->mx_nw, n, j->mx_woff, j->d_words);
    with_need(need, [&](auto nd) {
        k_mx_eval<decltype(nd)::value><<<stride_grid(3 * n, j->tune.eval_blocks), kBlock, 0, s>>>(
            3 * n, nullptr, j->mx_probe, 1, j->mx_plan, A->tri, B->tri, j->mx_key, j->stats,
            j->mx_pwb);
    });
    launched();
    mark_prep(j, is_last, n, lo_free ? 1 : 0);
    launched();
    with_need(need, [&](auto nd) {
        auto go = [&](auto cull) {
            k_mx_mark<decltype(nd)::value, decltype(cull)::value><<<gw, kBlock, 0, s>>>(
                n, j->mx_plan, A->tri, A->bnd, B->tri, B->bnd, j->mx_prep, j->mx_woff, j->v1[c],
                j->v2[c], A->sph, B->sph, slot, j->mx_mask, j->keep, j->mx_tested);
        };
        if (row_cull(j, need)) go(std::true_type{});
        else go(std::false_type{});
    });
    launched();
    k_mx_drain<<<nb, kBlock, 0, s>>>(n, j->mx_plan, A->tri, B->tri, j->mx_woff, j->mx_mask,
                                     j->mx_key, j->stats);
    launched();
    k_mx_fold<R><<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        n, j->v1[c], j->v2[c], j->vpp[c], (R *)j->vmin[c], (R *)j->vmax[c], A->vox_tri_cnt,
        B->vox_tri_cnt, A->vox_obj, B->vox_obj, A->obj_hd, B->obj_hd, slot, is_last,
        fold_level(j), j->mx_flags, j->mx_key, (KT *)j->acc_min, (KT *)j->acc_max,
        j->stats, j->mx_tested);
    launched();
    k_pair_finish<R><<<blocks_for(P, kBlock), kBlock, 0, s>>>(P, j->state, (const KT *)j->acc_min,
                                                              (const KT *)j->acc_max, rmin, rmax);
    launched();
    launch_decide<R>(j, rnd);
    launched();
    k_vp_keep<R><<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, j->vpp[c], (const R *)j->vmin[c],
                                                          j->state, rmax, evict, j->keep);
    launched();
    if (rc == TDP_OK) rc = scan_launch<uint32_t>(ctx, j->keep, n, j->pos, j->d_kept);
    k_vp_scatter<R><<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        n, j->keep, j->pos, j->v1[c], j->v2[c], j->vpp[c], (const R *)j->vmin[c],
        (const R *)j->vmax[c], j->v1[o], j->v2[o], j->vpp[o], (R *)j->vmin[o], (R *)j->vmax[o]);
    launched();
    hipError_t e = hipMemcpyAsync(j->h_kept, j->d_kept, sizeof(uint64_t), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess && rc == TDP_OK) {
        ctx->err = std::string("captured copy: ") + hipGetErrorString(e);
        rc = TDP_ERR_HIP;
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    capture_lock.unlock();
    if (e != hipSuccess && rc == TDP_OK) {
        ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        rc = TDP_ERR_HIP;
    }
    if (rc != TDP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    // reuse the executable when the round's graph has the same shape
    bool ok = false;
    if (j->gexec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        ok = hipGraphExecUpdate(j->gexec, graph, &bad, &res) == hipSuccess &&
             res == hipGraphExecUpdateSuccess;
        (void)hipGetLastError();
        if (!ok) {
            (void)hipGraphExecDestroy(j->gexec);
            j->gexec = nullptr;
        }
    }
    if (!ok) e = hipGraphInstantiate(&j->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    TDP_CHECK(ctx, e);
    TDP_CHECK(ctx, hipEventRecord(j->ev[2], s));
    TDP_CHECK(ctx, hipGraphLaunch(j->gexec, s));
    TDP_CHECK(ctx, hipEventRecord(j->ev[3], s));
    TDP_CHECK(ctx, hipStreamSynchronize(s));
    float ms = 0;
    TDP_CHECK(ctx, hipEventElapsedTime(&ms, j->ev[2], j->ev[3]));
    j->refine_ms += ms;
    j->stage_ms[TDP_STAGE_GRAPH] += ms;
    j->stage_n[TDP_STAGE_GRAPH]++;
    const uint64_t kept = *j->h_kept;
    j->cur = o;
    j->n_vp = kept;
    j->rounds++;
    if (pairs_out) *pairs_out = kept;
    return TDP_OK;
}

static int lod_slot(int lod) {
    for (int k = 0; k < 5; k++)
        if (lod == 20 * (k + 1)) return k;
    return -1;
}

// the tile of join side 1 or 2 when it is filled without records and the
// side has its own list (a self-join's tile is listed once, on side 1)
static tdp_tile *lazy_side(tdp_join *j, int side) {
    const tdp_tile *t = side == 1 ? j->t1 : j->t2;
    if (!t->lazy || (side == 2 && j->t2 == j->t1)) return nullptr;
    return const_cast<tdp_tile *>(t);
}

extern "C" int tdp_join_lazy_needs(tdp_join *j, int lod, int side, const uint32_t **voxels,
                                   size_t *n_voxels, uint64_t *n_records) {
    if (!j || (side != 1 && side != 2) || !voxels || !n_voxels) return TDP_ERR_ARG;
    tdp_ctx *ctx = j->ctx;
    *voxels = nullptr;
    *n_voxels = 0;
    if (n_records) *n_records = 0;
    if (!j->filtered) {
        ctx->err = "tdp_join_lazy_needs: call tdp_join_filter first";
        return TDP_ERR_STATE;
    }
    j->lz_ready[side - 1] = -1;
    tdp_tile *t = lazy_side(j, side);
    const int slot = lod_slot(lod);
    if (!t || slot < 0 || j->n_vp == 0 || t->n_vox == 0) {
        j->lz_ready[side - 1] = lod;
        if (t) t->lz_slot = slot;
        return TDP_OK;
    }
    // voxels listed before and never put are simply listed again (they are
    // marked held only once their records are in place)
    t->lz_pending = 0;
    hipStream_t s = ctx->stream;
    const int mask = side == 1 ? (j->t2 == j->t1 ? 3 : 1) : 2;
    k_fill_u32<<<blocks_for(t->n_vox, kBlock), kBlock, 0, s>>>(t->n_vox, t->lz_flag, 0u);
    TDP_LAUNCHED(ctx);
    k_lazy_flag<<<blocks_for(j->n_vp, kBlock), kBlock, 0, s>>>(
        j->n_vp, j->v1[j->cur], j->v2[j->cur], mask, slot, t->have, t->lz_flag);
    TDP_LAUNCHED(ctx);
    uint64_t n = 0, recs = 0;
    int rc = scan_exclusive<uint32_t>(ctx, t->lz_flag, t->n_vox, t->lz_pos, &n);
    if (rc != TDP_OK) return rc;
    if (n) {
        k_lazy_list<<<blocks_for(t->n_vox, kBlock), kBlock, 0, s>>>(
            t->n_vox, slot, t->lz_flag, t->lz_pos, t->vox_tri_cnt, t->lz_list, t->lz_cnt);
        TDP_LAUNCHED(ctx);
        rc = scan_exclusive<uint32_t>(ctx, t->lz_cnt, n, t->lz_off, &recs);
        if (rc != TDP_OK) return rc;
        uint32_t *&h = j->lz_host[side - 1];
        if (j->lz_host_cap[side - 1] < n) {
            const size_t cap = std::max<size_t>(n, 2 * j->lz_host_cap[side - 1]);
            if (h) tdp_host_free(ctx, h);
            h = nullptr;
            j->lz_host_cap[side - 1] = 0;
            rc = tdp_host_alloc(ctx, 4 * cap, (void **)&h);
            if (rc != TDP_OK) return rc;
            j->lz_host_cap[side - 1] = cap;
        }
        TDP_CHECK(ctx, hipMemcpyAsync(h, t->lz_list, 4 * n, hipMemcpyDeviceToHost, s));
        TDP_CHECK(ctx, hipStreamSynchronize(s));
        *voxels = h;
    }
    *n_voxels = n;
    if (n_records) *n_records = recs;
    t->lz_pending = n;
    t->lz_records = recs;
    t->lz_slot = slot;
    if (n == 0) j->lz_ready[side - 1] = lod;
    return TDP_OK;
}

extern "C" int tdp_join_lazy_put(tdp_join *j, int lod, int side, const float *records) {
    if (!j || (side != 1 && side != 2)) return TDP_ERR_ARG;
    tdp_ctx *ctx = j->ctx;
    tdp_tile *t = lazy_side(j, side);
    const int slot = lod_slot(lod);
    if (!t || !t->lz_pending) {
        if (!t || t->lz_slot == slot) j->lz_ready[side - 1] = lod;
        return TDP_OK;
    }
    if (t->lz_slot != slot) {
        ctx->err = "tdp_join_lazy_put: the listed voxels are of another LOD";
        return TDP_ERR_STATE;
    }
    if (!records) {
        ctx->err = "tdp_join_lazy_put: null records";
        return TDP_ERR_ARG;
    }
    void *dev = nullptr;
    hipError_t e = hipHostGetDevicePointer(&dev, const_cast<float *>(records), 0);
    if (e != hipSuccess || !dev) {
        (void)hipGetLastError();
        ctx->err = "tdp_join_lazy_put: the records are not in pinned memory from tdp_host_alloc";
        return TDP_ERR_ARG;
    }
    // the packed records go to device staging by DMA (the copy engines, at
    // the link's rate) in pieces of at most kPiece records, each scattered
    // into place from device memory
    constexpr uint64_t kPiece = uint64_t(1) << 20;
    const uint64_t recs = t->lz_records, cap = std::min(std::max<uint64_t>(recs, 1), kPiece);
    const size_t n = t->lz_pending;
    t->lz_pending = 0;  // on failure below nothing is marked held
    if (j->lz_stage_cap < cap) {
        pfree(ctx, j->lz_stage);
        j->lz_stage = nullptr;
        j->lz_stage_cap = 0;
        const int rc = palloc(ctx, &j->lz_stage, 3 * cap);
        if (rc != TDP_OK) return rc;
        j->lz_stage_cap = cap;
    }
    for (uint64_t p0 = 0; p0 < std::max<uint64_t>(recs, 1); p0 += kPiece) {
        const uint64_t p1 = std::min(recs, p0 + kPiece);
        if (p1 > p0)
            TDP_CHECK(ctx, hipMemcpyAsync(j->lz_stage, records + 12 * p0, 48 * (p1 - p0),
                                          hipMemcpyHostToDevice, ctx->stream));
        k_lazy_put<<<stride_grid(n * kWave, 2048), kBlock, 0, ctx->stream>>>(
            n, slot, t->lz_list, t->lz_off, t->vox_tri_off, t->vox_tri_cnt, j->lz_stage, p0, p1,
            p1 >= recs ? 1 : 0, t->tri, t->bnd, t->have);
        TDP_LAUNCHED(ctx);
    }
    // the listed voxels' spheres at this slot, from their records in place
    k_vox_sph_list<<<stride_grid(n * kWave, 2048), kBlock, 0, ctx->stream>>>(
        n, slot, t->lz_list, t->vox_tri_off, t->vox_tri_cnt, t->tri, t->bnd, t->sph);
    TDP_LAUNCHED(ctx);
    t->put_records += recs;
    j->lz_ready[side - 1] = lod;
    return TDP_OK;
}

extern "C" int tdp_join_refine(tdp_join *j, int lod, int is_last, uint64_t *pairs_out) {
    if (!j) return TDP_ERR_ARG;
    if (!j->filtered) {
        j->ctx->err = "tdp_join_refine: call tdp_join_filter first";
        return TDP_ERR_STATE;
    }
    if (j->n_vp == 0) {
        if (pairs_out) *pairs_out = 0;
        return TDP_OK;
    }
    for (int side = 1; side <= 2; side++)
        if (lazy_side(j, side) && j->lz_ready[side - 1] != lod) {
            j->ctx->err = "tdp_join_refine: a tile filled without records needs this round's "
                          "records first (tdp_join_lazy_needs, tdp_join_lazy_put)";
            return TDP_ERR_STATE;
        }
    switch (j->prm.precision) {
        case TDP_PREC_REF32: return join_refine_t<float, MODE_REF32>(j, lod, is_last, pairs_out);
        case TDP_PREC_FP64: return join_refine_t<double, MODE_FP64>(j, lod, is_last, pairs_out);
        default:
            if (j->tune.graph) return join_refine_graph(j, lod, is_last, pairs_out);
            return join_refine_t<double, MODE_MIXED>(j, lod, is_last, pairs_out);
    }
}

extern "C" int tdp_join_results(tdp_join *j, uint8_t *report_round, double *rmin, double *rmax) {
    if (!j) return TDP_ERR_ARG;
    tdp_ctx *ctx = j->ctx;
    const size_t P = j->P;
    if (P == 0) return TDP_OK;
    if (report_round)
        TDP_CHECK(ctx, copy_sync(ctx, report_round, j->round, P, hipMemcpyDeviceToHost));
    if (j->rsz == 8) {  // straight into the caller's arrays
        if (rmin) TDP_CHECK(ctx, copy_sync(ctx, rmin, j->rmin, P * 8, hipMemcpyDeviceToHost));
        if (rmax) TDP_CHECK(ctx, copy_sync(ctx, rmax, j->rmax, P * 8, hipMemcpyDeviceToHost));
        return TDP_OK;
    }
    std::unique_ptr<float[]> a(new float[P]);
    for (int k = 0; k < 2; k++) {
        double *dst = k ? rmax : rmin;
        if (!dst) continue;
        TDP_CHECK(ctx, copy_sync(ctx, a.get(), k ? j->rmax : j->rmin, P * 4, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P; p++) dst[p] = a[p];
    }
    return TDP_OK;
}

// the statistics counters, their copies added up (stat_add)
static int read_stats(tdp_join *j, unsigned long long s[kStatN]) {
    unsigned long long c[kStatN * kStatSlots];
    TDP_CHECK(j->ctx, copy_sync(j->ctx, c, j->stats, sizeof c, hipMemcpyDeviceToHost));
    for (int k = 0; k < kStatN; k++) {
        s[k] = 0;
        for (int b = 0; b < kStatSlots; b++) s[k] += c[kStatN * b + k];
    }
    return TDP_OK;
}

extern "C" int tdp_join_stats(tdp_join *j, uint64_t *out5) {
    if (!j || !out5) return TDP_ERR_ARG;
    unsigned long long s[kStatN];
    const int rc = read_stats(j, s);
    if (rc != TDP_OK) return rc;
    out5[0] = s[0];
    out5[1] = s[1];
    out5[2] = s[2];
    out5[3] = j->filter_vp;
    // MIXED: every triangle pair not evaluated in FP64 was excluded by its
    // FP32 interval
    out5[4] = j->prm.precision == TDP_PREC_MIXED && s[0] > s[2] ? s[0] - s[2] : 0;
    return TDP_OK;
}

extern "C" int tdp_join_stats_n(tdp_join *j, uint64_t *out, int n) {
    if (!j || !out || n < 0) return TDP_ERR_ARG;
    uint64_t a[7] = {0, 0, 0, 0, 0, 0, 0};
    const int rc = tdp_join_stats(j, a);
    if (rc != TDP_OK) return rc;
    unsigned long long s[kStatN];
    const int rs = read_stats(j, s);
    if (rs != TDP_OK) return rs;
    a[5] = s[3];
    a[6] = s[4];
    for (int k = 0; k < n && k < 7; k++) out[k] = a[k];
    return TDP_OK;
}

extern "C" int tdp_join_stage_ms(tdp_join *j, double *ms, uint64_t *launches, int n_stages) {
    if (!j || n_stages < 0) return TDP_ERR_ARG;
    for (int k = 0; k < n_stages; k++) {
        if (ms) ms[k] = k < TDP_STAGE_COUNT ? j->stage_ms[k] : 0.0;
        if (launches) launches[k] = k < TDP_STAGE_COUNT ? j->stage_n[k] : 0;
    }
    return TDP_OK;
}

extern "C" int tdp_join_kernel_ms(tdp_join *j, double *refine_ms, double *filter_ms) {
    if (!j) return TDP_ERR_ARG;
    if (refine_ms) *refine_ms = j->refine_ms;
    if (filter_ms) *filter_ms = j->filter_ms;
    return TDP_OK;
}
