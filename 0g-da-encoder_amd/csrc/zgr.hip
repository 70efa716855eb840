// C ABI of the gfx950 encoder (include/zgr.h): context, resident params, the
// encode pipeline and the primitives.
//
// Execution model of one context:
//   - two blob slots, each with its own stream, blob-level device buffers (packed
//     blob, coefficients, AMT outputs, encoded data, Merkle layers) and pinned
//     host staging for the outputs;
//   - two coset lanes, each with its own stream and MSM buffers (evaluations,
//     digits, sorted entries, window trees).
// A blob is packed and inverse-transformed on its slot stream; each coset then runs
// on the next lane (extension NTT, digits, sort, bucket kernel, window trees,
// Horner, Merkle leaves), so one coset's latency-bound tail overlaps the next
// coset's bucket kernel; the Merkle levels and the output copies run on the slot
// stream again.  With two slots, consecutive blobs of a batch overlap as well.
// All streams have the default priority.  Calls on one context are serialised by a
// mutex; the primitives run synchronously on lane 0.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <algorithm>
#include <map>
#include <atomic>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/zgr.h"
#include "kernels.h"
#include "recover.h"
#include "verify.h"
#include "params_pp.h"
#include "ntt.h"
#include "params_kernels.h"

using namespace zgr;

// Intermediate Merkle layers of zgr_merkle_noalloc: a fixed array per GPU, so that
// entry point allocates nothing.  Host code never uses this name as an address (it is a
// host-side stand-in there): it uses the device address kept per device by
// noalloc_scratch(), taken with hipGetSymbolAddress when a context is created.
static __device__ __align__(256) uint8_t scratch[ZGR_MERKLE_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(ZGR_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                                     \
    } while (0)

#define CHECK(expr)           \
    do {                      \
        int rc_ = (expr);     \
        if (rc_) return rc_;  \
    } while (0)

// Host-side range of an entry point or pipeline stage, shown by rocprofv3
// --marker-trace (roctx; a cheap no-op when no tool is attached).
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

inline uint32_t blocks(size_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// device address of `scratch` on each device (each GPU holds its own copy of the array)
std::mutex g_scratch_mu;
std::map<int, uint8_t*> g_scratch_dev;

// Resolve `scratch` on `device` once; the device must be current.
int noalloc_scratch(int device, uint8_t** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_dev.find(device);
    if (it == g_scratch_dev.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ZGR_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                        hipGetErrorString(e));
        it = g_scratch_dev.emplace(device, (uint8_t*)p).first;
    }
    *out = it->second;
    return ZGR_OK;
}

// host memcpy of large staged outputs, split over a few threads (one thread moves about
// 10 GB/s; the encoded data is 32 MB per coset)
void host_copy(uint8_t* dst, const uint8_t* src, size_t n) {
    constexpr size_t PARTS = 4, MIN_PART = 4u << 20;
    if (n < PARTS * MIN_PART) {
        memcpy(dst, src, n);
        return;
    }
    std::thread th[PARTS - 1];
    const size_t part = (n / PARTS + 63) & ~(size_t)63;
    for (size_t k = 1; k < PARTS; k++) {
        const size_t off = k * part, len = off < n ? std::min(part, n - off) : 0;
        th[k - 1] = std::thread([=] { memcpy(dst + off, src + off, len); });
    }
    memcpy(dst, src, std::min(part, n));
    for (auto& t : th) t.join();
}

// dynamic LDS of the bucket kernel: staged bases, reused for the lane reduction
inline size_t accum_lds_bytes(size_t bs) {
    size_t a = bs * LDS_PT * 4, r = (size_t)ACC_THREADS * LDS_XP * 4;
    return a > r ? a : r;
}

struct Timer {
    hipEvent_t a, b;
    std::string name;
};

int dmalloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(ZGR_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return ZGR_OK;
}

int hmalloc(void** p, size_t bytes) {
    hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(ZGR_ERR_OOM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return ZGR_OK;
}

#define DALLOC(ptr, bytes) CHECK(dmalloc((void**)&(ptr), (bytes)))
#define HALLOC(ptr, bytes) CHECK(hmalloc((void**)&(ptr), (bytes)))

struct Lane {
    hipStream_t st = nullptr;
    uint32_t* y = nullptr;        // n Fr: coset evaluations / scalars
    int8_t* digits = nullptr;     // 32 n
    uint32_t* entries = nullptr;  // 32 n u32 (windowed) / FB_W n uint2 (fixed-base level 0)
    uint32_t* counts = nullptr;   // 32 nb
    uint32_t* heap = nullptr;     // nlines * 32 * 2nb XYZZ (36 u32 each)
    uint32_t* bstart = nullptr;   // fixed-base: nb * FB_BSTART bucket starts
    uint32_t* bkt = nullptr;      // fixed-base: nlines * nb * FB_REGION bucket sums (XYZZ)
    // fixed-base chunk groups, levels 1..FB_GL_MAX (level 0: entries / counts / bstart)
    uint2* gent[FB_GL_MAX + 1] = {};
    uint32_t* gcnt[FB_GL_MAX + 1] = {};
    uint32_t* gbst[FB_GL_MAX + 1] = {};
    // wide lines (FbWide): the sort's counts and offsets (nblk x FW_BUCKETS each), its
    // tile sums / bases and entry count, the sorted list (FW_W n), segment boundaries,
    // bucket sums, hand-off partials and reduction partials
    uint32_t* fw_cnt = nullptr;
    uint32_t* fw_off = nullptr;
    uint32_t* fw_tile = nullptr;
    uint2* fw_ent = nullptr;
    uint32_t* fw_seg = nullptr;
    uint32_t* fw_bkt = nullptr;
    uint32_t* fw_hand = nullptr;
    uint32_t* fw_part = nullptr;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev_pre = nullptr, ev_done = nullptr;
    hipEvent_t ev_coset[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_data[3] = {nullptr, nullptr, nullptr};  // coset's encoded data staged
    uint8_t* raw = nullptr;       // staging for host input (31 n bytes)
    uint32_t* pts = nullptr;      // n Fr, column-major blob
    uint32_t* coeff = nullptr;    // n Fr, coefficients (bit-reversed)
    uint32_t* outpts = nullptr;   // per coset: out_slots points (16 u32)
    uint32_t* data = nullptr;     // coset_n * n * 8 u32 canonical encoded data
    std::vector<uint32_t*> mlayers;  // Merkle layers, leaves first
    uint32_t* root = nullptr;
    // pinned host staging
    uint8_t* h_pts = nullptr;
    uint8_t* h_data = nullptr;
    uint8_t* h_merkle = nullptr;
    uint8_t* h_root = nullptr;
    // the request this slot is serving
    const zgr_encoded* out = nullptr;
    bool busy = false;
};

// A chunk-group plan of the fixed-base bucket kernels: per AMT line the log2 of the
// chunks that share one bucket set, and the work lists built from it ([0] every line,
// [1] without the high-basis line).
// A plan of the bucket kernels for one coset, in two variants: v = 0 every line, v = 1
// without the high-basis line (cosets > 0, whose high commitment is coset 0's).
struct FbPlan {
    std::string name;
    uint2* work[2] = {nullptr, nullptr};   // k_fb_sums items, longest first
    uint32_t nwork[2] = {0, 0};
    uint2* rwork[2] = {nullptr, nullptr};  // k_fb_reduce items
    uint32_t nrwork[2] = {0, 0};
    uint64_t plan[2] = {0, 0};  // 4 bits per line: log2 chunks per bucket set
    int gl_used = 0;            // deepest narrow chunk-group level any line uses
    uint32_t nseg = 0;          // bucket-segment waves per wide line
    int red_lanes = 64;         // lanes per k_fb_reduce item (32: two items per wave)
    FbWide wide[2] = {};        // the wide lines (host fields only: counts, bases, lines)
};

// Density classes of a coset's scalars for the plan choice: the share of non-zero
// chunks (coset 0 of a blob shorter than MAX_RAW_DATA_SIZE has zero chunks; cosets 1
// and 2 are dense): at most 1/16, 1/4, 1/2, more.
constexpr int ZGR_DENSITY_CLASSES = 4;
constexpr int ZGR_DENSE = ZGR_DENSITY_CLASSES - 1;

}  // namespace

struct zgr_ctx {
    int device = 0;
    int log_col = 0, log_row = 0, log_n = 0, h = 0, coset_n = 0, nlines = 0;
    size_t n = 0, nb = 0, bs = 0;
    std::mutex mu;
    // zgr_test_inject_fault: countdown to an injected failure per site (0: none); atomic
    // so that it can be armed while a server owns the context
    std::atomic<int> fault[ZGR_FAULT_SITES] = {};
    // zgr_msm's buffers for the sort and bucket kernels (allocated on its first call):
    // entries, counts and offsets per (sort block, bucket), scan tiles, segment bounds,
    // PM_W x FW_BUCKETS bucket slots, reduction partials, hand-off partials, window sums,
    // the work lists (sums waves, then reduction items), the scalars' GLV halves and the
    // table rows of P_i and phi(P_i)
    struct {
        uint2* ent = nullptr;
        uint32_t *cnt = nullptr, *off = nullptr, *tile = nullptr, *seg = nullptr, *bkt = nullptr, *part = nullptr,
                 *hand = nullptr, *heap = nullptr;
        uint2* work = nullptr;
        uint32_t *halves = nullptr, *table = nullptr;  // GLV halves of the scalars; P_i and phi(P_i) rows
        uint32_t* bases = nullptr;                      // the uploaded bases (any zgr_msm path)
        std::vector<uint2> host_work;
        bool lds_ok = false;  // the sort kernels' 128 KB of LDS were granted
    } pm;

    // domain tables
    uint32_t* consts = nullptr;   // K_COUNT Fr
    uint32_t* tw = nullptr;       // n/2 Fr: omega^k
    uint32_t* tw_inv = nullptr;   // n/2 Fr: omega^-k
    // NTT tables in the 2^261 form (k_ntt_tables): omega^+-x split over two small
    // tables (pw_lo, pw_hi), the stage twiddles of the in-LDS sub-transforms, and 2^k / n
    uint32_t* pw_lo[2] = {nullptr, nullptr};  // [0] forward, [1] inverse
    uint32_t* pw_hi[2] = {nullptr, nullptr};
    uint32_t* stw = nullptr;
    uint32_t* stw_inv = nullptr;
    uint32_t* ninv29 = nullptr;
    // coset scale tables, 2^261 form: scale[c][p] = cf_c^bitrev(p) / n, the factor of the
    // coefficient at DIF position p
    uint32_t* scale[3] = {nullptr, nullptr, nullptr};

    // resident params
    uint32_t* bases[3] = {nullptr, nullptr, nullptr};
    bool loaded[3] = {false, false, false};
    // verifier side of each coset's params: the vanishes (2^(h+1) - 2 G2 points, affine
    // Montgomery in the bamt order), then g2 and high_g2 (canonical affine G2)
    uint32_t* van[3] = {nullptr, nullptr, nullptr};
    uint32_t* vg2[3] = {nullptr, nullptr, nullptr};
    bool van_ready[3] = {false, false, false};
    // powers of tau (zgr_power_tau_setup / zgr_power_tau_load), canonical affine
    uint32_t* pt_g1 = nullptr;       // n x 16 u32: [tau^i] G1
    uint32_t* pt_high = nullptr;     // n x 16 u32: [tau^(3n + i)] G1
    uint32_t* pt_g2 = nullptr;       // n x 32 u32: [tau^i] G2
    uint32_t* pt_high_g2 = nullptr;  // 32 u32: [tau^(3n)] G2
    bool pt_ready = false;
    // fixed-base tables of the generators, built on first use
    uint32_t* g1tab = nullptr;  // 32 x 256 affine G1 (k_fixed_base_table)
    uint32_t* g2tab = nullptr;  // 32 x 256 affine G2 (k_g2_fixed_base_table)
    // fixed-base tables (shifted copies of the bases); fb == false: windowed MSM
    bool fb = false;
    uint32_t* fbtab[3] = {nullptr, nullptr, nullptr};
    // wide lines (16-bit windows, fb_wide_lines) and their shifted copies: slot w of
    // fwtab holds line wide_lines[w]
    std::vector<int> wide_lines;
    uint32_t* fwtab[3] = {nullptr, nullptr, nullptr};
    uint32_t fw_nseg = 1;  // sums waves per wide line in the default plan
    uint32_t fw_nseg_max = 1;  // the most any plan uses (sizes the segment buffers)
    // chunk-group plans of the bucket kernels (plans[0] the default), the plan each
    // density class of a coset's scalars uses (set by zgr_autotune), the tuned split of
    // the log_n-point NTT (0: ceil(log_n / 2)) and the tuner's report
    std::vector<FbPlan> plans;
    int plan_of_class[ZGR_DENSITY_CLASSES] = {};
    int ntt_n1 = 0;
    std::string tune_report;

    Lane lanes[2];
    // a third lane for zgr_verify_msm's third coset, with one line's bucket slots
    // (allocated on the first verification that needs it)
    Lane vlane;
    Slot slots[2];
    std::vector<size_t> mlens;
    int next_lane = 0;
    // the MSM stage of one coset (sort, bucket kernels, trees, outputs) captured once per
    // (lane, slot, coset, plan) and replayed on the lane's stream; ZGR_GRAPHS=0 launches
    // the kernels one by one instead
    bool graphs = true;
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> msm_graphs;
    int* flag = nullptr;

    // RS recovery workspace, grown on demand and kept until the context ends
    std::vector<void*> rec_buf;
    std::vector<size_t> rec_cap;

    // profiling
    bool profiling = false;
    std::vector<Timer> pending;
    std::map<std::string, std::pair<double, int>> times;
};

namespace {

size_t tree_pts(const zgr_ctx* c) { return 2 * c->nb - 2; }
size_t out_slots(const zgr_ctx* c) { return 2 * tree_pts(c) + 2; }
size_t merkle_top_nodes(const zgr_ctx* c) { return (size_t)c->coset_n * ((2 * c->nb) - 1); }

// launch on stream st, with optional event timing
template <typename F>
int timed(zgr_ctx* c, hipStream_t st, const char* name, F&& launch) {
    Timer t{};
    if (c->profiling) {
        HIPCHK(hipEventCreate(&t.a));
        HIPCHK(hipEventCreate(&t.b));
        HIPCHK(hipEventRecord(t.a, st));
    }
    launch();
    HIPCHK(hipGetLastError());
    if (c->profiling) {
        HIPCHK(hipEventRecord(t.b, st));
        t.name = name;
        c->pending.push_back(t);
    }
    return ZGR_OK;
}

#define LAUNCH(ctx, st, name, kern, grid, block, shm, ...)                                        \
    CHECK(timed(ctx, st, name, [&] {                                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, st, __VA_ARGS__);                  \
    }))

// wait for every stream of the context, then fold the finished timers
int sync_all(zgr_ctx* c) {
    for (auto& l : c->lanes) HIPCHK(hipStreamSynchronize(l.st));
    if (c->vlane.st) HIPCHK(hipStreamSynchronize(c->vlane.st));
    for (auto& s : c->slots) HIPCHK(hipStreamSynchronize(s.st));
    for (auto& t : c->pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        auto& slot = c->times[t.name];
        slot.first += ms;
        slot.second += 1;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c->pending.clear();
    return ZGR_OK;
}

// ---------------------------------------------------------------------------
// Fr NTT (ntt.h): one pass up to 2^11 points, two passes (four-step) up to 2^22
// ---------------------------------------------------------------------------
int ntt_launch(zgr_ctx* c, hipStream_t st, const NttPass& p, uint32_t groups, uint32_t rows) {
    const uint32_t threads = std::max(64u, std::min(1024u, (1u << p.log_m) / 2));
    LAUNCH(c, st, "ntt", k_ntt_pass, dim3(groups, rows), threads, ntt_lds_bytes(p.log_m), p);
    return ZGR_OK;
}

// high-bit part n1 of a two-pass 2^n transform: the tuned one for the context's own
// size, else ceil(n / 2)
int ntt_split(const zgr_ctx* c, int n) { return n == c->log_n && c->ntt_n1 ? c->ntt_n1 : (n + 1) / 2; }

NttPass ntt_pass_base(const zgr_ctx* c, bool inverse, int n) {
    NttPass p{};
    p.pw_lo = c->pw_lo[inverse ? 1 : 0];
    p.pw_hi = c->pw_hi[inverse ? 1 : 0];
    p.pw_s = ntt_pw_bits(c->log_n);
    p.stw = inverse ? c->stw_inv : c->stw;
    p.log_ctx = c->log_n;
    p.row_len = 1ull << n;
    return p;
}

// natural -> bit-reversed order (DIF) on `rows` consecutive rows of 2^n points, src ->
// dst (they may be the same buffer); no 1/N scaling
int ntt_dif(zgr_ctx* c, hipStream_t st, const uint32_t* src, uint32_t* dst, int n, uint32_t rows, bool inverse) {
    NttPass p = ntt_pass_base(c, inverse, n);
    p.src = src;
    p.dst = dst;
    p.in_rev = p.out_rev = 1;
    if (n <= NTT_MAX_LOG) {
        p.estride = 1;
        p.log_m = n;
        return ntt_launch(c, st, p, 1, rows);
    }
    const int n1 = ntt_split(c, n), n2 = n - n1;
    p.gstride = 1;  // pass A: columns
    p.estride = 1u << n2;
    p.log_m = n1;
    p.log_g = n2;
    p.post = 1;
    p.log_post = n;
    CHECK(ntt_launch(c, st, p, 1u << n2, rows));
    p.src = dst;  // pass B: blocks, in place
    p.gstride = 1u << n2;
    p.estride = 1;
    p.log_m = n2;
    p.log_g = n1;
    p.post = 0;
    return ntt_launch(c, st, p, 1u << n1, rows);
}

// bit-reversed -> natural order (DIT), src -> dst; pre (2^261 form, one factor per
// input point, or null) multiplies the input
int ntt_dit(zgr_ctx* c, hipStream_t st, const uint32_t* src, uint32_t* dst, int n, uint32_t rows, bool inverse,
            const uint32_t* pre) {
    NttPass p = ntt_pass_base(c, inverse, n);
    p.src = src;
    p.dst = dst;
    p.pre = pre;
    if (n <= NTT_MAX_LOG) {
        p.estride = 1;
        p.log_m = n;
        return ntt_launch(c, st, p, 1, rows);
    }
    const int n1 = ntt_split(c, n), n2 = n - n1;
    p.gstride = 1u << n2;  // pass 1: blocks
    p.estride = 1;
    p.log_m = n2;
    p.log_g = n1;
    p.post = 2;
    p.log_post = n;
    CHECK(ntt_launch(c, st, p, 1u << n1, rows));
    p.src = dst;  // pass 2: columns, in place
    p.pre = nullptr;
    p.gstride = 1;
    p.estride = 1u << n2;
    p.log_m = n1;
    p.log_g = n2;
    p.post = 0;
    return ntt_launch(c, st, p, 1u << n2, rows);
}

// ark-poly fft / ifft (natural -> natural, ifft scaled by 1/2^n) of `rows` rows in place;
// above 2^11 points the bit-reversed result goes through tmp (rows * 2^n points)
int ntt_natural(zgr_ctx* c, hipStream_t st, uint32_t* data, uint32_t* tmp, int n, uint32_t rows, bool inverse) {
    const uint32_t* scale = inverse ? c->ninv29 + 8 * (size_t)(c->log_n - n) : nullptr;
    if (n <= NTT_MAX_LOG) {
        NttPass p = ntt_pass_base(c, inverse, n);
        p.src = data;
        p.dst = data;
        p.scale = scale;
        p.estride = 1;
        p.log_m = n;
        p.in_rev = 1;
        return ntt_launch(c, st, p, 1, rows);
    }
    CHECK(ntt_dif(c, st, data, data, n, rows, inverse));
    LAUNCH(c, st, "ntt_unrev", k_ntt_unrev, dim3(blocks((size_t)1 << n, 256), rows), 256, 0, data, tmp, n, scale);
    HIPCHK(hipMemcpyAsync(data, tmp, ((size_t)rows << n) * 32, hipMemcpyDeviceToDevice, st));
    return ZGR_OK;
}

// rows of one ntt_natural launch: grid.y <= 65535, and above 2^11 points what the
// context's staging buffer (n_ctx points) holds
uint32_t ntt_chunk_rows(const zgr_ctx* c, int n, int batch) {
    size_t r = n <= NTT_MAX_LOG ? 65535 : std::min<size_t>(65535, c->n >> n);
    return (uint32_t)std::max<size_t>(1, std::min<size_t>(r, (size_t)batch));
}

// windows per heap tree set: 1 with the fixed-base tables, else MSM_W
int heap_windows(const zgr_ctx* c) { return c->fb ? 1 : MSM_W; }

// multi-line MSM of one coset on a lane: scalars `src` -> chunk sums in the lane heap
// (one tree per line, or one per (line, window)), for the first `lines` lines
// (quotients[0..h), basis, high_basis).  With a plan (fixed-base only, lines == nlines or
// nlines - 1) each line gets only the chunk-group sums its plan gives it (at most what
// its AMT outputs read, fb_group_log), at heap level h - gl; without one every chunk of
// every line.
int msm_lines(zgr_ctx* c, Lane& ln, int coset, const uint32_t* src, int from_y, int lines, const FbPlan* grouped) {
    const int log_bs = c->log_n - c->h;
    if (c->fb) {
        const int gl_used = grouped ? grouped->gl_used : 0;
        const int v = lines == c->nlines ? 0 : 1;
        FbLevels lv{};
        FbOut fo{};
        lv.ent[0] = fo.ent[0] = (uint2*)ln.entries;
        lv.cnt[0] = ln.counts;
        lv.bst[0] = ln.bstart;
        for (int gl = 1; gl <= gl_used; gl++) {
            lv.ent[gl] = fo.ent[gl] = ln.gent[gl];
            lv.cnt[gl] = fo.cnt[gl] = ln.gcnt[gl];
            lv.bst[gl] = fo.bst[gl] = ln.gbst[gl];
        }
        FbWide wd = grouped ? grouped->wide[v] : FbWide{};
        bool narrow = true;  // any narrow line
        if (grouped) {
            narrow = false;
            for (int L = 0; L < lines; L++)
                narrow |= std::find(wd.line, wd.line + wd.nw, (uint32_t)L) == wd.line + wd.nw;
        }
        if (narrow) {
            LAUNCH(c, ln.st, "msm_sort", k_fb_count, (uint32_t)c->nb, (uint32_t)c->bs, 0, src, c->log_n, log_bs,
                   from_y, ln.counts, ln.bstart);
            LAUNCH(c, ln.st, gl_used ? "msm_group_scatter" : "msm_scatter", k_fb_scatter, (uint32_t)c->nb,
                   (uint32_t)c->bs, 0, src, c->log_n, log_bs, from_y, ln.bstart, gl_used, fo);
        }
        if (wd.nw) {  // the wide sort, segment boundaries
            wd.ent = ln.fw_ent;
            wd.off = ln.fw_off;
            wd.tot = ln.fw_tile + 2 * FW_TILES;
            wd.seg = ln.fw_seg;
            wd.table = c->fwtab[coset];
            wd.bkt = ln.fw_bkt;
            wd.hand = ln.fw_hand;
            wd.part = ln.fw_part;
            const size_t shm = FW_BUCKETS * 4;
            LAUNCH(c, ln.st, "msm_wide_sort", k_fw_count, wd.nblk, FW_SORT_T, shm, src, c->log_n, log_bs, from_y,
                   ln.fw_cnt);
            LAUNCH(c, ln.st, "msm_wide_sort", k_fw_tile_sums, FW_TILES, FW_TILE * FW_PARTS, 0, ln.fw_cnt, wd.nblk,
                   ln.fw_tile);
            LAUNCH(c, ln.st, "msm_wide_sort", k_fw_tile_scan, 1, FW_TILES, 0, ln.fw_tile, ln.fw_tile + FW_TILES,
                   ln.fw_tile + 2 * FW_TILES);
            LAUNCH(c, ln.st, "msm_wide_sort", k_fw_offsets, FW_TILES, FW_TILE * FW_PARTS, 0, ln.fw_cnt, wd.nblk,
                   ln.fw_tile + FW_TILES, ln.fw_off);
            LAUNCH(c, ln.st, "msm_wide_sort", k_fw_scatter, wd.nblk, FW_SORT_T, shm, src, c->log_n, log_bs, from_y,
                   ln.fw_off, ln.fw_ent);
            const uint32_t nbnd = wd.segbase[wd.nw] + (uint32_t)wd.nw;
            LAUNCH(c, ln.st, "msm_wide_sort", k_fw_segments, blocks(nbnd, 64), 64, 0, wd, ln.fw_seg);
        }
        const uint2* work = nullptr;
        uint32_t waves = (uint32_t)(c->nb * lines), rwaves = waves;
        const uint2* rwork = nullptr;
        if (grouped) {
            work = grouped->work[v];
            waves = grouped->nwork[v];
            rwork = grouped->rwork[v];
            rwaves = grouped->nrwork[v];
        }
        LAUNCH(c, ln.st, "msm_accum", k_fb_sums, waves, FB_LANES, 0, c->fbtab[coset], lv, wd, work, ln.bkt,
               c->log_n, c->h, lines);
        if (grouped && grouped->red_lanes == 32)
            LAUNCH(c, ln.st, "msm_reduce", k_fb_reduce<32>, (rwaves + 1) / 2, FB_LANES, 0, lv, wd, rwork, rwaves,
                   ln.bkt, ln.heap, c->h, lines);
        else
            LAUNCH(c, ln.st, "msm_reduce", k_fb_reduce<64>, rwaves, FB_LANES, 0, lv, wd, rwork, rwaves, ln.bkt,
                   ln.heap, c->h, lines);
        if (wd.nw)
            LAUNCH(c, ln.st, "msm_combine", k_fw_combine<>, wd.gbase[wd.nw], FB_LANES, 0, wd, ln.heap, c->h);
        return ZGR_OK;
    }
    LAUNCH(c, ln.st, "msm_digits", k_msm_digits, blocks(c->n, 256), 256, 0, src, c->log_n, log_bs, ln.digits,
           from_y);
    uint32_t sort_threads = c->bs < 256 ? (uint32_t)c->bs : 256u;
    LAUNCH(c, ln.st, "msm_sort", k_msm_sort, (uint32_t)(c->nb * MSM_W), sort_threads, 0, ln.digits, log_bs,
           ln.entries, ln.counts);
    LAUNCH(c, ln.st, "msm_accum", k_msm_accum, (uint32_t)(c->nb * lines), ACC_THREADS, accum_lds_bytes(c->bs),
           c->bases[coset], ln.entries, ln.counts, ln.heap, c->log_n, c->h, lines);
    return ZGR_OK;
}

// trees over the chunks (+ Horner over the windows): AMT outputs of one coset into
// `out` (out_slots points).  With lines == h + 1 the high-basis line was not computed
// and its output slot (the last one) is left for the caller to fill.
int amt_outputs(zgr_ctx* c, Lane& ln, uint32_t* out, int lines, const FbPlan* grouped) {
    const int nw = heap_windows(c);
    const uint64_t plan = grouped ? grouped->plan[lines == c->nlines ? 0 : 1] : 0;
    for (int d = c->h - 1; d >= 0; d--) {
        uint32_t total = (uint32_t)(lines * nw) << d;
        LAUNCH(c, ln.st, "msm_tree", k_msm_tree_level, blocks(total, 64), 64, 0, ln.heap, c->h, d, total, nw, plan);
    }
    uint32_t slots = (uint32_t)out_slots(c) - (lines == c->nlines ? 0u : 1u);
    LAUNCH(c, ln.st, "msm_horner", k_msm_horner, blocks(slots, 64), 64, 0, ln.heap, c->h, nw, out, slots);
    return ZGR_OK;
}

// Lines that go wide (16-bit windows): those whose AMT outputs are sums over groups of
// at least 2^FW_MIN_GL chunks -- quotient lines 0..h-1-FW_MIN_GL and the high-basis line
// -- when the chunks split into whole blocks of the wide sort.  ZGR_WIDE=0 keeps every
// line narrow; ZGR_WIDE_MIN_GL raises the threshold (log2 chunks per group, at least
// FW_MIN_GL).
std::vector<int> fb_wide_lines(const zgr_ctx* c) {
    std::vector<int> w;
    const char* env = getenv("ZGR_WIDE");
    if ((env && env[0] == '0') || c->h < FW_MIN_GL || c->nlines > FW_MAX_LINES) return w;
    const char* genv = getenv("ZGR_WIDE_MIN_GL");
    const int min_gl = std::max(FW_MIN_GL, genv ? atoi(genv) : FW_MIN_GL);
    for (int L = 0; L < c->nlines; L++) {
        const int g = L < c->h ? c->h - 1 - L : (L == c->h ? 0 : c->h);
        if (g >= min_gl) w.push_back(L);
    }
    return w;
}

// log2 of the chunks per bucket set of wide line L (its outputs' group, uncapped)
int fw_group_log(const zgr_ctx* c, int L) { return L < c->h ? c->h - 1 - L : c->h; }

// Work lists of the bucket kernels, one item per wave (FB_WIDE and the narrow / wide
// item layouts in kernels.h).  k_fb_sums takes them longest first (dense estimate:
// entries / 64 per lane), so that the long waves start first; within a narrow level the
// lines of one chunk group are adjacent (they read the same sorted entries), and the
// wide lines of one segment are adjacent (they read nearly the same entry range).
// The default plan puts the wide lines (fb_wide_lines) in 16-bit windows and every
// other line at the largest narrow level its outputs allow (fb_group_log).  With no
// wide lines it has one exception: when the waves of the deepest level outnumber the
// wave slots k_fb_sums has on the device (4 SIMDs per CU, 2 waves each at its register
// count), the ones left over would run last and alone, so the high-basis line (read only
// at the root) drops one level.  At 2^20 x 12 lines that turns 2304 four-chunk waves into
// 2048 and measured 4 % faster on dense scalars.  The other candidates of zgr_autotune:
// "narrow" (the default without wide lines), "max" (no drop), "half" (at most 2-chunk
// groups) and "chunk" (one bucket set per chunk), plus "wide-coarse" / "wide-fine" (the
// default with half / twice its wide segments).  ZGR_GROUP_PLAN=none makes the only plan
// "chunk", a candidate's name makes it that candidate, and a comma-separated list sets
// each line's narrow level explicitly (each capped at fb_group_log); each turns the tuner
// off.
constexpr int FB_SUMS_WAVES_PER_SIMD = 2;

int build_plan(zgr_ctx* c, const std::string& kind, const std::vector<int>& given, uint32_t nseg, FbPlan* out) {
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const size_t slots = (size_t)cus * 4 * FB_SUMS_WAVES_PER_SIMD;
    const int log_bs = c->log_n - c->h;
    out->name = kind;
    out->nseg = nseg;
    const bool wide_plan = kind == "default" || kind.compare(0, 5, "wide-") == 0;
    // the dense-scalar plans reduce with two items per wave (fewer multiplications); the
    // sparse-oriented ones ("max", "half", "chunk", "given") one item per wave (shorter
    // waves: their few non-zero groups leave the device idle)
    out->red_lanes = (wide_plan || kind == "narrow") ? 32 : 64;
    for (int v = 0; v < 2; v++) {
        const int lines = c->nlines - v;
        std::vector<int> gl(lines, 0);
        std::vector<bool> wide(lines, false);
        if (wide_plan)
            for (int L : c->wide_lines)
                if (L < lines) wide[L] = true;
        if (kind != "chunk")
            for (int L = 0; L < lines; L++) {
                gl[L] = fb_group_log(L, c->h);
                if (kind == "half") gl[L] = std::min(gl[L], 1);
                if (kind == "given" && L < (int)given.size()) gl[L] = std::max(0, std::min(gl[L], given[L]));
            }
        if ((kind == "default" || kind == "narrow") && lines == c->h + 2 && gl[c->h + 1] > 0 && !wide[c->h + 1]) {
            int top = 0;
            for (int L = 0; L < lines; L++)
                if (!wide[L]) top = std::max(top, gl[L]);
            size_t deepest = 0;
            for (int L = 0; L < lines; L++) deepest += (!wide[L] && gl[L] == top) ? c->nb >> top : 0;
            if (gl[c->h + 1] == top && deepest > slots && deepest - (c->nb >> top) <= slots) gl[c->h + 1]--;
        }
        FbWide& wd = out->wide[v];
        wd = FbWide{};
        wd.nblk = (uint32_t)(c->nb >> FW_BLK_LOG);
        for (int L = 0; L < lines; L++) {
            if (!wide[L]) continue;
            const int w = wd.nw++;
            const int g = fw_group_log(c, L);
            wd.line[w] = (uint32_t)L;
            wd.sh[w] = (uint32_t)(g - FW_BLK_LOG);
            wd.segbase[w + 1] = wd.segbase[w] + nseg;
            wd.gbase[w + 1] = wd.gbase[w] + (uint32_t)(c->nb >> g);
            gl[L] = g;
        }
        out->plan[v] = 0;
        for (int L = 0; L < lines; L++) {
            out->plan[v] |= (uint64_t)gl[L] << (4 * L);
            if (!wide[L]) out->gl_used = std::max(out->gl_used, gl[L]);
        }
        // (dense length, item, region) of every sums wave; narrow regions in list order
        struct Item {
            uint64_t len;
            uint32_t x, y;
        };
        std::vector<Item> sums, red, red_n;
        uint32_t region = 0;
        for (int g = FB_GL_MAX; g >= 0; g--)
            for (uint32_t S = 0; S < (uint32_t)(c->nb >> g); S++)
                for (int L = 0; L < lines; L++)
                    if (!wide[L] && gl[L] == g) {
                        const uint32_t x = (uint32_t)L | (uint32_t)g << 8 | S << 10;
                        sums.push_back({((uint64_t)FB_W << (g + log_bs)) / FB_LANES, x, region});
                        red_n.push_back({0, x, region++});
                    }
        for (uint32_t s = 0; s < nseg; s++)
            for (int w = 0; w < wd.nw; w++)
                sums.push_back({((uint64_t)FW_W * c->n / nseg) / FB_LANES, FB_WIDE | (uint32_t)w | s << 8,
                                wd.segbase[w] + s});
        for (int w = 0; w < wd.nw; w++)  // wide reductions first, then the narrow ones
            for (uint32_t S = 0; S < wd.gbase[w + 1] - wd.gbase[w]; S++)
                for (uint32_t q = 0; q < (uint32_t)FW_RSEGS; q++)
                    red.push_back({0, FB_WIDE | (uint32_t)w | q << 4 | S << 8, 0});
        red.insert(red.end(), red_n.begin(), red_n.end());
        std::stable_sort(sums.begin(), sums.end(), [](const Item& a, const Item& b) { return a.len > b.len; });
        std::vector<uint2> ws, wr;
        for (const Item& it : sums) ws.push_back(make_uint2(it.x, it.y));
        for (const Item& it : red) wr.push_back(make_uint2(it.x, it.y));
        DALLOC(out->work[v], ws.size() * sizeof(uint2));
        HIPCHK(hipMemcpy(out->work[v], ws.data(), ws.size() * sizeof(uint2), hipMemcpyHostToDevice));
        out->nwork[v] = (uint32_t)ws.size();
        DALLOC(out->rwork[v], wr.size() * sizeof(uint2));
        HIPCHK(hipMemcpy(out->rwork[v], wr.data(), wr.size() * sizeof(uint2), hipMemcpyHostToDevice));
        out->nrwork[v] = (uint32_t)wr.size();
    }
    return ZGR_OK;
}

int build_work_lists(zgr_ctx* c) {
    const char* mode = getenv("ZGR_GROUP_PLAN");
    std::vector<int> given;
    std::vector<std::string> kinds;
    static const char* named[] = {"default", "narrow", "wide-coarse", "wide-fine", "max", "half", "chunk"};
    const bool is_named = mode && std::any_of(std::begin(named), std::end(named),
                                              [&](const char* k) { return !strcmp(mode, k); });
    if (mode && !strcmp(mode, "none")) {
        kinds = {"chunk"};
    } else if (is_named) {  // one of the tuner's candidates, alone
        if (c->wide_lines.empty() && std::string(mode).compare(0, 4, "wide") == 0)
            return fail(ZGR_ERR_INVALID_ARG, "ZGR_GROUP_PLAN=%s: this geometry has no wide lines", mode);
        kinds = {mode};
    } else if (mode) {  // one level per line, comma-separated
        for (const char* q = mode; *q;) {
            given.push_back((int)strtol(q, (char**)&q, 10));
            if (*q == ',') q++;
            else if (*q) return fail(ZGR_ERR_INVALID_ARG, "ZGR_GROUP_PLAN=%s: expected none or levels like 2,2,1,0", mode);
        }
        kinds = {"given"};
    } else {
        kinds = {"default", "max", "half", "chunk"};
        // with wide lines the tuner also weighs the window width (every line narrow) and
        // the wide sort's wave granularity (half and twice the default segments per line)
        if (!c->wide_lines.empty()) kinds.insert(kinds.begin() + 1, {"narrow", "wide-coarse", "wide-fine"});
    }
    c->plans.resize(kinds.size());
    for (size_t k = 0; k < kinds.size(); k++) {
        uint32_t nseg = c->fw_nseg;
        if (kinds[k] == "wide-coarse") nseg = std::max(1u, c->fw_nseg / 2);
        if (kinds[k] == "wide-fine") nseg = c->fw_nseg_max;
        CHECK(build_plan(c, kinds[k], given, nseg, &c->plans[k]));
    }
    return ZGR_OK;
}

// the plan for a coset's scalars of density class cls, or null without fixed-base tables
const FbPlan* plan_for(const zgr_ctx* c, int cls) { return c->fb ? &c->plans[c->plan_of_class[cls]] : nullptr; }

// density class of coset 0's scalars for a blob of len raw bytes: the raw rows that hold
// data, plus the last row (it holds the u32 length footer unless len = 0)
int density_class(const zgr_ctx* c, size_t len) {
    const size_t row_bytes = 31 * ((size_t)1 << c->log_col), rows = (size_t)1 << c->log_row;
    size_t nz = std::min(rows, (len + row_bytes - 1) / row_bytes);
    if (len && nz < rows) nz++;
    if (16 * nz <= rows) return 0;
    if (4 * nz <= rows) return 1;
    if (2 * nz <= rows) return 2;
    return ZGR_DENSE;
}

// shifted-copy tables of one coset from its resident bases (after every params upload):
// the narrow table of every line, and the wide table of the wide lines
int build_fb(zgr_ctx* c, int coset) {
    if (!c->fb) return ZGR_OK;
    const int nw = (int)c->wide_lines.size();
    if (!c->fbtab[coset]) DALLOC(c->fbtab[coset], (size_t)c->nlines * c->n * FB_W * 64);
    if (nw && !c->fwtab[coset]) DALLOC(c->fwtab[coset], (size_t)nw * c->n * FW_W * 64);
    // lane 0's bucket-sum buffer is the chain scratch: its stream runs the build and
    // nothing else uses it meanwhile (the context lock is held)
    hipStream_t st = c->lanes[0].st;
    const size_t bkt_bytes = (size_t)c->nlines * c->nb * FB_REGION * X29_WORDS * 4;
    auto build = [&](auto kern, int W, uint32_t* table, uint32_t src, uint32_t dst) -> int {
        const size_t per_base = (size_t)(W - 1) * 40 * 4;
        const size_t step = std::min(c->n, bkt_bytes / per_base);
        for (size_t first = 0; first < c->n; first += step) {
            const uint32_t count = (uint32_t)std::min(step, c->n - first);
            LAUNCH(c, st, "fb_build", kern, blocks(count, 64), 64, 0, c->bases[coset], table, c->lanes[0].bkt,
                   c->log_n, c->h, src, dst, (uint32_t)first, count);
        }
        return ZGR_OK;
    };
    for (int L = 0; L < c->nlines; L++)
        CHECK(build(k_fb_build<FB_C, FB_W>, FB_W, c->fbtab[coset], (uint32_t)L, (uint32_t)L));
    for (int w = 0; w < nw; w++)
        CHECK(build(k_fb_build<FW_C, FW_W>, FW_W, c->fwtab[coset], (uint32_t)c->wide_lines[w], (uint32_t)w));
    return ZGR_OK;
}

int ensure_params(zgr_ctx* c, int coset) {
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    if (!c->loaded[coset]) return fail(ZGR_ERR_NOT_LOADED, "params for coset %d are not resident", coset);
    return ZGR_OK;
}

int alloc_bases(zgr_ctx* c, int coset) {
    if (!c->bases[coset]) DALLOC(c->bases[coset], (size_t)c->nlines * c->n * 64);
    return ZGR_OK;
}

int check_len(const zgr_ctx* c, size_t len) {
    size_t max_raw = 31 * c->n - 4;
    if (len > max_raw)
        return fail(ZGR_ERR_TOO_LARGE, "TooLargeBlob { actual: %zu, expected_max: %zu }", len, max_raw);
    return ZGR_OK;
}

// Issue the whole encode of one blob (raw bytes in device memory) into slot s.
// Returns once everything is queued; the outputs land in the slot's pinned staging.
// an armed fault at `site` counts down and fires when it reaches zero
bool fault_fires(zgr_ctx* c, int site) {
    int v = c->fault[site].load();
    while (v > 0 && !c->fault[site].compare_exchange_weak(v, v - 1)) {
    }
    return v == 1;
}

int issue_blob(zgr_ctx* c, Slot& s, const uint8_t* raw_dev, size_t len, const zgr_encoded* out) {
    Range range_("issue_blob");
    if (fault_fires(c, ZGR_FAULT_ISSUE)) return fail(ZGR_ERR_DEVICE, "injected fault: blob issue");
    const uint32_t n = (uint32_t)c->n;
    LAUNCH(c, s.st, "pack", k_pack, blocks(n, 256), 256, 0, raw_dev, len, c->log_col, c->log_row, s.pts);
    if (c->coset_n > 1) CHECK(ntt_dif(c, s.st, s.pts, s.coeff, c->log_n, 1, true));  // coefficients, bit-reversed
    HIPCHK(hipEventRecord(s.ev_pre, s.st));
    const uint32_t slots = (uint32_t)out_slots(c);
    for (int cs = 0; cs < c->coset_n; cs++) {
        Lane& ln = c->lanes[c->next_lane];
        c->next_lane ^= 1;
        HIPCHK(hipStreamWaitEvent(ln.st, s.ev_pre, 0));
        const uint32_t* y = s.pts;
        if (cs > 0) {  // evaluations on coset cs: DIT of the coefficients times cf^i / N
            CHECK(ntt_dit(c, ln.st, s.coeff, ln.y, c->log_n, 1, false, c->scale[cs]));
            y = ln.y;
        }
        // high_commitment = [tau^(3N) f(tau)] G1 is the same group element for every
        // coset (coset c's evaluations interpolate f(cf_c X)), so the high-basis line
        // runs for coset 0 only and finish_blob copies its result to the others
        const int lines = cs == 0 ? c->nlines : c->nlines - 1;
        const FbPlan* plan = plan_for(c, cs == 0 ? density_class(c, len) : ZGR_DENSE);
        uint32_t* outp = s.outpts + (size_t)cs * slots * 16;
        auto stage = [&] {
            CHECK(msm_lines(c, ln, cs, y, 1, lines, plan));
            return amt_outputs(c, ln, outp, lines, plan);
        };
        if (c->graphs && plan && !c->profiling) {  // (event timers cannot be captured)
            const auto key = std::make_tuple((int)(&ln - c->lanes), (int)(&s - c->slots), cs,
                                             (int)(plan - c->plans.data()));
            auto it = c->msm_graphs.find(key);
            if (it == c->msm_graphs.end()) {
                hipGraph_t g = nullptr;
                hipGraphExec_t ex = nullptr;
                HIPCHK(hipStreamBeginCapture(ln.st, hipStreamCaptureModeThreadLocal));
                const int rc = stage();
                const hipError_t ec = hipStreamEndCapture(ln.st, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ec != hipSuccess) return fail(ZGR_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
                const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) return fail(ZGR_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
                it = c->msm_graphs.emplace(key, ex).first;
            }
            HIPCHK(hipGraphLaunch(it->second, ln.st));
        } else {
            CHECK(stage());
        }
        LAUNCH(c, ln.st, "merkle_leaves", k_merkle_leaves, blocks(n / 8, 128), 128, 0, y, c->log_col, c->log_row,
               cs, s.data, s.mlayers[0]);
        HIPCHK(hipEventRecord(s.ev_coset[cs], ln.st));
    }
    // each coset's encoded data goes to the pinned staging as soon as its leaves are
    // hashed, so finish_blob copies it out while later cosets still compute
    for (int cs = 0; cs < c->coset_n; cs++) {
        HIPCHK(hipStreamWaitEvent(s.st, s.ev_coset[cs], 0));
        if (out->data)
            HIPCHK(hipMemcpyAsync(s.h_data + (size_t)cs * c->n * 32, (const uint8_t*)s.data + (size_t)cs * c->n * 32,
                                  c->n * 32, hipMemcpyDeviceToHost, s.st));
        HIPCHK(hipEventRecord(s.ev_data[cs], s.st));
    }
    for (size_t l = 1; l < s.mlayers.size(); l++)
        LAUNCH(c, s.st, "merkle_level", k_merkle_level, blocks(c->mlens[l], 128), 128, 0, s.mlayers[l - 1],
               s.mlayers[l], (uint32_t)c->mlens[l]);
    LAUNCH(c, s.st, "file_root", k_file_root, 1, 1, 0, s.mlayers.back(), (int)c->mlens.back(), s.root);
    // outputs into pinned staging
    HIPCHK(hipMemcpyAsync(s.h_pts, s.outpts, (size_t)c->coset_n * slots * 64, hipMemcpyDeviceToHost, s.st));
    if (out->merkle) {
        size_t off = 0;
        const int L = (int)s.mlayers.size();
        for (int k = 0; k <= c->h; k++) {  // tree[k] = layer (L - 1 - k)
            const int li = L - 1 - k;
            HIPCHK(hipMemcpyAsync(s.h_merkle + off, s.mlayers[li], c->mlens[li] * 32, hipMemcpyDeviceToHost, s.st));
            off += c->mlens[li] * 32;
        }
    }
    HIPCHK(hipMemcpyAsync(s.h_root, s.root, 32, hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipEventRecord(s.ev_done, s.st));
    s.out = out;
    s.busy = true;
    return ZGR_OK;
}

// Wait for slot s, copy its staged outputs into the caller's buffers, check the
// coset commitments (crates/encoder/src/amt/blob.rs:42-47).
int finish_blob(zgr_ctx* c, Slot& s) {
    Range range_("finish_blob");
    if (!s.busy) return ZGR_OK;
    s.busy = false;
    const zgr_encoded* out = s.out;
    s.out = nullptr;
    if (out->data)
        for (int cs = 0; cs < c->coset_n; cs++) {  // coset by coset, as the staging fills
            HIPCHK(hipEventSynchronize(s.ev_data[cs]));
            host_copy(out->data + (size_t)cs * c->n * 32, s.h_data + (size_t)cs * c->n * 32, c->n * 32);
        }
    HIPCHK(hipEventSynchronize(s.ev_done));
    const size_t slots = out_slots(c), T = tree_pts(c);
    const uint8_t* pts = s.h_pts;
    for (int cs = 0; cs < c->coset_n; cs++) {
        const uint8_t* base = pts + (size_t)cs * slots * 64;
        if (out->proof_tree) memcpy(out->proof_tree + cs * T * 64, base, T * 64);
        if (out->comm_tree) memcpy(out->comm_tree + cs * T * 64, base + T * 64, T * 64);
        if (out->high) memcpy(out->high + cs * 64, pts + (2 * T + 1) * 64, 64);  // coset 0 (see issue_blob)
    }
    if (out->commitment) memcpy(out->commitment, pts + 2 * T * 64, 64);
    if (out->merkle) memcpy(out->merkle, s.h_merkle, merkle_top_nodes(c) * 32);
    if (out->file_root) memcpy(out->file_root, s.h_root, 32);
    for (int cs = 1; cs < c->coset_n; cs++)
        if (memcmp(pts + ((size_t)cs * slots + 2 * T) * 64, pts + 2 * T * 64, 64))
            return fail(ZGR_ERR_INCONSISTENT, "coset %d commitment differs from coset 0", cs);
    return ZGR_OK;
}

// Abandon whatever the slots hold: wait for the device, then mark both slots free so
// that no later call copies their staging into output buffers of a call that has
// returned (the slot's `out` points into the caller's array of that call).
void drop_slots(zgr_ctx* c) {
    (void)sync_all(c);
    for (auto& s : c->slots) {
        s.busy = false;
        s.out = nullptr;
    }
}

// encode nblobs blobs already in device memory, two in flight
int encode_pipeline(zgr_ctx* c, int nblobs, const uint8_t* const* raws, const size_t* lens,
                    const zgr_encoded* outs) {
    for (int cs = 0; cs < c->coset_n; cs++) CHECK(ensure_params(c, cs));
    if (c->slots[0].busy || c->slots[1].busy) drop_slots(c);  // never inherit another call's slot
    int first_err = ZGR_OK;
    for (int i = 0; i < nblobs; i++) {
        Slot& s = c->slots[i & 1];
        int rc = finish_blob(c, s);  // the blob issued two steps ago
        if (rc && !first_err) first_err = rc;
        rc = issue_blob(c, s, raws[i], lens[i], &outs[i]);
        if (rc) {  // the other slot may still be in flight into this call's outputs
            drop_slots(c);
            return rc;
        }
    }
    for (int k = 0; k < 2; k++) {
        Slot& s = c->slots[(nblobs + k) & 1];
        int rc = finish_blob(c, s);
        if (rc && !first_err) first_err = rc;
    }
    int rc = sync_all(c);
    if (rc) {
        drop_slots(c);
        return rc;
    }
    return first_err;
}

// ---------------------------------------------------------------------------
// serving: continuous batching over the two blob slots
// ---------------------------------------------------------------------------

// first-fit allocator over one device buffer (the raw-input staging arena)
struct Arena {
    uint8_t* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length, coalesced

    static size_t round(size_t n) { return (n + 255) & ~(size_t)255; }
    bool alloc(size_t n, size_t* off) {
        n = round(n ? n : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            size_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = round(n ? n : 1);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    uint64_t id = 0;
    const uint8_t* raw = nullptr;
    size_t len = 0;
    const zgr_encoded* out = nullptr;
    size_t off = 0;  // raw bytes' place in the staging arena
};

}  // namespace

// Serving over one context's two blob slots.  Three threads:
//   worker    admits queued requests (staging-arena footprint, in-flight cap), and
//             issues the next admitted one into each free slot: the raw bytes' upload
//             on the slot's own stream, then the encode, then a host callback;
//   callback  (HIP runtime thread, after the slot's last copy into pinned staging) marks
//             the slot done and wakes the worker -- no polling;
//   finisher  copies a done slot's staged outputs into the request's buffers and
//             completes its ticket, off the issuing thread; the slot is free after that.
// The context's four streams (two coset lanes, two slots) are all there is: an extra
// upload stream would share a hardware queue with a lane (four queues per process) and
// wait behind its bucket kernels.
struct zgr_server {
    zgr_ctx* c = nullptr;
    Arena arena;
    int max_inflight = 1;
    std::mutex m;
    std::condition_variable cv_done, cv_work, cv_finish;
    std::deque<Request*> queue;   // submitted, waiting for admission
    std::deque<Request*> staged;  // admitted (arena space reserved), waiting for a free slot
    Request* in_slot[2] = {nullptr, nullptr};  // issued, until its outputs are copied out
    bool gpu_done[2] = {false, false};         // set by the slot's host callback
    std::deque<int> to_finish;                 // slots for the finisher
    struct CbArg {
        zgr_server* sv;
        int slot;
    } cb[2];
    // ticket -> final status and, for a failure, the error message of the thread that
    // failed it (zgr_last_error is per thread), until polled
    std::unordered_map<uint64_t, std::pair<int, std::string>> done;
    std::unordered_map<uint64_t, bool> pending;  // submitted tickets not yet completed
    uint64_t next_id = 1, n_completed = 0;
    int n_admitted = 0;
    bool stop = false;
    bool worker_exited = false;  // set once the worker has completed every request
    std::thread worker, finisher;
};

namespace {

// complete a request (caller holds sv->m): its ticket, the arena, the in-flight count
void server_complete(zgr_server* sv, Request* r, int st, const std::string& msg) {
    sv->arena.release(r->off, r->len);
    sv->n_admitted--;
    sv->n_completed++;
    sv->done[r->id] = {st, st ? msg : std::string()};
    sv->pending.erase(r->id);
    delete r;
}

// host callback queued after a slot's work: no HIP calls here
void server_slot_done(void* arg) {
    auto* a = (zgr_server::CbArg*)arg;
    {
        std::lock_guard<std::mutex> lk(a->sv->m);
        a->sv->gpu_done[a->slot] = true;
    }
    a->sv->cv_work.notify_all();
}

void server_finisher(zgr_server* sv) {
    zgr_ctx* c = sv->c;
    (void)hipSetDevice(c->device);
    std::unique_lock<std::mutex> lk(sv->m);
    while (true) {
        sv->cv_finish.wait(lk, [&] { return !sv->to_finish.empty() || sv->worker_exited; });
        if (sv->to_finish.empty()) break;
        const int k = sv->to_finish.front();
        sv->to_finish.pop_front();
        Request* r = sv->in_slot[k];
        lk.unlock();
        const int st = finish_blob(c, c->slots[k]);  // host copies of the staged outputs
        const std::string msg = st ? g_err : std::string();
        lk.lock();
        server_complete(sv, r, st, msg);
        sv->in_slot[k] = nullptr;  // the slot's staging may be written again
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();
    }
}

void server_loop(zgr_server* sv) {
    zgr_ctx* c = sv->c;
    std::lock_guard<std::mutex> own(c->mu);  // the server owns its context
    (void)hipSetDevice(c->device);
    std::unique_lock<std::mutex> lk(sv->m);
    while (true) {
        // admission: by staging-arena footprint and in-flight count
        while (!sv->queue.empty() && sv->n_admitted < sv->max_inflight) {
            Request* r = sv->queue.front();
            if (!sv->arena.alloc(r->len, &r->off)) break;
            sv->queue.pop_front();
            sv->n_admitted++;
            sv->staged.push_back(r);
        }
        // slots whose work is done go to the finisher
        for (int k = 0; k < 2; k++)
            if (sv->gpu_done[k]) {
                sv->gpu_done[k] = false;
                sv->to_finish.push_back(k);
                sv->cv_finish.notify_all();
            }
        // issue into free slots
        for (int k = 0; k < 2 && !sv->staged.empty(); k++) {
            if (sv->in_slot[k]) continue;
            Request* r = sv->staged.front();
            sv->staged.pop_front();
            sv->in_slot[k] = r;
            lk.unlock();  // no HIP call under the server lock (the callback takes it)
            Slot& s = c->slots[k];
            int st = ZGR_OK;
            if (r->len && hipMemcpyAsync(sv->arena.base + r->off, r->raw, r->len, hipMemcpyHostToDevice, s.st) != hipSuccess)
                st = fail(ZGR_ERR_DEVICE, "upload of a request's raw bytes failed");
            if (!st) st = issue_blob(c, s, sv->arena.base + r->off, r->len, r->out);
            if (!st && hipLaunchHostFunc(s.st, server_slot_done, &sv->cb[k]) != hipSuccess)
                st = fail(ZGR_ERR_DEVICE, "hipLaunchHostFunc failed");
            if (st) {  // could not be issued: complete it with the error, slot free again
                s.busy = false;
                s.out = nullptr;
                (void)hipStreamSynchronize(s.st);
            }
            const std::string msg = st ? g_err : std::string();
            lk.lock();
            if (st) {
                sv->in_slot[k] = nullptr;
                server_complete(sv, r, st, msg);
                sv->cv_done.notify_all();
            }
        }
        const bool busy = sv->in_slot[0] || sv->in_slot[1] || !sv->staged.empty() || !sv->to_finish.empty();
        if (sv->stop && sv->queue.empty() && !busy) break;
        // sleep until a submission, a slot callback or a finished copy-out wakes us (the
        // timeout only guards against a lost wake-up)
        const bool can_issue = !sv->staged.empty() && (!sv->in_slot[0] || !sv->in_slot[1]);
        if (!can_issue && !sv->gpu_done[0] && !sv->gpu_done[1]) sv->cv_work.wait_for(lk, std::chrono::milliseconds(50));
    }
    lk.unlock();
    (void)sync_all(c);
}

void destroy_streams(zgr_ctx* c) {
    for (auto& l : c->lanes)
        if (l.st) (void)hipStreamDestroy(l.st);
    if (c->vlane.st) (void)hipStreamDestroy(c->vlane.st);
    for (auto& s : c->slots) {
        if (s.st) (void)hipStreamDestroy(s.st);
        for (hipEvent_t e : {s.ev_pre, s.ev_done, s.ev_coset[0], s.ev_coset[1], s.ev_coset[2], s.ev_data[0],
                             s.ev_data[1], s.ev_data[2]})
            if (e) (void)hipEventDestroy(e);
    }
}

// ---------------------------------------------------------------------------
// public-parameter pipeline (params_pp.h, params_io.h)
// ---------------------------------------------------------------------------
size_t van_points(const zgr_ctx* c) { return ((size_t)2 << c->h) - 2; }

// fixed-base tables of the G1 and G2 generators, built once per context
int ensure_gen_tables(zgr_ctx* c, hipStream_t st) {
    if (!c->g1tab) {
        DALLOC(c->g1tab, 32 * 256 * 64);
        LAUNCH(c, st, "params_table", k_fixed_base_table, blocks(32 * 256, 64), 64, 0, c->g1tab);
    }
    if (!c->g2tab) {
        DALLOC(c->g2tab, 32 * 256 * G2_WORDS * 4);
        LAUNCH(c, st, "params_table", k_g2_fixed_base_table, blocks(32 * 256, 64), 64, 0, c->g2tab);
    }
    return ZGR_OK;
}

int ensure_van(zgr_ctx* c, int coset) {
    if (!c->van[coset]) DALLOC(c->van[coset], van_points(c) * G2_WORDS * 4);
    if (!c->vg2[coset]) DALLOC(c->vg2[coset], 2 * G2_WORDS * 4);
    return ZGR_OK;
}

// AMTParams::from_pp's G1 half for one coset (generate.rs:152-195, 204-232) from
// canonical powers already in device memory: coset scaling, the EC-FFTs of gen_basis /
// gen_quotients and enact, straight into the resident layout, then the shifted-copy
// tables.  The caller holds the context lock and has synchronised the streams.
int pp_bases(zgr_ctx* c, int coset, const uint32_t* g1pp, const uint32_t* high_g1pp) {
    CHECK(alloc_bases(c, coset));
    c->loaded[coset] = false;
    hipStream_t st = c->lanes[0].st;
    const size_t n = c->n;
    const int log_n = c->log_n;
    uint32_t *gp = nullptr, *hp = nullptr, *pts = nullptr, *w = nullptr;
    auto body = [&]() -> int {
        DALLOC(gp, n * X29_WORDS * 4);
        DALLOC(hp, n * X29_WORDS * 4);
        DALLOC(pts, n * X29_WORDS * 4);
        DALLOC(w, 32);
        LAUNCH(c, st, "params_pp", k_pp_coset_w, 1, 64, 0, c->consts, coset, w);
        // coset-scaled powers (generate.rs:152-170; coset 0 is unscaled), times 1 / N: the
        // size_inv that every transform below ends with, moved to its inputs
        LAUNCH(c, st, "params_pp", k_pp_scale, blocks(n, 64), 64, 0, g1pp, w, coset != 0, c->consts, (uint32_t)n,
               gp);
        LAUNCH(c, st, "params_pp", k_pp_scale, blocks(n, 64), 64, 0, high_g1pp, w, coset != 0, c->consts,
               (uint32_t)n, hp);
        const uint32_t half_n = (uint32_t)(n / 2);
        auto transform = [&](const uint32_t* src, uint32_t m, bool inverse, int line) -> int {
            LAUNCH(c, st, "params_pp", k_pp_load, blocks(n, 256), 256, 0, src, pts, log_n, m);
            const uint32_t* tw = inverse ? c->tw_inv : c->tw;
            for (int s = 1; s <= log_n; s++)
                LAUNCH(c, st, "params_pp_fft", k_ec_dit_stage, blocks(half_n, 64), 64, 0, pts, tw, s, log_n - s, half_n);
            // both directions end with size_inv (the iDFT's own, gen_quotients' explicit
            // one), already applied to gp and hp
            LAUNCH(c, st, "params_pp", k_pp_store, blocks(n, 64), 64, 0, pts, c->consts, 0, log_n,
                   c->bases[coset] + (size_t)line * n * 16);
            return ZGR_OK;
        };
        for (int d = 1; d <= c->h; d++) CHECK(transform(gp, (uint32_t)(n >> d), false, d - 1));  // quotients
        CHECK(transform(gp, 0, true, c->h));                                                    // basis
        CHECK(transform(hp, 0, true, c->h + 1));                                                // high_basis
        CHECK(build_fb(c, coset));
        return sync_all(c);
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)gp, (void*)hp, (void*)pts, (void*)w})
        if (p) (void)hipFree(p);
    if (rc == ZGR_OK) c->loaded[coset] = true;
    return rc;
}

// AMTParams::from_pp's G2 half for one coset: vanishes[d - 1] = enact(gen_vanishes(g2pp',
// d)) for d = 1..h (generate.rs:177-179, 234-263), g2 = g2pp[0], high_g2 (both unscaled:
// w^0 = 1), from canonical powers in device memory.  Caller holds the lock.
int pp_vanishes(zgr_ctx* c, int coset, const uint32_t* g2pp, const uint32_t* high_g2) {
    CHECK(ensure_van(c, coset));
    c->van_ready[coset] = false;
    hipStream_t st = c->lanes[0].st;
    uint32_t *work = nullptr, *w = nullptr;
    auto body = [&]() -> int {
        DALLOC(work, ((size_t)G2X_WORDS * 4) << (c->h + 1));  // every level: 2^(h+1) - 2 points
        DALLOC(w, 32);
        LAUNCH(c, st, "params_vanish", k_rec_clear, 1, 64, 0, c->flag);
        LAUNCH(c, st, "params_vanish", k_pp_coset_w, 1, 64, 0, c->consts, coset, w);
        // the h levels' transforms side by side: one launch per stage for all of them
        const size_t pts = ((size_t)2 << c->h) - 2;
        LAUNCH(c, st, "params_vanish", k_vanish_load, blocks(pts, 64), 64, 0, g2pp, w, coset != 0, c->log_n, c->h,
               work);
        for (int s = 1; s <= c->h; s++)
            LAUNCH(c, st, "params_vanish", k_vanish_stage, blocks(((size_t)1 << c->h) - ((size_t)1 << (s - 1)), 64), 64,
                   0, work, c->tw_inv, s, c->log_n, c->h);
        LAUNCH(c, st, "params_vanish", k_vanish_store, blocks(pts, 64), 64, 0, work, c->h, c->van[coset], c->flag);
        HIPCHK(hipMemcpyAsync(c->vg2[coset], g2pp, G2_WORDS * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(c->vg2[coset] + G2_WORDS, high_g2, G2_WORDS * 4, hipMemcpyDeviceToDevice, st));
        int bad = 0;
        HIPCHK(hipMemcpyAsync(&bad, c->flag, sizeof(int), hipMemcpyDeviceToHost, st));
        CHECK(sync_all(c));
        if (bad) return fail(ZGR_ERR_INVALID_ARG, "a vanish point is infinity (tau' on a level's domain)");
        return ZGR_OK;
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)work, (void*)w})
        if (p) (void)hipFree(p);
    if (rc == ZGR_OK) c->van_ready[coset] = true;
    return rc;
}

// little-endian writer of a parameter file image (counts bytes when out is null)
struct ImageWriter {
    uint8_t* out;
    size_t cap, len = 0;
    void bytes(const void* p, size_t n) {
        if (out && len + n <= cap) memcpy(out + len, p, n);
        len += n;
    }
    void u64(uint64_t v) { bytes(&v, 8); }
    void u8(uint8_t v) { bytes(&v, 1); }
    uint8_t* at(size_t n) {  // the next n bytes, to be filled by a device copy
        uint8_t* p = out && len + n <= cap ? out + len : nullptr;
        len += n;
        return p;
    }
};

}  // namespace

extern "C" {

const char* zgr_last_error(void) { return g_err.c_str(); }

int zgr_ctx_create(int device, int log_col, int log_row, int coset_n, zgr_ctx** out) {
    if (!out) return fail(ZGR_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (log_col < 3 || log_row < 1 || log_col + log_row > NTT_MAX_TWO_PASS || coset_n < 1 || coset_n > 3 ||
        (1 << (log_col + log_row)) * coset_n > (1 << 28))
        return fail(ZGR_ERR_INVALID_ARG, "unsupported geometry log_col=%d log_row=%d coset_n=%d", log_col,
                    log_row, coset_n);
    if (log_col > 15) return fail(ZGR_ERR_INVALID_ARG, "log_col > 15 not supported");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(ZGR_ERR_INVALID_ARG, "device %d not present", device);
    HIPCHK(hipSetDevice(device));
    uint8_t* scratch_dev = nullptr;  // set-up of zgr_merkle_noalloc's scratch on this device
    CHECK(noalloc_scratch(device, &scratch_dev));
    zgr_ctx* c = new zgr_ctx();
    c->device = device;
    c->log_col = log_col;
    c->log_row = log_row;
    c->log_n = log_col + log_row;
    c->h = log_row;
    c->coset_n = coset_n;
    c->nlines = log_row + 2;
    c->n = (size_t)1 << c->log_n;
    c->nb = (size_t)1 << c->h;
    c->bs = c->n / c->nb;
    // fixed-base tables unless the chunk is too large for their entry format, or
    // ZGR_FIXED_BASE=0 asks for the windowed MSM (same results, far less memory)
    const char* genv = getenv("ZGR_GRAPHS");
    c->graphs = !(genv && genv[0] == '0');
    const char* fbenv = getenv("ZGR_FIXED_BASE");
    c->fb = (c->log_n - c->h) <= FB_MAX_LOG_BS && c->nlines <= 16 && !(fbenv && fbenv[0] == '0');
    if (c->fb) {
        c->wide_lines = fb_wide_lines(c);
        // about 256 dense steps per lane in each wide sums wave (ZGR_WIDE_SEGS overrides):
        // at 2^20, 1024 segments per line measured 2-3 % faster than 512 in the tuner's
        // dense timings, 512 1 % faster than 256 or 768, and 128 12 % slower (the long
        // waves' tail)
        const char* segenv = getenv("ZGR_WIDE_SEGS");
        const size_t nseg = segenv ? (size_t)strtoul(segenv, nullptr, 10) : ((size_t)FW_W * c->n) / (FB_LANES * 256);
        c->fw_nseg = (uint32_t)std::max<size_t>(1, std::min<size_t>(nseg, 1u << 20));
        c->fw_nseg_max = 2 * c->fw_nseg;  // the tuner's finer-segment plan
    }
    auto init = [&]() -> int {
        int max_shm = 0;
        HIPCHK(hipDeviceGetAttribute(&max_shm, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        size_t need = accum_lds_bytes(c->bs);
        if (need + 1024 > (size_t)max_shm)
            return fail(ZGR_ERR_INVALID_ARG, "bucket kernel needs %zu B of LDS, device allows %d", need, max_shm);
        HIPCHK(hipFuncSetAttribute((const void*)k_msm_accum, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
        const size_t ntt_lds = ntt_lds_bytes(std::min(c->log_n, NTT_MAX_LOG));
        if (ntt_lds > (size_t)max_shm)
            return fail(ZGR_ERR_INVALID_ARG, "NTT pass needs %zu B of LDS, device allows %d", ntt_lds, max_shm);
        HIPCHK(hipFuncSetAttribute((const void*)k_ntt_pass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ntt_lds));
        if (FW_BUCKETS * 4 <= max_shm) {  // zgr_msm's sort (otherwise it keeps the windowed kernels)
            for (const void* k : {(const void*)k_fwp_count, (const void*)k_fwp_scatter})
                HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, FW_BUCKETS * 4));
            c->pm.lds_ok = true;
        }
        if (!c->wide_lines.empty()) {  // the wide sort's LDS counters / cursors
            if (FW_BUCKETS * 4 > max_shm) c->wide_lines.clear();
            else
                for (const void* k : {(const void*)k_fw_count, (const void*)k_fw_scatter})
                    HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, FW_BUCKETS * 4));
        }
        DALLOC(c->consts, K_COUNT * 32);
        DALLOC(c->tw, c->n / 2 * 32);
        DALLOC(c->tw_inv, c->n / 2 * 32);
        for (int d = 0; d < 2; d++) {
            DALLOC(c->pw_lo[d], ((size_t)32) << ntt_pw_bits(c->log_n));
            DALLOC(c->pw_hi[d], ((size_t)32) << (c->log_n - ntt_pw_bits(c->log_n)));
        }
        DALLOC(c->stw, 9 * NTT_STAGE_TW * 4);
        DALLOC(c->stw_inv, 9 * NTT_STAGE_TW * 4);
        DALLOC(c->ninv29, (c->log_n + 1) * 32);
        DALLOC(c->flag, sizeof(int));
        // Merkle layer lengths: leaves, then halving down to coset_n roots
        for (size_t len = (size_t)coset_n * c->n / 8;; len /= 2) {
            c->mlens.push_back(len);
            if (len <= (size_t)coset_n || (len & 1)) break;
        }
        if ((int)c->mlens.size() <= c->h) return fail(ZGR_ERR_INVALID_ARG, "merkle tree too shallow");
        for (auto& l : c->lanes) {
            HIPCHK(hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking));
            DALLOC(l.y, c->n * 32);
            DALLOC(l.digits, c->n * MSM_W);
            DALLOC(l.entries, std::max(c->n * MSM_W * 4, c->fb ? c->n * FB_W * sizeof(uint2) : 0));
            DALLOC(l.counts, c->nb * MSM_W * 4);
            DALLOC(l.heap, (size_t)c->nlines * MSM_W * 2 * c->nb * X29_WORDS * 4);
            if (c->fb) {
                DALLOC(l.bstart, c->nb * FB_BSTART * 4);
                DALLOC(l.bkt, (size_t)c->nlines * c->nb * FB_REGION * X29_WORDS * 4);
                for (int gl = 1; gl <= FB_GL_MAX && gl <= c->h; gl++) {
                    DALLOC(l.gent[gl], c->n * FB_W * sizeof(uint2));
                    DALLOC(l.gcnt[gl], (c->nb >> gl) * 4);
                    DALLOC(l.gbst[gl], (c->nb >> gl) * FB_BSTART * 4);
                }
                if (!c->wide_lines.empty()) {
                    const size_t nw = c->wide_lines.size(), nblk = c->nb >> FW_BLK_LOG;
                    size_t groups = 0;
                    for (int L : c->wide_lines) groups += c->nb >> fw_group_log(c, L);
                    DALLOC(l.fw_cnt, nblk * FW_BUCKETS * 4);
                    DALLOC(l.fw_off, nblk * FW_BUCKETS * 4);
                    DALLOC(l.fw_tile, (2 * FW_TILES + 1) * 4);
                    DALLOC(l.fw_ent, c->n * FW_W * sizeof(uint2));
                    DALLOC(l.fw_seg, nw * (c->fw_nseg_max + 1) * 4);
                    DALLOC(l.fw_bkt, groups * FW_BUCKETS * X29_WORDS * 4);
                    DALLOC(l.fw_hand, nw * c->fw_nseg_max * 2 * FB_LANES * X29_WORDS * 4);
                    DALLOC(l.fw_part, groups * FW_RSEGS * X29_WORDS * 4);
                }
            }
        }
        if (c->fb) CHECK(build_work_lists(c));
        // lane 0's point buffer also serves multiple_multiexp (nlines * nb points)
        size_t opts = out_slots(c) * coset_n;
        if (opts < (size_t)c->nlines * c->nb) opts = (size_t)c->nlines * c->nb;
        for (auto& s : c->slots) {
            HIPCHK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&s.ev_pre, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&s.ev_done, hipEventDisableTiming));
            for (auto& e : s.ev_coset) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            for (auto& e : s.ev_data) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            DALLOC(s.raw, 31 * c->n);
            DALLOC(s.pts, c->n * 32);
            DALLOC(s.coeff, c->n * 32);
            DALLOC(s.outpts, opts * 64);
            DALLOC(s.data, (size_t)coset_n * c->n * 32);
            DALLOC(s.root, 32);
            for (size_t len : c->mlens) {
                uint32_t* p = nullptr;
                DALLOC(p, len * 32);
                s.mlayers.push_back(p);
            }
            HALLOC(s.h_pts, out_slots(c) * coset_n * 64);
            HALLOC(s.h_data, (size_t)coset_n * c->n * 32);
            HALLOC(s.h_merkle, merkle_top_nodes(c) * 32);
            HALLOC(s.h_root, 32);
        }
        hipStream_t st = c->lanes[0].st;
        LAUNCH(c, st, "setup", k_setup_consts, 1, 1, 0, c->consts, c->log_n, coset_n);
        LAUNCH(c, st, "setup", k_powers, blocks(c->n / 2, 256), 256, 0, c->tw, c->consts, (int)K_OMEGA,
               (uint32_t)(c->n / 2));
        LAUNCH(c, st, "setup", k_powers, blocks(c->n / 2, 256), 256, 0, c->tw_inv, c->consts, (int)K_OMEGA_INV,
               (uint32_t)(c->n / 2));
        LAUNCH(c, st, "setup", k_ntt_tables, blocks(c->n / 2, 256), 256, 0, c->tw, c->tw_inv, c->consts, c->log_n,
               c->pw_lo[0], c->pw_lo[1], c->pw_hi[0], c->pw_hi[1], c->stw, c->stw_inv, c->ninv29);
        for (int cs = 1; cs < coset_n; cs++) {
            DALLOC(c->scale[cs], c->n * 32);
            LAUNCH(c, st, "setup", k_coset_scale_table, blocks(c->n, 256), 256, 0, c->scale[cs], c->consts, cs,
                   c->log_n);
            LAUNCH(c, st, "setup", k_to261_inplace, blocks(c->n, 256), 256, 0, c->scale[cs], (uint32_t)c->n);
        }
        return sync_all(c);
    };
    int rc = init();
    if (rc) {
        zgr_ctx_destroy(c);
        return rc;
    }
    c->times.clear();
    *out = c;
    return ZGR_OK;
}

void zgr_ctx_destroy(zgr_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (auto& l : c->lanes)
        if (l.st) (void)hipStreamSynchronize(l.st);
    if (c->vlane.st) (void)hipStreamSynchronize(c->vlane.st);
    for (auto& s : c->slots)
        if (s.st) (void)hipStreamSynchronize(s.st);
    void* ptrs[] = {c->consts, c->tw,      c->tw_inv, c->pw_lo[0], c->pw_lo[1], c->pw_hi[0],
                    c->pw_hi[1], c->stw,   c->stw_inv, c->ninv29,  c->flag};
    for (auto& pl : c->plans)
        for (void* p : {(void*)pl.work[0], (void*)pl.work[1], (void*)pl.rwork[0], (void*)pl.rwork[1]})
            if (p) (void)hipFree(p);
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (void* p : c->rec_buf)
        if (p) (void)hipFree(p);
    for (void* p : {(void*)c->pm.ent, (void*)c->pm.cnt, (void*)c->pm.off, (void*)c->pm.tile, (void*)c->pm.seg,
                    (void*)c->pm.bkt, (void*)c->pm.part, (void*)c->pm.hand, (void*)c->pm.heap, (void*)c->pm.work,
                    (void*)c->pm.halves, (void*)c->pm.table, (void*)c->pm.bases})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)c->pt_g1, (void*)c->pt_high, (void*)c->pt_g2, (void*)c->pt_high_g2, (void*)c->g1tab,
                    (void*)c->g2tab})
        if (p) (void)hipFree(p);
    for (int i = 0; i < 3; i++) {
        if (c->van[i]) (void)hipFree(c->van[i]);
        if (c->vg2[i]) (void)hipFree(c->vg2[i]);
        if (c->scale[i]) (void)hipFree(c->scale[i]);
        if (c->bases[i]) (void)hipFree(c->bases[i]);
        if (c->fbtab[i]) (void)hipFree(c->fbtab[i]);
        if (c->fwtab[i]) (void)hipFree(c->fwtab[i]);
    }
    for (void* p : {(void*)c->vlane.y, (void*)c->vlane.entries, (void*)c->vlane.counts, (void*)c->vlane.heap,
                    (void*)c->vlane.bstart, (void*)c->vlane.bkt})
        if (p) (void)hipFree(p);
    for (auto& l : c->lanes) {
        void* lp[] = {l.y,      l.digits, l.entries, l.counts,  l.heap,    l.bstart,  l.bkt,
                      l.fw_cnt, l.fw_off, l.fw_tile, l.fw_ent, l.fw_seg, l.fw_bkt, l.fw_hand, l.fw_part};
        for (void* p : lp)
            if (p) (void)hipFree(p);
        for (int gl = 0; gl <= FB_GL_MAX; gl++)
            for (void* p : {(void*)l.gent[gl], (void*)l.gcnt[gl], (void*)l.gbst[gl]})
                if (p) (void)hipFree(p);
    }
    for (auto& s : c->slots) {
        void* sp[] = {s.raw, s.pts, s.coeff, s.outpts, s.data, s.root};
        for (void* p : sp)
            if (p) (void)hipFree(p);
        for (auto* p : s.mlayers) (void)hipFree(p);
        void* hp[] = {s.h_pts, s.h_data, s.h_merkle, s.h_root};
        for (void* p : hp)
            if (p) (void)hipHostFree(p);
    }
    for (auto& t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto& g : c->msm_graphs) (void)hipGraphExecDestroy(g.second);
    destroy_streams(c);
    delete c;
}

int zgr_params_load(zgr_ctx* c, int coset, const void* bases_mont, size_t nbytes) {
    Range range_("zgr_params_load");
    if (!c || !bases_mont) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    size_t want = (size_t)c->nlines * c->n * 64;
    if (nbytes != want) return fail(ZGR_ERR_INVALID_ARG, "params size %zu, expected %zu", nbytes, want);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));  // blobs still in flight read the bases and tables about to be replaced
    CHECK(alloc_bases(c, coset));
    c->loaded[coset] = false;
    HIPCHK(hipMemcpyAsync(c->bases[coset], bases_mont, want, hipMemcpyHostToDevice, c->lanes[0].st));
    CHECK(build_fb(c, coset));
    CHECK(sync_all(c));
    c->loaded[coset] = true;
    return ZGR_OK;
}

int zgr_params_load_bamt(zgr_ctx* c, int coset, const uint8_t* bamt, size_t nbytes) {
    Range range_("zgr_params_load_bamt");
    if (!c || !bamt) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    // write_amt_params layout: "bamt", u8 degree, u8 sub_degree, g2, high_g2 (ark
    // uncompressed G2), basis, quotients[0..sub), vanishes[d] (2^(d+1) G2 of 128 B),
    // high_basis; G1 points are x_mont || y_mont (64 B)
    if (nbytes < 6 || memcmp(bamt, "bamt", 4)) return fail(ZGR_ERR_INVALID_ARG, "Incorrect format");
    const int degree = bamt[4], sub = bamt[5];
    if (degree != c->log_n || sub != c->h)
        return fail(ZGR_ERR_INVALID_ARG, "bamt degree %d / sub-degree %d, context has %d / %d", degree, sub,
                    c->log_n, c->h);
    const size_t line = c->n * 64;
    const size_t vanish = van_points(c) * 128;
    // g2 / high_g2: ark uncompressed G2 affine = x.c0 x.c1 y.c0 y.c1, 32 B each, the
    // infinity / sign flags in the spare top bits of the last byte
    const size_t g2 = 128;
    const size_t want = 6 + 2 * g2 + (size_t)(sub + 2) * line + vanish;
    if (nbytes != want)
        return fail(ZGR_ERR_INVALID_ARG, "bamt length %zu, expected %zu for degree %d", nbytes, want, degree);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));  // blobs still in flight read the bases and tables about to be replaced
    CHECK(alloc_bases(c, coset));
    CHECK(ensure_van(c, coset));
    c->loaded[coset] = false;
    c->van_ready[coset] = false;
    hipStream_t st = c->lanes[0].st;
    uint8_t* dev = (uint8_t*)c->bases[coset];
    // the verifier half: g2 and high_g2 as canonical points (the flag bits cleared; the
    // infinity flag gives zeros), the vanishes verbatim (Montgomery, the device layout)
    uint8_t vg2[2][128];
    for (int k = 0; k < 2; k++) {
        memcpy(vg2[k], bamt + 6 + k * g2, g2);
        const bool inf = vg2[k][127] & 0x40;
        vg2[k][127] &= 0x3f;
        if (inf) memset(vg2[k], 0, 128);
    }
    HIPCHK(hipMemcpy(c->vg2[coset], vg2, sizeof(vg2), hipMemcpyHostToDevice));
    const uint8_t* p = bamt + 6 + 2 * g2;
    HIPCHK(hipMemcpyAsync(dev + (size_t)c->h * line, p, line, hipMemcpyHostToDevice, st));  // basis
    p += line;
    for (int d = 0; d < sub; d++, p += line)  // quotients
        HIPCHK(hipMemcpyAsync(dev + (size_t)d * line, p, line, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->van[coset], p, vanish, hipMemcpyHostToDevice, st));  // vanishes
    p += vanish;
    HIPCHK(hipMemcpyAsync(dev + (size_t)(c->h + 1) * line, p, line, hipMemcpyHostToDevice, st));  // high_basis
    CHECK(build_fb(c, coset));
    CHECK(sync_all(c));
    c->loaded[coset] = true;
    c->van_ready[coset] = true;
    return ZGR_OK;
}

int zgr_params_from_pp(zgr_ctx* c, int coset, const uint8_t* g1pp, const uint8_t* high_g1pp, size_t n_points) {
    Range range_("zgr_params_from_pp");
    if (!c || !g1pp || !high_g1pp) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    if (n_points != c->n)
        return fail(ZGR_ERR_INVALID_ARG, "%zu powers of tau, the context's domain has %zu", n_points, c->n);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));  // blobs still in flight read the bases and tables about to be replaced
    hipStream_t st = c->lanes[0].st;
    uint32_t *din = nullptr, *hin = nullptr;
    auto body = [&]() -> int {
        DALLOC(din, c->n * 64);
        DALLOC(hin, c->n * 64);
        HIPCHK(hipMemcpyAsync(din, g1pp, c->n * 64, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(hin, high_g1pp, c->n * 64, hipMemcpyHostToDevice, st));
        c->van_ready[coset] = false;  // the G2 half is not given here
        return pp_bases(c, coset, din, hin);
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)din, (void*)hin})
        if (p) (void)hipFree(p);
    return rc;
}

int zgr_params_generate(zgr_ctx* c, const uint8_t tau[32]) {
    Range range_("zgr_params_generate");
    if (!c || !tau) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));
    hipStream_t st = c->lanes[0].st;
    uint32_t *tau_d = nullptr, *pc = nullptr;
    auto body = [&]() -> int {
        DALLOC(tau_d, 32);
        DALLOC(pc, (3 + c->h) * 32);
        HIPCHK(hipMemcpyAsync(tau_d, tau, 32, hipMemcpyHostToDevice, st));
        LAUNCH(c, st, "params_gen", k_rec_clear, 1, 64, 0, c->flag);
        CHECK(ensure_gen_tables(c, st));
        for (int cs = 0; cs < c->coset_n; cs++) {
            CHECK(alloc_bases(c, cs));
            CHECK(ensure_van(c, cs));
            c->loaded[cs] = c->van_ready[cs] = false;
            LAUNCH(c, st, "params_consts", k_params_consts, 1, 1, 0, tau_d, c->consts, cs, c->log_n, c->h, pc);
            LAUNCH(c, st, "params_gen", k_params_gen, blocks(c->n, 64), 64, 0, c->bases[cs], c->consts, pc, c->g1tab,
                   c->log_n, c->h, c->flag);
            // the verifier half in closed form: vanishes, g2 = [1] G2, high_g2 = [tau^(3N)] G2
            LAUNCH(c, st, "params_vanish", k_vanish_closed, blocks(van_points(c), 64), 64, 0, pc, c->consts, c->log_n,
                   c->h, c->g2tab, c->van[cs], c->flag);
            LAUNCH(c, st, "params_gen", k_power_tau, 1, 64, 0, tau_d, (uint64_t)0, 1u, (const uint32_t*)nullptr,
                   c->g2tab, (uint32_t*)nullptr, c->vg2[cs], c->flag);
            LAUNCH(c, st, "params_gen", k_power_tau, 1, 64, 0, tau_d, (uint64_t)3 * c->n, 1u, (const uint32_t*)nullptr,
                   c->g2tab, (uint32_t*)nullptr, c->vg2[cs] + G2_WORDS, c->flag);
        }
        int bad = 0;
        HIPCHK(hipMemcpyAsync(&bad, c->flag, sizeof(int), hipMemcpyDeviceToHost, st));
        CHECK(sync_all(c));
        if (bad) return fail(ZGR_ERR_INVALID_ARG, "tau hits the evaluation domain (or a base is infinity)");
        for (int cs = 0; cs < c->coset_n; cs++) CHECK(build_fb(c, cs));
        CHECK(sync_all(c));
        for (int cs = 0; cs < c->coset_n; cs++) c->loaded[cs] = c->van_ready[cs] = true;
        return ZGR_OK;
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)tau_d, (void*)pc})
        if (p) (void)hipFree(p);
    return rc;
}

int zgr_power_tau_setup(zgr_ctx* c, const uint8_t tau[32]) {
    Range range_("zgr_power_tau_setup");
    if (!c || !tau) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));
    hipStream_t st = c->lanes[0].st;
    const size_t n = c->n;
    uint32_t* tau_d = nullptr;
    auto body = [&]() -> int {
        c->pt_ready = false;
        if (!c->pt_g1) DALLOC(c->pt_g1, n * 64);
        if (!c->pt_high) DALLOC(c->pt_high, n * 64);
        if (!c->pt_g2) DALLOC(c->pt_g2, n * G2_WORDS * 4);
        if (!c->pt_high_g2) DALLOC(c->pt_high_g2, G2_WORDS * 4);
        DALLOC(tau_d, 32);
        HIPCHK(hipMemcpyAsync(tau_d, tau, 32, hipMemcpyHostToDevice, st));
        LAUNCH(c, st, "power_tau", k_rec_clear, 1, 64, 0, c->flag);
        CHECK(ensure_gen_tables(c, st));
        // H = 2^(log_n + 2) - 2^log_n = 3N (power_tau.rs:66-67)
        LAUNCH(c, st, "power_tau", k_power_tau, blocks(n, 64), 64, 0, tau_d, (uint64_t)0, (uint32_t)n, c->g1tab,
               c->g2tab, c->pt_g1, c->pt_g2, c->flag);
        LAUNCH(c, st, "power_tau", k_power_tau, blocks(n, 64), 64, 0, tau_d, (uint64_t)3 * n, (uint32_t)n, c->g1tab,
               (const uint32_t*)nullptr, c->pt_high, (uint32_t*)nullptr, c->flag);
        LAUNCH(c, st, "power_tau", k_power_tau, 1, 64, 0, tau_d, (uint64_t)3 * n, 1u, (const uint32_t*)nullptr, c->g2tab,
               (uint32_t*)nullptr, c->pt_high_g2, c->flag);
        int bad = 0;
        HIPCHK(hipMemcpyAsync(&bad, c->flag, sizeof(int), hipMemcpyDeviceToHost, st));
        CHECK(sync_all(c));
        if (bad) return fail(ZGR_ERR_INVALID_ARG, "tau = 0 gives points at infinity");
        c->pt_ready = true;
        return ZGR_OK;
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    if (tau_d) (void)hipFree(tau_d);
    return rc;
}

int zgr_power_tau_load(zgr_ctx* c, const uint8_t* g1pp, const uint8_t* g2pp, const uint8_t* high_g1pp,
                       const uint8_t* high_g2, size_t n_points) {
    Range range_("zgr_power_tau_load");
    if (!c || !g1pp || !g2pp || !high_g1pp || !high_g2) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (n_points != c->n)
        return fail(ZGR_ERR_INVALID_ARG, "%zu powers of tau, the context's domain has %zu", n_points, c->n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));
    const size_t n = c->n;
    c->pt_ready = false;
    if (!c->pt_g1) DALLOC(c->pt_g1, n * 64);
    if (!c->pt_high) DALLOC(c->pt_high, n * 64);
    if (!c->pt_g2) DALLOC(c->pt_g2, n * G2_WORDS * 4);
    if (!c->pt_high_g2) DALLOC(c->pt_high_g2, G2_WORDS * 4);
    HIPCHK(hipMemcpy(c->pt_g1, g1pp, n * 64, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->pt_high, high_g1pp, n * 64, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->pt_g2, g2pp, n * G2_WORDS * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->pt_high_g2, high_g2, G2_WORDS * 4, hipMemcpyHostToDevice));
    c->pt_ready = true;
    return ZGR_OK;
}

int zgr_ppot_decode(zgr_ctx* c, int kind, const uint8_t* src, size_t count, uint8_t* out, size_t* first_bad) {
    Range range_("zgr_ppot_decode");
    if (!c || (count && (!src || !out))) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (kind < ZGR_PPOT_G1_UNCOMPRESSED || kind > ZGR_PPOT_G2_COMPRESSED)
        return fail(ZGR_ERR_INVALID_ARG, "point kind %d is not one of ZGR_PPOT_*", kind);
    if (first_bad) *first_bad = count;
    if (count == 0) return ZGR_OK;
    const bool g2 = kind >= ZGR_PPOT_G2_UNCOMPRESSED, comp = (kind & 1) != 0;
    const size_t in_b = (g2 ? 64 : 32) * (comp ? 1 : 2), out_b = g2 ? 128 : 64;
    static const char* const reason[] = {"accepted",
                                         "unexpected compression mode (0x80 on an uncompressed point)",
                                         "unexpected information (infinity flag with other bits set)",
                                         "point at infinity",
                                         "coordinate not below p",
                                         "not on the curve",
                                         "not in the order-r subgroup"};
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));
    hipStream_t st = c->lanes[0].st;
    const size_t chunk = std::min(count, (size_t)1 << 20);
    uint32_t *d_src = nullptr, *d_out = nullptr, *d_st = nullptr;
    std::vector<uint32_t> status(chunk);
    auto body = [&]() -> int {
        DALLOC(d_src, chunk * in_b);
        DALLOC(d_out, chunk * out_b);
        DALLOC(d_st, chunk * 4);
        for (size_t off = 0; off < count; off += chunk) {
            const size_t n = std::min(chunk, count - off);
            HIPCHK(hipMemcpyAsync(d_src, src + off * in_b, n * in_b, hipMemcpyHostToDevice, st));
            const uint32_t t = g2 ? PPOT_G2_THREADS : PPOT_G1_THREADS;
            if (!g2 && !comp) LAUNCH(c, st, "ppot", k_ppot_g1u, blocks(n, t), t, 0, d_src, n, d_out, d_st);
            if (!g2 && comp) LAUNCH(c, st, "ppot", k_ppot_g1c, blocks(n, t), t, 0, d_src, n, d_out, d_st);
            if (g2 && !comp) LAUNCH(c, st, "ppot", k_ppot_g2u, blocks(n, t), t, 0, d_src, n, d_out, d_st);
            if (g2 && comp) LAUNCH(c, st, "ppot", k_ppot_g2c, blocks(n, t), t, 0, d_src, n, d_out, d_st);
            HIPCHK(hipMemcpyAsync(out + off * out_b, d_out, n * out_b, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(status.data(), d_st, n * 4, hipMemcpyDeviceToHost, st));
            CHECK(sync_all(c));
            for (size_t j = 0; j < n; j++)
                if (status[j]) {
                    if (first_bad) *first_bad = off + j;
                    return fail(ZGR_ERR_INVALID_ARG, "point %zu: %s", off + j,
                                status[j] < sizeof(reason) / sizeof(reason[0]) ? reason[status[j]] : "bad status");
                }
        }
        return ZGR_OK;
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)d_src, (void*)d_out, (void*)d_st})
        if (p) (void)hipFree(p);
    return rc;
}

int zgr_power_tau_download(zgr_ctx* c, uint8_t* g1pp, uint8_t* g2pp, uint8_t* high_g1pp, uint8_t* high_g2) {
    if (!c) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->pt_ready) return fail(ZGR_ERR_NOT_LOADED, "no powers of tau in this context");
    HIPCHK(hipSetDevice(c->device));
    const size_t n = c->n;
    if (g1pp) HIPCHK(hipMemcpy(g1pp, c->pt_g1, n * 64, hipMemcpyDeviceToHost));
    if (g2pp) HIPCHK(hipMemcpy(g2pp, c->pt_g2, n * G2_WORDS * 4, hipMemcpyDeviceToHost));
    if (high_g1pp) HIPCHK(hipMemcpy(high_g1pp, c->pt_high, n * 64, hipMemcpyDeviceToHost));
    if (high_g2) HIPCHK(hipMemcpy(high_g2, c->pt_high_g2, G2_WORDS * 4, hipMemcpyDeviceToHost));
    return ZGR_OK;
}

int zgr_params_from_power_tau(zgr_ctx* c, int coset) {
    Range range_("zgr_params_from_power_tau");
    if (!c) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    if (!c->pt_ready) return fail(ZGR_ERR_NOT_LOADED, "no powers of tau in this context");
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));  // blobs still in flight read the bases and tables about to be replaced
    CHECK(pp_bases(c, coset, c->pt_g1, c->pt_high));
    return pp_vanishes(c, coset, c->pt_g2, c->pt_high_g2);
}

int zgr_params_g2(zgr_ctx* c, int coset, uint8_t* vanishes, uint8_t* g2, uint8_t* high_g2) {
    Range range_("zgr_params_g2");
    if (!c) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    if (!c->van_ready[coset]) return fail(ZGR_ERR_NOT_LOADED, "no verifier params (vanishes) for coset %d", coset);
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->lanes[0].st;
    uint32_t* tmp = nullptr;
    auto body = [&]() -> int {
        if (vanishes) {  // Montgomery in the device -> canonical
            DALLOC(tmp, van_points(c) * G2_WORDS * 4);
            LAUNCH(c, st, "params_io", k_fq_convert, blocks(van_points(c) * 4, 256), 256, 0, c->van[coset], tmp,
                   (uint64_t)van_points(c) * 4, 0);
            HIPCHK(hipMemcpyAsync(vanishes, tmp, van_points(c) * G2_WORDS * 4, hipMemcpyDeviceToHost, st));
        }
        if (g2) HIPCHK(hipMemcpyAsync(g2, c->vg2[coset], G2_WORDS * 4, hipMemcpyDeviceToHost, st));
        if (high_g2) HIPCHK(hipMemcpyAsync(high_g2, c->vg2[coset] + G2_WORDS, G2_WORDS * 4, hipMemcpyDeviceToHost, st));
        return sync_all(c);
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    if (tmp) (void)hipFree(tmp);
    return rc;
}

int zgr_write_params_file(zgr_ctx* c, int kind, int coset, int sub_depth, uint8_t* out, size_t cap, size_t* len) {
    Range range_("zgr_write_params_file");
    if (!c || !len) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    const bool amt = kind == ZGR_FILE_AMT_MONT || kind == ZGR_FILE_AMT || kind == ZGR_FILE_AMT_VERIFY;
    const bool ptau = kind == ZGR_FILE_PTAU || kind == ZGR_FILE_PTAU_MONT;
    if (!amt && !ptau) return fail(ZGR_ERR_INVALID_ARG, "unknown parameter file kind %d", kind);
    if (amt && (coset < 0 || coset >= c->coset_n)) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    if (kind == ZGR_FILE_AMT_VERIFY && (sub_depth < 0 || sub_depth > c->h))
        return fail(ZGR_ERR_INVALID_ARG, "verify depth %d outside [0, %d]", sub_depth, c->h);
    if (amt && !c->loaded[coset]) return fail(ZGR_ERR_NOT_LOADED, "params for coset %d are not resident", coset);
    if (amt && !c->van_ready[coset])
        return fail(ZGR_ERR_NOT_LOADED, "no verifier params (vanishes, g2) for coset %d", coset);
    if (ptau && !c->pt_ready) return fail(ZGR_ERR_NOT_LOADED, "no powers of tau in this context");
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));
    const size_t n = c->n;
    const int h = c->h;
    // one pass counts the image, a second one fills it (when out has room)
    ImageWriter w{out, cap};
    hipStream_t st = c->lanes[0].st;
    uint32_t* stage = nullptr;  // device staging of one converted line (at most n G2 points)
    auto emit = [&](bool fill) -> int {
        w.len = 0;
        // ark form of n_pts points from device memory (g2: G2 else G1; mont: Montgomery input)
        auto ark = [&](const uint32_t* src, size_t n_pts, bool g2, bool mont, bool compressed) -> int {
            const size_t bytes = n_pts * (g2 ? 64 : 32) * (compressed ? 1 : 2);
            uint8_t* dst = w.at(bytes);
            if (!fill || !dst) return ZGR_OK;
            if (g2)
                LAUNCH(c, st, "params_io", k_g2_ark, blocks(n_pts, 256), 256, 0, src, (uint32_t)n_pts, (int)mont,
                       (int)compressed, stage);
            else
                LAUNCH(c, st, "params_io", k_g1_ark, blocks(n_pts, 256), 256, 0, src, (uint32_t)n_pts, (int)mont,
                       (int)compressed, stage);
            HIPCHK(hipMemcpyAsync(dst, stage, bytes, hipMemcpyDeviceToHost, st));
            return sync_all(c);  // the staging buffer is reused by the next section
        };
        // raw copy (Montgomery in the device layout), or canonical -> Montgomery first
        auto raw = [&](const uint32_t* src, size_t bytes, bool to_mont) -> int {
            uint8_t* dst = w.at(bytes);
            if (!fill || !dst) return ZGR_OK;
            if (to_mont) {
                LAUNCH(c, st, "params_io", k_fq_convert, blocks(bytes / 32, 256), 256, 0, src, stage,
                       (uint64_t)(bytes / 32), 1);
                src = stage;
            }
            HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
            return sync_all(c);
        };
        const uint32_t* line = amt ? c->bases[coset] : nullptr;
        auto lineptr = [&](int L) { return line + (size_t)L * n * 16; };
        const uint32_t* van = amt ? c->van[coset] : nullptr;
        const uint32_t* vg2 = amt ? c->vg2[coset] : nullptr;
        switch (kind) {
            case ZGR_FILE_AMT_MONT:  // write_amt_params (fast_serde_bn254.rs:20-55)
                w.bytes("bamt", 4);
                w.u8((uint8_t)c->log_n);
                w.u8((uint8_t)h);
                CHECK(ark(vg2, 1, true, false, false));             // g2
                CHECK(ark(vg2 + G2_WORDS, 1, true, false, false));  // high_g2
                CHECK(raw(lineptr(h), n * 64, false));              // basis
                for (int d = 0; d < h; d++) CHECK(raw(lineptr(d), n * 64, false));  // quotients
                CHECK(raw(van, van_points(c) * 128, false));        // vanishes
                CHECK(raw(lineptr(h + 1), n * 64, false));          // high_basis
                break;
            case ZGR_FILE_AMT:  // serialize_uncompressed of AMTParams (prove_params/serde.rs)
                w.u64(n);
                CHECK(ark(lineptr(h), n, false, true, false));
                w.u64((uint64_t)h);
                for (int d = 0; d < h; d++) {
                    w.u64(n);
                    CHECK(ark(lineptr(d), n, false, true, false));
                }
                w.u64((uint64_t)h);
                for (int d = 1; d <= h; d++) {
                    w.u64((uint64_t)1 << d);
                    CHECK(ark(van + G2_WORDS * (((size_t)1 << d) - 2), (size_t)1 << d, true, true, false));
                }
                CHECK(ark(vg2, 1, true, false, false));
                w.u64(n);
                CHECK(ark(lineptr(h + 1), n, false, true, false));
                CHECK(ark(vg2 + G2_WORDS, 1, true, false, false));
                break;
            case ZGR_FILE_AMT_VERIFY:  // serialize_uncompressed of AMTVerifyParams (verify_params.rs:22-28, 81-98)
                w.u64(n);
                CHECK(ark(lineptr(h), n, false, true, false));
                w.u64((uint64_t)sub_depth);
                for (int d = 1; d <= sub_depth; d++) {
                    w.u64((uint64_t)1 << d);
                    CHECK(ark(van + G2_WORDS * (((size_t)1 << d) - 2), (size_t)1 << d, true, true, false));
                }
                CHECK(ark(vg2, 1, true, false, false));
                CHECK(ark(vg2 + G2_WORDS, 1, true, false, false));
                break;
            case ZGR_FILE_PTAU:  // serialize_compressed of PowerTau (power_tau.rs:20-26)
                w.u64(n);
                CHECK(ark(c->pt_g1, n, false, false, true));
                w.u64(n);
                CHECK(ark(c->pt_g2, n, true, false, true));
                w.u64(n);
                CHECK(ark(c->pt_high, n, false, false, true));
                CHECK(ark(c->pt_high_g2, 1, true, false, true));
                break;
            case ZGR_FILE_PTAU_MONT:  // write_power_tau (fast_serde_bn254.rs:90-113)
                w.bytes("ptau", 4);
                w.u8((uint8_t)c->log_n);
                CHECK(ark(c->pt_high_g2, 1, true, false, false));
                CHECK(raw(c->pt_g1, n * 64, true));
                CHECK(raw(c->pt_g2, n * G2_WORDS * 4, true));
                CHECK(raw(c->pt_high, n * 64, true));
                break;
        }
        return ZGR_OK;
    };
    CHECK(emit(false));
    *len = w.len;
    if (!out) return ZGR_OK;
    if (cap < w.len) return fail(ZGR_ERR_INVALID_ARG, "output holds %zu bytes, the file image needs %zu", cap, w.len);
    int rc = dmalloc((void**)&stage, n * G2_WORDS * 4);
    if (!rc) rc = emit(true);
    (void)hipStreamSynchronize(st);
    if (stage) (void)hipFree(stage);
    return rc;
}

int zgr_params_download(zgr_ctx* c, int coset, void* out, size_t nbytes) {
    if (!c || !out) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, coset));
    size_t want = (size_t)c->nlines * c->n * 64;
    if (nbytes != want) return fail(ZGR_ERR_INVALID_ARG, "params size %zu, expected %zu", nbytes, want);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(out, c->bases[coset], want, hipMemcpyDeviceToHost, c->lanes[0].st));
    return sync_all(c);
}

int zgr_multiple_multiexp(zgr_ctx* c, int coset, const uint8_t* ri_mont, size_t n, uint8_t* out_points) {
    Range range_("zgr_multiple_multiexp");
    if (!c || !ri_mont || !out_points) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, coset));
    if (n != c->n) return fail(ZGR_ERR_INVALID_ARG, "scalar count %zu, expected %zu", n, c->n);
    HIPCHK(hipSetDevice(c->device));
    Lane& ln = c->lanes[0];
    uint32_t* opts = c->slots[0].outpts;
    HIPCHK(hipMemcpyAsync(ln.y, ri_mont, n * 32, hipMemcpyHostToDevice, ln.st));
    CHECK(msm_lines(c, ln, coset, ln.y, 0, c->nlines, nullptr));  // every chunk of every line
    uint32_t total = (uint32_t)(c->nlines * c->nb);
    LAUNCH(c, ln.st, "msm_leaves_out", k_msm_leaves_out, blocks(total, 64), 64, 0, ln.heap, c->h, heap_windows(c),
           c->nlines, opts);
    HIPCHK(hipMemcpyAsync(out_points, opts, (size_t)total * 64, hipMemcpyDeviceToHost, ln.st));
    return sync_all(c);
}

int zgr_amt_prove(zgr_ctx* c, int coset, const uint8_t* ri_mont, size_t n, uint8_t* commitment,
                  uint8_t* comm_tree, uint8_t* proof_tree, uint8_t* high) {
    Range range_("zgr_amt_prove");
    if (!c || !ri_mont) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, coset));
    if (n != c->n) return fail(ZGR_ERR_INVALID_ARG, "scalar count %zu, expected %zu", n, c->n);
    HIPCHK(hipSetDevice(c->device));
    Lane& ln = c->lanes[0];
    uint32_t* opts = c->slots[0].outpts;
    HIPCHK(hipMemcpyAsync(ln.y, ri_mont, n * 32, hipMemcpyHostToDevice, ln.st));
    CHECK(msm_lines(c, ln, coset, ln.y, 0, c->nlines, plan_for(c, ZGR_DENSE)));
    CHECK(amt_outputs(c, ln, opts, c->nlines, plan_for(c, ZGR_DENSE)));
    const size_t slots = out_slots(c), T = tree_pts(c);
    std::vector<uint8_t> pts(slots * 64);
    HIPCHK(hipMemcpyAsync(pts.data(), opts, pts.size(), hipMemcpyDeviceToHost, ln.st));
    CHECK(sync_all(c));
    if (proof_tree) memcpy(proof_tree, pts.data(), T * 64);
    if (comm_tree) memcpy(comm_tree, pts.data() + T * 64, T * 64);
    if (commitment) memcpy(commitment, pts.data() + 2 * T * 64, 64);
    if (high) memcpy(high, pts.data() + (2 * T + 1) * 64, 64);
    return ZGR_OK;
}

int zgr_encode_device(zgr_ctx* c, const void* raw_dev, size_t len, const zgr_encoded* out) {
    Range range_("zgr_encode_device");
    if (!c || !out || (!raw_dev && len)) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(check_len(c, len));
    HIPCHK(hipSetDevice(c->device));
    const uint8_t* r = raw_dev ? (const uint8_t*)raw_dev : c->slots[0].raw;
    return encode_pipeline(c, 1, &r, &len, out);
}

int zgr_encode(zgr_ctx* c, const uint8_t* raw, size_t len, const zgr_encoded* out) {
    Range range_("zgr_encode");
    if (!c || !out || (!raw && len)) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(check_len(c, len));
    HIPCHK(hipSetDevice(c->device));
    Slot& s = c->slots[0];
    if (len) HIPCHK(hipMemcpyAsync(s.raw, raw, len, hipMemcpyHostToDevice, s.st));
    const uint8_t* r = s.raw;
    return encode_pipeline(c, 1, &r, &len, out);
}

int zgr_encode_batch_device(zgr_ctx* c, int nblobs, const void* const* raw_dev, const size_t* lens,
                            const zgr_encoded* outs) {
    Range range_("zgr_encode_batch_device");
    if (!c || nblobs < 0 || (nblobs && (!raw_dev || !lens || !outs)))
        return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    for (int i = 0; i < nblobs; i++) {
        if (!raw_dev[i] && lens[i]) return fail(ZGR_ERR_INVALID_ARG, "blob %d: null data", i);
        CHECK(check_len(c, lens[i]));
    }
    HIPCHK(hipSetDevice(c->device));
    std::vector<const uint8_t*> r(nblobs);
    for (int i = 0; i < nblobs; i++) r[i] = raw_dev[i] ? (const uint8_t*)raw_dev[i] : c->slots[0].raw;
    return encode_pipeline(c, nblobs, r.data(), lens, outs);
}

int zgr_server_create(zgr_ctx* c, size_t hbm_budget, int max_inflight, zgr_server** out) {
    if (!c || !out || max_inflight < 1) return fail(ZGR_ERR_INVALID_ARG, "bad argument");
    *out = nullptr;
    for (int cs = 0; cs < c->coset_n; cs++) CHECK(ensure_params(c, cs));
    HIPCHK(hipSetDevice(c->device));
    zgr_server* sv = new zgr_server();
    sv->c = c;
    sv->max_inflight = max_inflight;
    sv->arena.size = Arena::round(hbm_budget ? hbm_budget : 1);
    for (int k = 0; k < 2; k++) sv->cb[k] = {sv, k};
    int rc = dmalloc((void**)&sv->arena.base, sv->arena.size);
    if (rc) {
        delete sv;
        return rc;
    }
    sv->arena.free_[0] = sv->arena.size;
    sv->worker = std::thread(server_loop, sv);
    sv->finisher = std::thread(server_finisher, sv);
    *out = sv;
    return ZGR_OK;
}

void zgr_server_destroy(zgr_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->m);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();  // every request completed
    {
        std::lock_guard<std::mutex> lk(sv->m);
        sv->worker_exited = true;
    }
    sv->cv_finish.notify_all();
    if (sv->finisher.joinable()) sv->finisher.join();
    (void)hipSetDevice(sv->c->device);
    if (sv->arena.base) (void)hipFree(sv->arena.base);
    delete sv;
}

int zgr_server_submit(zgr_server* sv, const uint8_t* raw, size_t len, const zgr_encoded* out, uint64_t* ticket) {
    if (!sv || !out || !ticket || (!raw && len)) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    CHECK(check_len(sv->c, len));
    if (Arena::round(len ? len : 1) > sv->arena.size)
        return fail(ZGR_ERR_OOM, "request of %zu B can never fit the %zu B staging arena", len, sv->arena.size);
    Request* r = new Request();
    r->raw = raw;
    r->len = len;
    r->out = out;
    {
        std::lock_guard<std::mutex> lk(sv->m);
        if (sv->stop) {
            delete r;
            return fail(ZGR_ERR_INVALID_ARG, "server is stopping");
        }
        r->id = sv->next_id++;
        sv->pending[r->id] = true;
        sv->queue.push_back(r);
        *ticket = r->id;
    }
    sv->cv_work.notify_all();
    return ZGR_OK;
}

int zgr_server_poll(zgr_server* sv, uint64_t ticket, int timeout_ms) {
    if (!sv) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(sv->m);
    if (!sv->pending.count(ticket) && !sv->done.count(ticket))
        return fail(ZGR_ERR_INVALID_ARG, "unknown or already collected ticket %llu", (unsigned long long)ticket);
    auto ready = [&] { return sv->done.count(ticket) != 0; };
    if (timeout_ms < 0)
        sv->cv_done.wait(lk, ready);
    else if (!sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return 1;
    const auto res = sv->done[ticket];
    sv->done.erase(ticket);
    if (res.first != ZGR_OK) g_err = res.second;  // the worker's message, on the polling thread
    return res.first;
}

int zgr_server_cancel(zgr_server* sv, uint64_t ticket) {
    if (!sv) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    {
        std::lock_guard<std::mutex> lk(sv->m);
        if (!sv->pending.count(ticket))
            return fail(ZGR_ERR_INVALID_ARG, "ticket %llu is not pending", (unsigned long long)ticket);
        auto take = [&](std::deque<Request*>& q) -> Request* {
            for (auto it = q.begin(); it != q.end(); ++it)
                if ((*it)->id == ticket) {
                    Request* r = *it;
                    q.erase(it);
                    return r;
                }
            return nullptr;
        };
        if (Request* r = take(sv->queue)) {  // not admitted: no arena space, no in-flight count
            sv->n_completed++;
            sv->done[r->id] = {ZGR_ERR_CANCELLED, "cancelled"};
            sv->pending.erase(r->id);
            delete r;
        } else if (Request* r2 = take(sv->staged)) {  // admitted, not issued
            server_complete(sv, r2, ZGR_ERR_CANCELLED, "cancelled");
        } else {
            return fail(ZGR_ERR_INVALID_ARG, "ticket %llu is already on the GPU", (unsigned long long)ticket);
        }
    }
    sv->cv_done.notify_all();
    sv->cv_work.notify_all();
    return ZGR_OK;
}

int zgr_test_inject_fault(zgr_ctx* c, int site, int countdown) {
    if (!c) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (site < 0 || site >= ZGR_FAULT_SITES) return fail(ZGR_ERR_INVALID_ARG, "fault site %d out of range", site);
    if (countdown < 0) return fail(ZGR_ERR_INVALID_ARG, "negative countdown");
    c->fault[site].store(countdown);
    return ZGR_OK;
}

void zgr_server_stats(zgr_server* sv, uint64_t* queued, uint64_t* admitted, uint64_t* completed) {
    if (!sv) return;
    std::lock_guard<std::mutex> lk(sv->m);
    if (queued) *queued = sv->queue.size();
    if (admitted) *admitted = (uint64_t)sv->n_admitted;
    if (completed) *completed = sv->n_completed;
}

int zgr_ntt(zgr_ctx* c, uint8_t* data_mont, int log_n, int batch, int inverse) {
    Range range_("zgr_ntt");
    if (!c || (!data_mont && batch) || batch < 0) return fail(ZGR_ERR_INVALID_ARG, "bad argument");
    if (log_n < 1 || log_n > c->log_n)
        return fail(ZGR_ERR_INVALID_ARG, "log_n %d outside [1, %d] of this context", log_n, c->log_n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->lanes[0].st;
    uint32_t* buf = c->slots[0].pts;   // rows staged here (n_ctx points)
    uint32_t* tmp = c->slots[0].coeff;
    const size_t row = (size_t)32 << log_n;
    for (int r = 0; r < batch;) {
        const uint32_t rows = std::min<uint32_t>(ntt_chunk_rows(c, log_n, batch - r), c->n >> log_n);
        HIPCHK(hipMemcpyAsync(buf, data_mont + r * row, rows * row, hipMemcpyHostToDevice, st));
        CHECK(ntt_natural(c, st, buf, tmp, log_n, rows, inverse != 0));
        HIPCHK(hipMemcpyAsync(data_mont + r * row, buf, rows * row, hipMemcpyDeviceToHost, st));
        r += rows;
    }
    return sync_all(c);
}

int zgr_ntt_device(zgr_ctx* c, void* data_dev, int log_n, int batch, int inverse) {
    Range range_("zgr_ntt_device");
    if (!c || (!data_dev && batch) || batch < 0) return fail(ZGR_ERR_INVALID_ARG, "bad argument");
    if (log_n < 1 || log_n > c->log_n)
        return fail(ZGR_ERR_INVALID_ARG, "log_n %d outside [1, %d] of this context", log_n, c->log_n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->lanes[0].st;
    uint32_t* data = (uint32_t*)data_dev;
    for (int r = 0; r < batch;) {
        const uint32_t rows = ntt_chunk_rows(c, log_n, batch - r);
        CHECK(ntt_natural(c, st, data + ((size_t)r << log_n) * 8, c->slots[0].coeff, log_n, rows, inverse != 0));
        r += rows;
    }
    return sync_all(c);
}

int zgr_coset_extend(zgr_ctx* c, const uint8_t* data_mont, int coset, uint8_t* out_mont) {
    Range range_("zgr_coset_extend");
    if (!c || !data_mont || !out_mont) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    Lane& ln = c->lanes[0];
    uint32_t* coeff = c->slots[0].coeff;
    HIPCHK(hipMemcpyAsync(coeff, data_mont, c->n * 32, hipMemcpyHostToDevice, ln.st));
    if (coset == 0) {
        HIPCHK(hipMemcpyAsync(ln.y, coeff, c->n * 32, hipMemcpyDeviceToDevice, ln.st));
    } else {
        CHECK(ntt_dif(c, ln.st, coeff, coeff, c->log_n, 1, true));
        CHECK(ntt_dit(c, ln.st, coeff, ln.y, c->log_n, 1, false, c->scale[coset]));
    }
    HIPCHK(hipMemcpyAsync(out_mont, ln.y, c->n * 32, hipMemcpyDeviceToHost, ln.st));
    return sync_all(c);
}

// Windowed MSM sum_i s_i P_i over 2^log_np device-resident points (param layout) with
// the scalars (Montgomery) already in ln.y; the canonical affine result goes to out.
// zgr_msm through the wide sort and bucket kernels: each scalar split by GLV into two
// 127-bit halves over P_i and phi(P_i) (k_pm_split, k_pm_table), the PM_W windows of the
// 2n halves as virtual scalars of one digit (k_fwp_count / k_fwp_scatter), one group per
// window, bucket-segment sums waves (k_fb_sums), the groups' segmented reductions and
// combine, then Horner over the windows (k_pm_horner).  Returns 1 without launching
// anything when the device cannot give the sort its LDS; the caller then runs the
// windowed kernels.
constexpr uint32_t PM_MAX_SEGS = 8192;
constexpr int PM_RB = 512;                     // buckets per reduction item
constexpr int PM_RSEGS = FW_BUCKETS / PM_RB;  // reduction partials per window

int msm_plain(zgr_ctx* c, Lane& ln, const uint32_t* dbases, int log_np, uint32_t* out) {
    if (!c->pm.lds_ok) return 1;
    const uint32_t np = 1u << log_np, nv = 2 * np;  // points, virtual points (P, phi(P))
    const uint32_t ns = std::min(nv, 1u << 16);      // virtual points per sort block
    const uint32_t bpw = nv / ns;                    // sort blocks per window
    int sh = 0;
    while ((1u << sh) < bpw) sh++;
    const uint32_t nblk = PM_W * bpw;
    // bucket segments (one sums wave each): one line of work, so more and shorter waves
    // than an AMT line's 1024 to fill the device; ZGR_MSM_SEGS overrides
    uint32_t nseg = std::max(1u, std::min(PM_MAX_SEGS, (uint32_t)((uint64_t)PM_W * nv / 8192)));
    if (const char* e = getenv("ZGR_MSM_SEGS")) {
        const long v = atol(e);
        if (v >= 1) nseg = std::min(PM_MAX_SEGS, (uint32_t)v);
    }
    auto& pm = c->pm;
    {  // sized for the context's largest MSM, on the first call; each buffer allocated
       // only if missing, so a call after a failed allocation completes the set
        auto need = [](auto*& p, size_t bytes) { return p ? ZGR_OK : dmalloc((void**)&p, bytes); };
        const size_t blk_max = (size_t)PM_W * std::max<size_t>(1, (2 * c->n) >> 16);
        CHECK(need(pm.ent, (size_t)PM_W * 2 * c->n * sizeof(uint2)));
        CHECK(need(pm.cnt, blk_max * FW_BUCKETS * 4));
        CHECK(need(pm.off, blk_max * FW_BUCKETS * 4));
        CHECK(need(pm.tile, (2 * FW_TILES + 1) * 4));
        CHECK(need(pm.seg, (PM_MAX_SEGS + 1) * 4));
        CHECK(need(pm.bkt, (size_t)PM_W * FW_BUCKETS * X29_WORDS * 4));
        CHECK(need(pm.part, (size_t)PM_W * PM_RSEGS * X29_WORDS * 4));
        CHECK(need(pm.hand, (size_t)PM_MAX_SEGS * 2 * FB_LANES * X29_WORDS * 4));
        CHECK(need(pm.heap, (size_t)(PM_W + 1) * X29_WORDS * 4));
        CHECK(need(pm.work, (PM_MAX_SEGS + PM_W * PM_RSEGS) * sizeof(uint2)));
        CHECK(need(pm.halves, (size_t)PM_SPLIT_WORDS * c->n * 4));
        CHECK(need(pm.table, (size_t)2 * c->n * 64));
    }
    FbWide wd{};
    wd.ent = pm.ent;
    wd.off = pm.off;
    wd.tot = pm.tile + 2 * FW_TILES;
    wd.seg = pm.seg;
    wd.table = pm.table;
    wd.bkt = pm.bkt;
    wd.hand = pm.hand;
    wd.part = pm.part;
    wd.nblk = nblk;
    wd.nw = 1;
    wd.segbase[0] = 0;
    wd.segbase[1] = nseg;
    wd.gbase[0] = 0;
    wd.gbase[1] = PM_W;
    wd.line[0] = 0;
    wd.sh[0] = (uint32_t)sh;
    // work lists: the sums waves (wide line 0, segment s, hand-off slot s), then the
    // reduction items (window S, segment q)
    pm.host_work.clear();
    for (uint32_t sg = 0; sg < nseg; sg++) pm.host_work.push_back(make_uint2(FB_WIDE | (sg << 8), sg));
    for (uint32_t S = 0; S < (uint32_t)PM_W; S++)
        for (uint32_t q = 0; q < (uint32_t)PM_RSEGS; q++)
            pm.host_work.push_back(make_uint2(FB_WIDE | (q << 4) | (S << 10), 0));
    HIPCHK(hipMemcpyAsync(pm.work, pm.host_work.data(), pm.host_work.size() * sizeof(uint2), hipMemcpyHostToDevice,
                          ln.st));
    const size_t shm = FW_BUCKETS * 4;
    const uint32_t nred = (uint32_t)(PM_W * PM_RSEGS);
    LAUNCH(c, ln.st, "msm_pm_table", k_pm_split, blocks(np, 256), 256, 0, ln.y, np, pm.halves);
    LAUNCH(c, ln.st, "msm_pm_table", k_pm_table, blocks(np, 256), 256, 0, dbases, np, pm.table);
    LAUNCH(c, ln.st, "msm_wide_sort", k_fwp_count, nblk, FW_SORT_T, shm, pm.halves, np, sh, ns, pm.cnt);
    LAUNCH(c, ln.st, "msm_wide_sort", k_fw_tile_sums, FW_TILES, FW_TILE * FW_PARTS, 0, pm.cnt, nblk, pm.tile);
    LAUNCH(c, ln.st, "msm_wide_sort", k_fw_tile_scan, 1, FW_TILES, 0, pm.tile, pm.tile + FW_TILES,
           pm.tile + 2 * FW_TILES);
    LAUNCH(c, ln.st, "msm_wide_sort", k_fw_offsets, FW_TILES, FW_TILE * FW_PARTS, 0, pm.cnt, nblk,
           pm.tile + FW_TILES, pm.off);
    LAUNCH(c, ln.st, "msm_wide_sort", k_fwp_scatter, nblk, FW_SORT_T, shm, pm.halves, np, sh, ns, pm.off, pm.ent);
    LAUNCH(c, ln.st, "msm_wide_sort", k_fw_segments, blocks(nseg + 1, 64), 64, 0, wd, pm.seg);
    LAUNCH(c, ln.st, "msm_accum", k_fb_sums, nseg, FB_LANES, 0, (const uint32_t*)pm.table, FbLevels{}, wd, pm.work,
           ln.bkt, c->log_n, c->h, 1);
    // the groups' node level: heap node (2^h >> gl) + S with h = gl is node 1 + S
    const int gl = sh + FW_BLK_LOG;
    LAUNCH(c, ln.st, "msm_reduce", (k_fb_reduce<64, PM_RB>), nred, FB_LANES, 0, FbLevels{}, wd, pm.work + nseg, nred,
           ln.bkt, pm.heap, gl, 1);
    LAUNCH(c, ln.st, "msm_combine", k_fw_combine<PM_RSEGS>, (uint32_t)PM_W, FB_LANES, 0, wd, pm.heap, gl);
    LAUNCH(c, ln.st, "msm_root", k_pm_horner, 1, 64, 0, pm.heap, out);
    return ZGR_OK;
}

int msm_windowed(zgr_ctx* c, Lane& ln, const uint32_t* dbases, int log_np, uint32_t* out) {
    const int log_bs = log_np < 10 ? log_np : 10;
    const int h = log_np - log_bs;
    const size_t np = (size_t)1 << log_np;
    LAUNCH(c, ln.st, "msm_digits", k_msm_digits, blocks(np, 256), 256, 0, ln.y, log_np, log_bs, ln.digits, 0);
    uint32_t bs = 1u << log_bs, nb = 1u << h;
    LAUNCH(c, ln.st, "msm_sort", k_msm_sort, nb * MSM_W, bs < 256 ? bs : 256u, 0, ln.digits, log_bs, ln.entries,
           ln.counts);
    LAUNCH(c, ln.st, "msm_accum", k_msm_accum, nb, ACC_THREADS, accum_lds_bytes(bs), dbases, ln.entries, ln.counts,
           ln.heap, log_np, h, 1);
    for (int d = h - 1; d >= 0; d--) {
        uint32_t total = (uint32_t)MSM_W << d;
        LAUNCH(c, ln.st, "msm_tree", k_msm_tree_level, blocks(total, 64), 64, 0, ln.heap, h, d, total, MSM_W,
               (uint64_t)0);
    }
    LAUNCH(c, ln.st, "msm_root", k_msm_root_out, 1, 64, 0, ln.heap, h, MSM_W, out);
    return ZGR_OK;
}

// sum_i s_i B_i over line L of a coset's resident bases, through that line of the
// shifted-copy tables: scalars in ln.y (Montgomery, in order), one bucket set per
// chunk (one wave each), the chunk sums summed by the tree kernel, canonical affine
// result to out.  Uses the lane's sort, bucket and heap buffers.
int msm_fb_line(zgr_ctx* c, Lane& ln, int coset, int L, uint32_t* out) {
    const int log_bs = c->log_n - c->h;
    FbLevels lv{};
    FbOut fo{};
    lv.ent[0] = fo.ent[0] = (uint2*)ln.entries;
    lv.cnt[0] = ln.counts;
    lv.bst[0] = ln.bstart;
    LAUNCH(c, ln.st, "msm_sort", k_fb_count, (uint32_t)c->nb, (uint32_t)c->bs, 0, ln.y, c->log_n, log_bs, 0, ln.counts,
           ln.bstart);
    LAUNCH(c, ln.st, "msm_scatter", k_fb_scatter, (uint32_t)c->nb, (uint32_t)c->bs, 0, ln.y, c->log_n, log_bs, 0,
           ln.bstart, 0, fo);
    const size_t cap = (size_t)FB_W << log_bs;
    const uint32_t* table = c->fbtab[coset] + ((size_t)L << c->h) * cap * 16;  // line L's chunks
    uint32_t* heap = ln.heap + (size_t)L * 2 * c->nb * X29_WORDS;               // line L's tree
    LAUNCH(c, ln.st, "msm_accum", k_fb_sums, (uint32_t)c->nb, FB_LANES, 0, table, lv, FbWide{},
           (const uint2*)nullptr, ln.bkt, c->log_n, c->h, 1);
    LAUNCH(c, ln.st, "msm_reduce", k_fb_reduce<64>, (uint32_t)c->nb, FB_LANES, 0, lv, FbWide{},
           (const uint2*)nullptr, (uint32_t)c->nb, ln.bkt, heap, c->h, 1);
    for (int d = c->h - 1; d >= 0; d--)
        LAUNCH(c, ln.st, "msm_tree", k_msm_tree_level, blocks((size_t)1 << d, 64), 64, 0, heap, c->h, d,
               (uint32_t)1 << d, 1, (uint64_t)0);
    LAUNCH(c, ln.st, "msm_root", k_msm_root_out, 1, 64, 0, heap, c->h, 1, out);
    return ZGR_OK;
}

int zgr_msm(zgr_ctx* c, const uint8_t* bases_mont, const uint8_t* scalars_mont, size_t n, uint8_t* out_point) {
    Range range_("zgr_msm");
    if (!c || !bases_mont || !scalars_mont || !out_point) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (n == 0 || n > c->n) return fail(ZGR_ERR_INVALID_ARG, "msm size %zu outside [1, %zu]", n, c->n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    int log_np = 0;
    while (((size_t)1 << log_np) < n) log_np++;
    const size_t np = (size_t)1 << log_np;
    const int log_bs = log_np < 10 ? log_np : 10;
    const int h = log_np - log_bs;
    Lane& ln = c->lanes[0];
    if (!c->pm.bases) CHECK(dmalloc((void**)&c->pm.bases, c->n * 64));  // kept for later calls
    uint32_t* dbases = c->pm.bases;
    auto body = [&]() -> int {
        // zero-padded scalars: the padding points are never read by the bucket kernel
        if (np > n) {
            HIPCHK(hipMemsetAsync(ln.y + n * 8, 0, (np - n) * 32, ln.st));
            HIPCHK(hipMemsetAsync(dbases + n * 16, 0, (np - n) * 64, ln.st));
        }
        HIPCHK(hipMemcpyAsync(ln.y, scalars_mont, n * 32, hipMemcpyHostToDevice, ln.st));
        HIPCHK(hipMemcpyAsync(dbases, bases_mont, n * 64, hipMemcpyHostToDevice, ln.st));
        LAUNCH(c, ln.st, "msm_inf", k_msm_clear_inf, blocks(n, 256), 256, 0, dbases, (uint32_t)n, ln.y);
        uint32_t* opts = c->slots[0].outpts;
        // ZGR_MSM=windowed keeps the per-chunk windowed kernels (same result)
        const char* mode = getenv("ZGR_MSM");
        int rc = 1;  // 1: the windowed kernels
        if (!(mode && !strcmp(mode, "windowed"))) rc = msm_plain(c, ln, dbases, log_np, opts);
        if (rc < 0) return rc;
        if (rc == 1) CHECK(msm_windowed(c, ln, dbases, log_np, opts));
        HIPCHK(hipMemcpyAsync(out_point, opts, 64, hipMemcpyDeviceToHost, ln.st));
        return sync_all(c);
    };
    int rc = body();
    (void)hipStreamSynchronize(ln.st);
    return rc;
}

int zgr_ecfft(zgr_ctx* c, const uint8_t* points, int log_n, int inverse, uint8_t* out) {
    Range range_("zgr_ecfft");
    if (!c || !points || !out) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (log_n < 1 || log_n > c->log_n)
        return fail(ZGR_ERR_INVALID_ARG, "log_n %d outside [1, %d] of this context", log_n, c->log_n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)1 << log_n;
    hipStream_t st = c->lanes[0].st;
    uint32_t *din = nullptr, *dpts = nullptr;
    auto body = [&]() -> int {
        DALLOC(din, n * 64);
        DALLOC(dpts, n * X29_WORDS * 4);
        HIPCHK(hipMemcpyAsync(din, points, n * 64, hipMemcpyHostToDevice, st));
        LAUNCH(c, st, "ecfft", k_ec_load_bitrev, blocks(n, 256), 256, 0, din, dpts, log_n);
        const uint32_t* tw = inverse ? c->tw_inv : c->tw;
        uint32_t half_n = (uint32_t)(n / 2);
        for (int s = 1; s <= log_n; s++)
            LAUNCH(c, st, "ecfft", k_ec_dit_stage, blocks(half_n, 64), 64, 0, dpts, tw, s, c->log_n - s, half_n);
        LAUNCH(c, st, "ecfft", k_ec_store, blocks(n, 64), 64, 0, dpts, c->consts, c->log_n - log_n, inverse ? 1 : 0,
               din, (uint32_t)n);
        HIPCHK(hipMemcpyAsync(out, din, n * 64, hipMemcpyDeviceToHost, st));
        return sync_all(c);
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)din, (void*)dpts})
        if (p) (void)hipFree(p);
    return rc;
}

int zgr_merkle(zgr_ctx* c, const uint8_t* data, size_t n_elems, int levels, uint8_t* out_layers,
               uint8_t* file_root) {
    Range range_("zgr_merkle");
    if (!c || !data || !file_root) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (n_elems == 0 || n_elems % 8) return fail(ZGR_ERR_INVALID_ARG, "n_elems %zu unsupported", n_elems);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->lanes[0].st;
    std::vector<uint32_t*> layers;
    std::vector<size_t> lens;
    auto body = [&]() -> int {
        uint32_t* dsrc = nullptr;
        DALLOC(dsrc, n_elems * 32);
        layers.push_back(dsrc);
        lens.push_back(n_elems);
        HIPCHK(hipMemcpyAsync(dsrc, data, n_elems * 32, hipMemcpyHostToDevice, st));
        size_t len = n_elems / 8;
        uint32_t* lv = nullptr;
        DALLOC(lv, len * 32);
        layers.push_back(lv);
        lens.push_back(len);
        LAUNCH(c, st, "merkle_leaves", k_merkle_leaves_flat, blocks(len, 128), 128, 0, dsrc, lv, (uint32_t)len);
        while (len > (size_t)c->coset_n && !(len & 1)) {
            len /= 2;
            uint32_t* nx = nullptr;
            DALLOC(nx, len * 32);
            LAUNCH(c, st, "merkle_level", k_merkle_level, blocks(len, 128), 128, 0, layers.back(), nx, (uint32_t)len);
            layers.push_back(nx);
            lens.push_back(len);
        }
        int L = (int)layers.size() - 1;  // tree layers are layers[1..L]
        if (levels >= L) return fail(ZGR_ERR_INVALID_ARG, "tree has only %d layers", L);
        uint32_t* droot = nullptr;
        DALLOC(droot, 32);
        layers.push_back(droot);
        LAUNCH(c, st, "file_root", k_file_root, 1, 1, 0, layers[L], (int)lens[L], droot);
        size_t off = 0;
        if (out_layers) {
            for (int k = 0; k <= levels; k++) {
                HIPCHK(hipMemcpyAsync(out_layers + off, layers[L - k], lens[L - k] * 32, hipMemcpyDeviceToHost, st));
                off += lens[L - k] * 32;
            }
        }
        HIPCHK(hipMemcpyAsync(file_root, droot, 32, hipMemcpyDeviceToHost, st));
        return sync_all(c);
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    for (auto* p : layers) (void)hipFree(p);
    return rc;
}

int zgr_merkle_noalloc(zgr_ctx* c, const void* data_dev, size_t n_elems, int levels, void* out_layers_dev,
                       void* file_root_dev) {
    Range range_("zgr_merkle_noalloc");
    if (!c || !data_dev || !file_root_dev || (levels >= 0 && !out_layers_dev) || levels < -1)
        return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (n_elems == 0 || n_elems % 8) return fail(ZGR_ERR_INVALID_ARG, "n_elems %zu unsupported", n_elems);
    // the layers below the root: n/8 leaves, halving while the count is even and above
    // coset_n; at most n/4 nodes of 32 B in all
    std::vector<size_t> lens;
    size_t total = 0;
    for (size_t len = n_elems / 8;; len /= 2) {
        lens.push_back(len);
        total += len * 32;
        if (len <= (size_t)c->coset_n || (len & 1)) break;
    }
    if (total > ZGR_MERKLE_SCRATCH_BYTES)
        return fail(ZGR_ERR_INVALID_ARG, "n_elems %zu needs %zu B of Merkle layers, the no-alloc scratch holds %zu B",
                    n_elems, total, (size_t)ZGR_MERKLE_SCRATCH_BYTES);
    const int L = (int)lens.size();  // tree layers 1..L = lens[0..L-1]
    if (levels >= L) return fail(ZGR_ERR_INVALID_ARG, "tree has only %d layers", L);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    uint8_t* base = nullptr;
    CHECK(noalloc_scratch(c->device, &base));  // resolved when the context was created
    hipStream_t st = c->lanes[0].st;
    std::vector<uint32_t*> layer(L);
    size_t off = 0;
    for (int i = 0; i < L; i++) {
        layer[i] = (uint32_t*)(base + off);
        off += lens[i] * 32;
    }
    // every node is written before it is read: nothing to clear
    LAUNCH(c, st, "merkle_leaves", k_merkle_leaves_flat, blocks(lens[0], 128), 128, 0, (const uint32_t*)data_dev,
           layer[0], (uint32_t)lens[0]);
    for (int i = 1; i < L; i++)
        LAUNCH(c, st, "merkle_level", k_merkle_level, blocks(lens[i], 128), 128, 0, layer[i - 1], layer[i],
               (uint32_t)lens[i]);
    LAUNCH(c, st, "file_root", k_file_root, 1, 1, 0, layer[L - 1], (int)lens[L - 1], (uint32_t*)file_root_dev);
    size_t out = 0;
    for (int k = 0; k <= levels; k++) {  // tree[k] = layer L - 1 - k
        HIPCHK(hipMemcpyAsync((uint8_t*)out_layers_dev + out, layer[L - 1 - k], lens[L - 1 - k] * 32,
                              hipMemcpyDeviceToDevice, st));
        out += lens[L - 1 - k] * 32;
    }
    return sync_all(c);
}

// zgr_verify_msm's third lane (msm_fb_line on line h: one line's bucket regions, every
// line's heap rows since the line index picks its tree); on failure it is released and
// the caller keeps two lanes
int ensure_vlane(zgr_ctx* c) {
    Lane& v = c->vlane;
    if (v.st) return ZGR_OK;
    auto alloc = [&]() -> int {
        HIPCHK(hipStreamCreateWithFlags(&v.st, hipStreamNonBlocking));
        DALLOC(v.y, c->n * 32);
        DALLOC(v.entries, c->n * FB_W * sizeof(uint2));
        DALLOC(v.counts, c->nb * MSM_W * 4);
        DALLOC(v.heap, (size_t)c->nlines * 2 * c->nb * X29_WORDS * 4);
        DALLOC(v.bstart, c->nb * FB_BSTART * 4);
        DALLOC(v.bkt, (size_t)c->nb * FB_REGION * X29_WORDS * 4);
        return ZGR_OK;
    };
    const int rc = alloc();
    if (rc != ZGR_OK) {
        for (void* p : {(void*)v.y, (void*)v.entries, (void*)v.counts, (void*)v.heap, (void*)v.bstart, (void*)v.bkt})
            if (p) (void)hipFree(p);
        if (v.st) (void)hipStreamDestroy(v.st);
        v = Lane{};
    }
    return rc;
}

int zgr_verify_msm(zgr_ctx* c, const uint32_t* row_ids, const uint8_t* rows, const uint8_t* alphas, size_t n_rows,
                   uint8_t* out_point) {
    Range range_("zgr_verify_msm");
    if (!c || !out_point || (n_rows && (!row_ids || !rows || !alphas))) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    const size_t ROW = (size_t)1 << c->log_row, COL = (size_t)1 << c->log_col;
    for (size_t k = 0; k < n_rows; k++) {
        if (row_ids[k] >= ROW * (size_t)c->coset_n)
            return fail(ZGR_ERR_INVALID_ARG, "row index %u past the encoded blob (%zu rows)", row_ids[k], ROW * c->coset_n);
        if (k && row_ids[k] <= row_ids[k - 1]) return fail(ZGR_ERR_INVALID_ARG, "row ids must be strictly increasing");
    }
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    std::vector<std::vector<uint32_t>> per(c->coset_n);
    for (size_t k = 0; k < n_rows; k++) per[row_ids[k] / ROW].push_back((uint32_t)k);
    for (int cs = 0; cs < c->coset_n; cs++)
        if (!per[cs].empty()) CHECK(ensure_params(c, cs));
    Lane& ln = c->lanes[0];
    hipStream_t st = ln.st;
    uint8_t *d_rows = nullptr, *d_alpha = nullptr;
    uint32_t *d_ids = nullptr, *d_ks = nullptr, *d_pts = nullptr;
    int* flag = nullptr;
    hipEvent_t ev_up = nullptr, ev_done = nullptr, ev_done2 = nullptr;
    auto body = [&]() -> int {
        HIPCHK(hipEventCreateWithFlags(&ev_up, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_done2, hipEventDisableTiming));
        CHECK(dmalloc((void**)&d_rows, n_rows * COL * 32 + 32));
        CHECK(dmalloc((void**)&d_alpha, n_rows * 32 + 32));
        CHECK(dmalloc((void**)&d_ids, n_rows * 4 + 4));
        CHECK(dmalloc((void**)&d_ks, n_rows * 4 + 4));
        CHECK(dmalloc((void**)&d_pts, (size_t)(c->coset_n + 1) * 64));
        CHECK(dmalloc((void**)&flag, sizeof(int)));
        LAUNCH(c, st, "verify", k_rec_clear, 1, 64, 0, flag);
        if (n_rows) {
            HIPCHK(hipMemcpyAsync(d_rows, rows, n_rows * COL * 32, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d_alpha, alphas, n_rows * 32, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d_ids, row_ids, n_rows * 4, hipMemcpyHostToDevice, st));
        }
        int npts = 0;
        size_t off = 0;
        std::vector<uint32_t> ks_all;
        for (int cs = 0; cs < c->coset_n; cs++) ks_all.insert(ks_all.end(), per[cs].begin(), per[cs].end());
        if (n_rows) HIPCHK(hipMemcpyAsync(d_ks, ks_all.data(), n_rows * 4, hipMemcpyHostToDevice, st));
        // with the shifted-copy tables the cosets run side by side, on lanes 0 and 1 and,
        // for a third, the verification lane (their streams wait for the uploads on lane
        // 0's, and lane 0's for their results); without it the third goes back to lane 0
        const bool fb = c->fb;
        int ncos = 0;
        for (int cs = 0; cs < c->coset_n; cs++) ncos += per[cs].empty() ? 0 : 1;
        const bool three = fb && ncos >= 3 && ensure_vlane(c) == ZGR_OK;
        HIPCHK(hipEventRecord(ev_up, st));
        bool used1 = false, used2 = false;
        for (int cs = 0; cs < c->coset_n; cs++) {
            const uint32_t nk = (uint32_t)per[cs].size();
            if (!nk) continue;
            Lane& lc = !fb ? ln : (three && npts == 2) ? c->vlane : c->lanes[npts & 1];
            if (&lc == &c->lanes[1] && !used1) {
                HIPCHK(hipStreamWaitEvent(lc.st, ev_up, 0));
                used1 = true;
            }
            if (&lc == &c->vlane && !used2) {
                HIPCHK(hipStreamWaitEvent(lc.st, ev_up, 0));
                used2 = true;
            }
            HIPCHK(hipMemsetAsync(lc.y, 0, c->n * 32, lc.st));
            LAUNCH(c, lc.st, "verify", k_ver_scalars, blocks((size_t)nk * COL, 256), 256, 0, d_rows, d_alpha,
                   d_ks + off, nk, d_ids, c->log_col, c->log_row, lc.y, flag);
            if (fb) {
                CHECK(msm_fb_line(c, lc, cs, c->h, d_pts + 16 * npts));  // line h: the basis
            } else {
                const uint32_t* basis = c->bases[cs] + (size_t)c->h * c->n * 16;
                CHECK(msm_windowed(c, lc, basis, c->log_n, d_pts + 16 * npts));
            }
            npts++;
            off += nk;
        }
        if (used1) {
            HIPCHK(hipEventRecord(ev_done, c->lanes[1].st));
            HIPCHK(hipStreamWaitEvent(st, ev_done, 0));
        }
        if (used2) {
            HIPCHK(hipEventRecord(ev_done2, c->vlane.st));
            HIPCHK(hipStreamWaitEvent(st, ev_done2, 0));
        }
        LAUNCH(c, st, "verify", k_ver_sum, 1, 64, 0, d_pts, npts, d_pts + 16 * c->coset_n);
        int hflag = 0;
        HIPCHK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(out_point, d_pts + 16 * c->coset_n, 64, hipMemcpyDeviceToHost, st));
        CHECK(sync_all(c));
        if (hflag) return fail(ZGR_ERR_INVALID_ARG, "row element or alpha is not a canonical field element");
        return ZGR_OK;
    };
    int rc = body();
    (void)sync_all(c);
    for (void* p : {(void*)d_rows, (void*)d_alpha, (void*)d_ids, (void*)d_ks, (void*)d_pts, (void*)flag})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {ev_up, ev_done, ev_done2})
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int zgr_recover(zgr_ctx* c, const uint32_t* row_ids, const uint8_t* rows, size_t n_rows, uint8_t* raw_out,
                size_t raw_cap, size_t* raw_len) {
    Range range_("zgr_recover");
    if (!c || (n_rows && (!row_ids || !rows)) || !raw_out || !raw_len) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    const int lc = c->log_col, h = c->log_row, log_n = c->log_n;
    const size_t ROW = (size_t)1 << h, COL = (size_t)1 << lc, n = c->n;
    const size_t max_raw = 31 * n - 4;
    if (lc < 1 || lc > 11) return fail(ZGR_ERR_INVALID_ARG, "recovery needs 1 <= log_col <= 11, context has %d", lc);
    if (raw_cap < max_raw) return fail(ZGR_ERR_INVALID_ARG, "raw_out holds %zu bytes, needs %zu", raw_cap, max_raw);
    for (size_t k = 1; k < n_rows; k++)
        if (row_ids[k] <= row_ids[k - 1]) return fail(ZGR_ERR_INVALID_ARG, "row ids must be strictly increasing");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->lanes[0].st;
    size_t slot = 0;
    auto body = [&]() -> int {
        // buffer i of the workspace, grown to at least `bytes`
        auto alloc = [&](auto*& p, size_t bytes) -> int {
            bytes = bytes ? bytes : 32;
            if (slot == c->rec_buf.size()) {
                c->rec_buf.push_back(nullptr);
                c->rec_cap.push_back(0);
            }
            if (c->rec_cap[slot] < bytes) {
                if (c->rec_buf[slot]) HIPCHK(hipFree(c->rec_buf[slot]));
                c->rec_buf[slot] = nullptr;
                c->rec_cap[slot] = 0;
                CHECK(dmalloc(&c->rec_buf[slot], bytes));
                c->rec_cap[slot] = bytes;
            }
            p = (std::remove_reference_t<decltype(p)>)c->rec_buf[slot++];
            return ZGR_OK;
        };
        int* flag = nullptr;
        uint8_t* d_rows = nullptr;
        CHECK(alloc(flag, sizeof(int)));
        CHECK(alloc(d_rows, n_rows * COL * 32));
        LAUNCH(c, st, "recover", k_rec_clear, 1, 64, 0, flag);
        if (n_rows) HIPCHK(hipMemcpyAsync(d_rows, rows, n_rows * COL * 32, hipMemcpyHostToDevice, st));
        int hflag = 0;
        // raw_slice_to_line runs over every row before the row count is looked at (lib.rs:21-36)
        if (n_rows)
            LAUNCH(c, st, "recover", k_rec_canon, blocks(n_rows * COL, 256), 256, 0, d_rows, (uint64_t)(n_rows * COL),
                   flag);
        HIPCHK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hflag & 1) return fail(ZGR_ERR_INVALID_ARG, "row element is not a canonical field element (the reference panics)");
        if (n_rows < ROW) return fail(ZGR_ERR_INVALID_ARG, "Not enough valid lines");
        if (row_ids[n_rows - 1] >= ROW * (size_t)c->coset_n)
            return fail(ZGR_ERR_INVALID_ARG, "Cannot recover data: RowIdOverflow");
        // interpolation set S = the first ROW rows (every received coset-0 row is among
        // them), extra rows X = the rest, targets = the erased coset-0 rows, then X
        const uint32_t m = (uint32_t)ROW, nx = (uint32_t)(n_rows - ROW);
        std::vector<uint32_t> erased;
        size_t n0 = 0;
        while (n0 < n_rows && row_ids[n0] < ROW) n0++;
        for (uint32_t l = 0, k = 0; l < ROW; l++) {
            if (k < n0 && row_ids[k] == l) k++;
            else erased.push_back(l);
        }
        const uint32_t e0 = (uint32_t)erased.size(), nt = e0 + nx;
        uint64_t g[4];  // gamma < 2^252 < r, fresh for every call
        {
            std::random_device rd;
            for (auto& w : g) w = ((uint64_t)rd() << 32) | rd();
            g[3] &= 0x0fffffffffffffffull;
        }
        uint32_t *d_ids, *d_gamma, *d_inv, *d_cpow, *d_gpow, *d_u, *d_V, *d_yv, *d_b, *d_t, *d_A, *d_L, *d_out,
            *d_er, *d_blob;
        uint8_t* d_raw;
        CHECK(alloc(d_ids, n_rows * 4));
        CHECK(alloc(d_gamma, 32));
        CHECK(alloc(d_inv, (size_t)(c->coset_n + 2) * 32));
        CHECK(alloc(d_cpow, (size_t)c->coset_n * COL * 32));
        CHECK(alloc(d_gpow, COL * 32));
        CHECK(alloc(d_u, n_rows * 32));
        CHECK(alloc(d_V, n_rows * COL * 32));
        CHECK(alloc(d_yv, n_rows * 32));
        CHECK(alloc(d_b, (size_t)m * 32));
        CHECK(alloc(d_t, (size_t)nt * 32));
        CHECK(alloc(d_A, (size_t)nt * 32));
        CHECK(alloc(d_L, (size_t)nt * m * 32));
        CHECK(alloc(d_out, (size_t)e0 * COL * 32));
        CHECK(alloc(d_er, (size_t)e0 * 4));
        CHECK(alloc(d_blob, n * 32));
        CHECK(alloc(d_raw, 31 * n));
        HIPCHK(hipMemcpyAsync(d_ids, row_ids, n_rows * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_gamma, g, 32, hipMemcpyHostToDevice, st));
        if (e0) HIPCHK(hipMemcpyAsync(d_er, erased.data(), (size_t)e0 * 4, hipMemcpyHostToDevice, st));
        const uint32_t nr = (uint32_t)n_rows;
        LAUNCH(c, st, "recover", k_rec_inverses, 1, 64, 0, c->consts, d_gamma, c->coset_n, lc, d_inv);
        LAUNCH(c, st, "recover", k_rec_consts, blocks((size_t)(c->coset_n + 1) * COL, 64), 64, 0, d_inv, c->coset_n,
               lc, d_cpow, d_gpow);
        LAUNCH(c, st, "recover", k_rec_points, blocks(nr, 64), 64, 0, d_ids, nr, c->consts, c->tw, lc, h, d_u);
        LAUNCH(c, st, "recover_rows", k_rec_rows, nr, 256, (COL + 256) * 32, d_rows, d_ids, d_cpow, d_gpow, c->tw_inv,
               lc, h, d_V, d_yv);
        LAUNCH(c, st, "recover", k_rec_prod, m, 256, 0, d_u, d_u, m, 1, 0, d_b);  // B_r, inverted in k_rec_lagrange
        if (e0) LAUNCH(c, st, "recover", k_rec_targets, blocks(e0, 64), 64, 0, d_er, e0, c->tw, lc, h, d_t);
        if (nx) HIPCHK(hipMemcpyAsync(d_t + 8 * (size_t)e0, d_u + 8 * (size_t)m, (size_t)nx * 32, hipMemcpyDeviceToDevice, st));
        if (nt) {
            const uint32_t chunks = (m + REC_CHUNK - 1) / REC_CHUNK;
            LAUNCH(c, st, "recover", k_rec_prod, nt, 256, 0, d_t, d_u, m, 0, 0, d_A);
            uint32_t *d_t29, *d_u29, *d_b29;  // targets, points and B_r in the 29-bit form
            CHECK(alloc(d_t29, (size_t)nt * 36));
            CHECK(alloc(d_u29, (size_t)m * 36));
            CHECK(alloc(d_b29, (size_t)m * 36));
            LAUNCH(c, st, "recover_lagrange", k_rec_to29, blocks(nt, 256), 256, 0, d_t, nt, d_t29);
            LAUNCH(c, st, "recover_lagrange", k_rec_to29, blocks(m, 256), 256, 0, d_u, m, d_u29);
            LAUNCH(c, st, "recover_lagrange", k_rec_to29, blocks(m, 256), 256, 0, d_b, m, d_b29);
            LAUNCH(c, st, "recover_lagrange", k_rec_lagrange, blocks((size_t)nt * chunks, 64), 64, 0, d_t29, nt, d_u29,
                   d_b29, m, d_L);
        }
        if (e0) {
            const char* gm = getenv("ZGR_REC_GEMM");
            if (gm && !strcmp(gm, "valu")) {  // 29-bit limb products on the VALU
                dim3 grid(blocks(COL, REC_TI), blocks(e0, REC_TL));
                LAUNCH(c, st, "recover_gemm", k_rec_gemm, grid, 256, 0, d_L, d_V, e0, m, (uint32_t)COL, d_out);
            } else {  // int8 limb planes on the matrix cores
                if (m > RM_MAX_M) return fail(ZGR_ERR_INVALID_ARG, "recovery GEMM over %u sources (max %u)", m, RM_MAX_M);
                const uint32_t tp = (e0 + 31) & ~31u, mp = (m + 31) & ~31u, cp = ((uint32_t)COL + 31) & ~31u;
                int8_t *d_lp, *d_vp;
                int* d_diag;
                CHECK(alloc(d_lp, (size_t)RM_LIMBS * tp * mp));
                CHECK(alloc(d_vp, (size_t)RM_LIMBS * cp * mp));
                CHECK(alloc(d_diag, (size_t)RM_DIAG * tp * cp * 4));
                LAUNCH(c, st, "recover_planes", k_rec_planes_l, blocks((size_t)tp * mp, 256), 256, 0, d_L, e0, m, tp, mp,
                       d_lp);
                LAUNCH(c, st, "recover_planes", k_rec_planes_v, blocks((size_t)cp * mp, 256), 256, 0, d_V, m,
                       (uint32_t)COL, mp, cp, d_vp);
                LAUNCH(c, st, "recover_gemm", k_rec_gemm_mfma, dim3(cp / 32, tp / 32), RM_WAVES * 64, 0, d_lp, d_vp,
                       tp, mp, cp, d_diag);
                LAUNCH(c, st, "recover_fold", k_rec_fold, dim3(blocks(COL, 64), e0), 64, 0, d_diag, tp, cp, e0,
                       (uint32_t)COL, d_out);
            }
        }
        uint32_t* d_c = nullptr;  // the degree-N coefficient, when there are extra rows
        if (nx) {
            uint32_t* d_cx;
            CHECK(alloc(d_cx, (size_t)nx * 32));
            CHECK(alloc(d_c, 32));
            LAUNCH(c, st, "recover", k_rec_check, nx, 256, 0, d_L + 8 * (size_t)e0 * m, d_A + 8 * (size_t)e0, d_yv, d_V,
                   m, (uint32_t)COL, d_cx, flag);
            LAUNCH(c, st, "recover", k_rec_fold_const, 1, 256, 0, d_cx, nx, d_c, flag);
        }
        if (n0) LAUNCH(c, st, "recover", k_rec_pass, blocks(n0 * COL, 256), 256, 0, d_rows, d_ids, (uint32_t)n0, lc, d_blob);
        if (e0)
            LAUNCH(c, st, "recover_rows", k_rec_out, e0, 256, COL * 32, d_out, d_A, d_er, c->tw, lc, h,
                   (const uint32_t*)d_c, d_blob);
        LAUNCH(c, st, "recover", k_rec_pack, blocks(31 * n / 4, 256), 256, 0, d_blob, (uint32_t)n, (uint32_t*)d_raw,
               flag);
        LAUNCH(c, st, "recover", k_rec_padding, blocks(max_raw, 256), 256, 0, d_raw, (uint64_t)max_raw, flag);
        uint8_t lenb[4];
        HIPCHK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(lenb, d_raw + max_raw, 4, hipMemcpyDeviceToHost, st));
        CHECK(sync_all(c));
        // the reference's order: the degree assert in data_poly, then RawBlob -> RawData
        if (hflag & 2)
            return fail(ZGR_ERR_INCONSISTENT, "rows do not lie on one polynomial of degree <= N (the reference panics)");
        if (hflag & 4) return fail(ZGR_ERR_INVALID_ARG, "Incorrect scalar");
        if (hflag & 8) return fail(ZGR_ERR_INVALID_ARG, "Incorrect length field");
        if (hflag & 16) return fail(ZGR_ERR_INVALID_ARG, "Non zero in padding range");
        const size_t len = (size_t)lenb[0] | ((size_t)lenb[1] << 8) | ((size_t)lenb[2] << 16) | ((size_t)lenb[3] << 24);
        if (len) HIPCHK(hipMemcpy(raw_out, d_raw, len, hipMemcpyDeviceToHost));
        *raw_len = len;
        return ZGR_OK;
    };
    int rc = body();
    (void)hipStreamSynchronize(st);
    return rc;
}

int zgr_autotune(zgr_ctx* c, char* report, size_t cap) {
    Range range_("zgr_autotune");
    if (!c) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, 0));  // the bucket kernels read coset 0's tables
    HIPCHK(hipSetDevice(c->device));
    CHECK(sync_all(c));
    Lane& ln = c->lanes[0];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const bool prof = c->profiling;
    c->profiling = false;
    // mean device time of `reps` runs of f on lane 0's stream, after one warm-up run
    auto time_ms = [&](auto&& f, int reps, double* ms) -> int {
        CHECK(f());
        HIPCHK(hipEventRecord(e0, ln.st));
        for (int r = 0; r < reps; r++) CHECK(f());
        HIPCHK(hipEventRecord(e1, ln.st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t / reps;
        return ZGR_OK;
    };
    std::string rep = "{";
    auto body = [&]() -> int {
        char buf[256];
        // 1. chunk-group plan per density class: scalars whose non-zero chunks are those
        //    of coset 0 of a blob filling a share of the raw rows (chunk b holds raw row
        //    bitrev(b)), plus the footer row; 12 lines, sort + bucket kernels
        if (c->fb && c->plans.size() > 1) {
            const size_t rows = c->nb;
            const double share[ZGR_DENSITY_CLASSES] = {1.0 / 32, 1.0 / 8, 3.0 / 8, 1.0};
            std::mt19937_64 rng(20260);
            std::vector<uint64_t> sc(c->n * 4);
            rep += "\"msm_plan\": {";
            for (int cls = 0; cls < ZGR_DENSITY_CLASSES; cls++) {
                const size_t nz = (size_t)(share[cls] * rows);
                for (size_t b = 0; b < c->nb; b++) {
                    size_t row = 0;
                    for (int k = 0; k < c->h; k++) row |= ((b >> k) & 1) << (c->h - 1 - k);
                    const bool on = row < nz || row == rows - 1;
                    for (size_t i = 0; i < c->bs; i++) {
                        uint64_t* v = &sc[4 * (b * c->bs + i)];
                        for (int q = 0; q < 4; q++) v[q] = on ? rng() : 0;
                        v[3] &= 0x0fffffffffffffffull;  // below 2^252 < r
                    }
                }
                HIPCHK(hipMemcpy(ln.y, sc.data(), c->n * 32, hipMemcpyHostToDevice));
                int best = 0;
                double best_ms = 1e30;
                snprintf(buf, sizeof(buf), "%s\"nonzero_rows_%zu_of_%zu\": {", cls ? ", " : "", nz == rows ? rows : nz + 1,
                         rows);
                rep += buf;
                for (size_t k = 0; k < c->plans.size(); k++) {
                    double ms = 0;
                    CHECK(time_ms([&] { return msm_lines(c, ln, 0, ln.y, 0, c->nlines, &c->plans[k]); }, 2, &ms));
                    snprintf(buf, sizeof(buf), "%s\"%s\": %.3f", k ? ", " : "", c->plans[k].name.c_str(), ms);
                    rep += buf;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = (int)k;
                    }
                }
                c->plan_of_class[cls] = best;
                snprintf(buf, sizeof(buf), ", \"chosen\": \"%s\"}", c->plans[best].name.c_str());
                rep += buf;
            }
            rep += "}";
        }
        // 2. split of the context's two-pass NTT: the iNTT and one coset's DIT, as the
        //    encode path runs them
        if (c->log_n > NTT_MAX_LOG && c->coset_n > 1) {
            const int n = c->log_n;
            std::vector<int> cand;
            for (int n1 : {(n + 1) / 2, n / 2, (n + 1) / 2 + 1})
                if (n1 <= NTT_MAX_LOG && n - n1 <= NTT_MAX_LOG &&
                    std::find(cand.begin(), cand.end(), n1) == cand.end())
                    cand.push_back(n1);
            uint32_t *a = c->slots[0].pts, *b = c->slots[0].coeff;
            HIPCHK(hipMemsetAsync(a, 0, c->n * 32, ln.st));
            int best = 0;
            double best_ms = 1e30;
            rep += std::string(rep.size() > 1 ? ", " : "") + "\"ntt_split\": {";
            for (size_t k = 0; k < cand.size(); k++) {
                c->ntt_n1 = cand[k];
                double ms = 0;
                CHECK(time_ms([&] {
                    CHECK(ntt_dif(c, ln.st, a, b, n, 1, true));
                    return ntt_dit(c, ln.st, b, ln.y, n, 1, false, c->scale[1]);
                }, 3, &ms));
                snprintf(buf, sizeof(buf), "%s\"%d+%d\": %.4f", k ? ", " : "", cand[k], n - cand[k], ms);
                rep += buf;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = cand[k];
                }
            }
            c->ntt_n1 = best;
            snprintf(buf, sizeof(buf), ", \"chosen\": \"%d+%d\"}", best, n - best);
            rep += buf;
        }
        rep += "}";
        return sync_all(c);
    };
    int rc = body();
    c->profiling = prof;
    (void)hipStreamSynchronize(ln.st);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    c->tune_report = rep;
    if (report && cap) {
        const size_t k = std::min(cap - 1, rep.size());
        memcpy(report, rep.data(), k);
        report[k] = 0;
    }
    return ZGR_OK;
}

double zgr_kernel_time_ms(zgr_ctx* c, const char* name, int* count) {
    if (!c || !name) return 0.0;
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->times.find(name);
    if (it == c->times.end()) {
        if (count) *count = 0;
        return 0.0;
    }
    if (count) *count = it->second.second;
    return it->second.second ? it->second.first / it->second.second : 0.0;
}

void zgr_kernel_time_reset(zgr_ctx* c) {
    if (!c) return;
    std::lock_guard<std::mutex> lk(c->mu);
    c->times.clear();
}

void zgr_set_profiling(zgr_ctx* c, int enabled) {
    if (!c) return;
    std::lock_guard<std::mutex> lk(c->mu);
    c->profiling = enabled != 0;
}

}  // extern "C"
