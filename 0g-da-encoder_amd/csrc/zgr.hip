// C ABI of the gfx950 encoder (include/zgr.h): context, resident params, the
// encode pipeline and the primitives.  All device work of a context runs on the
// context's own stream; calls on one context are serialised by a mutex.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zgr.h"
#include "kernels.h"

using namespace zgr;

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

inline uint32_t blocks(size_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// dynamic LDS of the bucket kernel: staged bases, reused for the lane reduction
inline size_t accum_lds_bytes(size_t bs) {
    size_t a = bs * LDS_PT * 4, r = (size_t)ACC_THREADS * LDS_XP * 4;
    return a > r ? a : r;
}

struct Timer {
    hipEvent_t a, b;
    std::string name;
};

}  // namespace

struct zgr_ctx {
    int device = 0;
    int log_col = 0, log_row = 0, log_n = 0, h = 0, coset_n = 0, nlines = 0;
    size_t n = 0, nb = 0, bs = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;

    // domain tables
    uint32_t* consts = nullptr;   // K_COUNT Fr
    uint32_t* tw = nullptr;       // n/2 Fr: omega^k
    uint32_t* tw_inv = nullptr;   // n/2 Fr: omega^-k
    uint32_t* scale[3] = {nullptr, nullptr, nullptr};  // coset scale tables (DIF order)

    // resident params
    uint32_t* bases[3] = {nullptr, nullptr, nullptr};
    bool loaded[3] = {false, false, false};

    // per-blob workspace
    uint8_t* raw = nullptr;       // staging for host input (31 n bytes)
    uint32_t* pts = nullptr;      // n Fr
    uint32_t* coeff = nullptr;    // n Fr
    uint32_t* y = nullptr;        // n Fr
    int8_t* digits = nullptr;     // 32 n
    uint32_t* entries = nullptr;  // 32 n u32
    uint32_t* counts = nullptr;   // 32 nb
    uint32_t* heap = nullptr;     // nlines * 32 * 2nb XYZZ
    uint32_t* outpts = nullptr;   // (4nb + 2) points * 16 u32 per coset, coset_n cosets
    uint32_t* data = nullptr;     // coset_n * n * 8 u32 (canonical encoded data)
    std::vector<uint32_t*> mlayers;  // merkle layers, bottom (leaves) first
    std::vector<size_t> mlens;
    uint32_t* root = nullptr;     // 8 u32
    int* flag = nullptr;

    // profiling
    bool profiling = false;
    std::vector<Timer> pending;
    std::map<std::string, std::pair<double, int>> times;
};

namespace {

// launch helper with optional event timing
template <typename F>
int timed(zgr_ctx* c, const char* name, F&& launch) {
    Timer t{};
    if (c->profiling) {
        HIPCHK(hipEventCreate(&t.a));
        HIPCHK(hipEventCreate(&t.b));
        HIPCHK(hipEventRecord(t.a, c->stream));
    }
    launch();
    HIPCHK(hipGetLastError());
    if (c->profiling) {
        HIPCHK(hipEventRecord(t.b, c->stream));
        t.name = name;
        c->pending.push_back(t);
    }
    return ZGR_OK;
}

#define LAUNCH(ctx, name, kern, grid, block, shm, ...)                                          \
    do {                                                                                        \
        int rc_ = timed(ctx, name, [&] {                                                        \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, (ctx)->stream, __VA_ARGS__); \
        });                                                                                     \
        if (rc_) return rc_;                                                                    \
    } while (0)

int sync_and_collect(zgr_ctx* c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& t : c->pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        auto& slot = c->times[t.name];
        slot.first += ms;
        slot.second += 1;
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    c->pending.clear();
    return ZGR_OK;
}

int dmalloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(ZGR_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return ZGR_OK;
}

#define DALLOC(ptr, bytes)                                         \
    do {                                                           \
        int rc_ = dmalloc((void**)&(ptr), (bytes));                \
        if (rc_) return rc_;                                       \
    } while (0)

#define CHECK(expr)           \
    do {                      \
        int rc_ = (expr);     \
        if (rc_) return rc_;  \
    } while (0)

// radix-2 transform of one 2^k row in place; dif: natural -> bitrev, else bitrev -> natural
int ntt_stages(zgr_ctx* c, uint32_t* a, int k, bool dif, bool inverse) {
    const uint32_t* tw = inverse ? c->tw_inv : c->tw;
    uint32_t half_n = 1u << (k - 1);
    if (dif) {
        for (int s = k; s >= 1; s--)
            LAUNCH(c, "ntt", k_ntt_dif_stage, blocks(half_n, 256), 256, 0, a, tw, s, c->log_n - s, half_n);
    } else {
        for (int s = 1; s <= k; s++)
            LAUNCH(c, "ntt", k_ntt_dit_stage, blocks(half_n, 256), 256, 0, a, tw, s, c->log_n - s, half_n);
    }
    return ZGR_OK;
}

// MSM of one coset's scalar vector (already as digits source `src`): fills c->heap
int msm_lines(zgr_ctx* c, int coset, const uint32_t* src, int from_y) {
    const int log_bs = c->log_n - c->h;
    LAUNCH(c, "msm_digits", k_msm_digits, blocks(c->n, 256), 256, 0, src, c->log_n, log_bs, c->digits, from_y);
    uint32_t sort_threads = c->bs < 256 ? (uint32_t)c->bs : 256u;
    LAUNCH(c, "msm_sort", k_msm_sort, (uint32_t)(c->nb * MSM_W), sort_threads, 0, c->digits, log_bs, c->entries,
           c->counts);
    size_t shm = accum_lds_bytes(c->bs);
    LAUNCH(c, "msm_accum", k_msm_accum, (uint32_t)(c->nb * c->nlines), ACC_THREADS, shm, c->bases[coset],
           c->entries, c->counts, c->heap, c->log_n, c->h, c->nlines);
    return ZGR_OK;
}

int msm_trees(zgr_ctx* c) {
    for (int d = c->h - 1; d >= 0; d--) {
        uint32_t total = (uint32_t)c->nlines * MSM_W << d;
        LAUNCH(c, "msm_tree", k_msm_tree_level, blocks(total, 64), 64, 0, c->heap, c->h, d, total);
    }
    return ZGR_OK;
}

size_t tree_pts(const zgr_ctx* c) { return 2 * c->nb - 2; }
size_t out_slots(const zgr_ctx* c) { return 2 * tree_pts(c) + 2; }

// AMT outputs of one coset into c->outpts + coset slot block
int amt_outputs(zgr_ctx* c, int coset) {
    uint32_t slots = (uint32_t)out_slots(c);
    CHECK(msm_trees(c));
    LAUNCH(c, "msm_horner", k_msm_horner, blocks(slots, 64), 64, 0, c->heap, c->h,
           c->outpts + (size_t)coset * slots * 16);
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

// copy the AMT outputs of all cosets + merkle to the host struct
int download_outputs(zgr_ctx* c, const zgr_encoded* out) {
    const size_t slots = out_slots(c), T = tree_pts(c);
    std::vector<uint8_t> pts(slots * 64 * c->coset_n);
    HIPCHK(hipMemcpyAsync(pts.data(), c->outpts, pts.size(), hipMemcpyDeviceToHost, c->stream));
    if (out->data)
        HIPCHK(hipMemcpyAsync(out->data, c->data, (size_t)c->coset_n * c->n * 32, hipMemcpyDeviceToHost,
                              c->stream));
    if (out->merkle) {
        // tree[k] = layer (L - 1 - k) for k = 0..h
        size_t off = 0;
        int L = (int)c->mlayers.size();
        for (int k = 0; k <= c->h; k++) {
            int li = L - 1 - k;
            HIPCHK(hipMemcpyAsync(out->merkle + off, c->mlayers[li], c->mlens[li] * 32, hipMemcpyDeviceToHost,
                                  c->stream));
            off += c->mlens[li] * 32;
        }
    }
    if (out->file_root) HIPCHK(hipMemcpyAsync(out->file_root, c->root, 32, hipMemcpyDeviceToHost, c->stream));
    CHECK(sync_and_collect(c));
    for (int cs = 0; cs < c->coset_n; cs++) {
        const uint8_t* base = pts.data() + (size_t)cs * slots * 64;
        if (out->proof_tree) memcpy(out->proof_tree + cs * T * 64, base, T * 64);
        if (out->comm_tree) memcpy(out->comm_tree + cs * T * 64, base + T * 64, T * 64);
        if (out->high) memcpy(out->high + cs * 64, base + (2 * T + 1) * 64, 64);
        if (cs > 0 && memcmp(base + 2 * T * 64, pts.data() + 2 * T * 64, 64))
            return fail(ZGR_ERR_INCONSISTENT, "coset %d commitment differs from coset 0", cs);
    }
    if (out->commitment) memcpy(out->commitment, pts.data() + 2 * T * 64, 64);
    return ZGR_OK;
}

// the device part of one encode: raw (device) -> outpts, data, merkle
int encode_on_device(zgr_ctx* c, const uint8_t* raw_dev, size_t len) {
    const uint32_t n = (uint32_t)c->n;
    for (int cs = 0; cs < c->coset_n; cs++) CHECK(ensure_params(c, cs));
    LAUNCH(c, "pack", k_pack, blocks(n, 256), 256, 0, raw_dev, len, c->log_col, c->log_row, c->pts);
    if (c->coset_n > 1) {
        HIPCHK(hipMemcpyAsync(c->coeff, c->pts, c->n * 32, hipMemcpyDeviceToDevice, c->stream));
        CHECK(ntt_stages(c, c->coeff, c->log_n, true, true));  // coefficients, bit-reversed
    }
    const uint32_t leaves_per_coset = n / 8;
    for (int cs = 0; cs < c->coset_n; cs++) {
        const uint32_t* y = c->pts;
        if (cs > 0) {
            LAUNCH(c, "coset_scale", k_mul_table, blocks(n, 256), 256, 0, c->y, c->coeff, c->scale[cs], n);
            CHECK(ntt_stages(c, c->y, c->log_n, false, false));
            y = c->y;
        }
        CHECK(msm_lines(c, cs, y, 1));
        CHECK(amt_outputs(c, cs));
        LAUNCH(c, "merkle_leaves", k_merkle_leaves, blocks(leaves_per_coset, 128), 128, 0, y, c->log_col,
               c->log_row, cs, c->data, c->mlayers[0]);
    }
    for (size_t l = 1; l < c->mlayers.size(); l++)
        LAUNCH(c, "merkle_level", k_merkle_level, blocks(c->mlens[l], 128), 128, 0, c->mlayers[l - 1],
               c->mlayers[l], (uint32_t)c->mlens[l]);
    LAUNCH(c, "file_root", k_file_root, 1, 1, 0, c->mlayers.back(), (int)c->mlens.back(), c->root);
    return ZGR_OK;
}

int check_len(const zgr_ctx* c, size_t len) {
    size_t max_raw = 31 * c->n - 4;
    if (len > max_raw)
        return fail(ZGR_ERR_TOO_LARGE, "TooLargeBlob { actual: %zu, expected_max: %zu }", len, max_raw);
    return ZGR_OK;
}

}  // namespace

extern "C" {

const char* zgr_last_error(void) { return g_err.c_str(); }

int zgr_ctx_create(int device, int log_col, int log_row, int coset_n, zgr_ctx** out) {
    if (!out) return fail(ZGR_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (log_col < 3 || log_row < 1 || log_col + log_row > 26 || coset_n < 1 || coset_n > 3 ||
        (1 << (log_col + log_row)) * coset_n > (1 << 28))
        return fail(ZGR_ERR_INVALID_ARG, "unsupported geometry log_col=%d log_row=%d coset_n=%d", log_col,
                    log_row, coset_n);
    if (log_col > 15) return fail(ZGR_ERR_INVALID_ARG, "log_col > 15 not supported");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(ZGR_ERR_INVALID_ARG, "device %d not present", device);
    HIPCHK(hipSetDevice(device));
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
    auto init = [&]() -> int {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        {
            int max_shm = 0;
            HIPCHK(hipDeviceGetAttribute(&max_shm, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
            size_t need = accum_lds_bytes(c->bs);
            if (need + 1024 > (size_t)max_shm)
                return fail(ZGR_ERR_INVALID_ARG, "bucket kernel needs %zu B of LDS, device allows %d", need, max_shm);
            HIPCHK(hipFuncSetAttribute((const void*)k_msm_accum, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)need));
        }
        DALLOC(c->consts, K_COUNT * 32);
        DALLOC(c->tw, c->n / 2 * 32);
        DALLOC(c->tw_inv, c->n / 2 * 32);
        DALLOC(c->raw, 31 * c->n);
        DALLOC(c->pts, c->n * 32);
        DALLOC(c->coeff, c->n * 32);
        DALLOC(c->y, c->n * 32);
        DALLOC(c->digits, c->n * MSM_W);
        DALLOC(c->entries, c->n * MSM_W * 4);
        DALLOC(c->counts, c->nb * MSM_W * 4);
        DALLOC(c->heap, (size_t)c->nlines * MSM_W * 2 * c->nb * 128);
        // holds either the AMT outputs of every coset or one multiple_multiexp result
        size_t opts = out_slots(c) * coset_n;
        if (opts < (size_t)c->nlines * c->nb) opts = (size_t)c->nlines * c->nb;
        DALLOC(c->outpts, opts * 64);
        DALLOC(c->data, (size_t)coset_n * c->n * 32);
        DALLOC(c->root, 32);
        DALLOC(c->flag, sizeof(int));
        size_t len = (size_t)coset_n * c->n / 8;
        while (true) {
            uint32_t* p = nullptr;
            DALLOC(p, len * 32);
            c->mlayers.push_back(p);
            c->mlens.push_back(len);
            if (len <= (size_t)coset_n || (len & 1)) break;
            len /= 2;
        }
        if ((int)c->mlayers.size() <= c->h) return fail(ZGR_ERR_INVALID_ARG, "merkle tree too shallow");
        LAUNCH(c, "setup", k_setup_consts, 1, 1, 0, c->consts, c->log_n, coset_n);
        LAUNCH(c, "setup", k_powers, blocks(c->n / 2, 256), 256, 0, c->tw, c->consts, (int)K_OMEGA,
               (uint32_t)(c->n / 2));
        LAUNCH(c, "setup", k_powers, blocks(c->n / 2, 256), 256, 0, c->tw_inv, c->consts, (int)K_OMEGA_INV,
               (uint32_t)(c->n / 2));
        for (int cs = 1; cs < coset_n; cs++) {
            DALLOC(c->scale[cs], c->n * 32);
            LAUNCH(c, "setup", k_coset_scale_table, blocks(c->n, 256), 256, 0, c->scale[cs], c->consts, cs,
                   c->log_n);
        }
        return sync_and_collect(c);
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
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->consts, c->tw, c->tw_inv, c->raw, c->pts, c->coeff, c->y, c->digits, c->entries,
                    c->counts, c->heap, c->outpts, c->data, c->root, c->flag};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (int i = 0; i < 3; i++) {
        if (c->scale[i]) hipFree(c->scale[i]);
        if (c->bases[i]) hipFree(c->bases[i]);
    }
    for (auto* p : c->mlayers) hipFree(p);
    for (auto& t : c->pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int zgr_params_load(zgr_ctx* c, int coset, const void* bases_mont, size_t nbytes) {
    if (!c || !bases_mont) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    size_t want = (size_t)c->nlines * c->n * 64;
    if (nbytes != want) return fail(ZGR_ERR_INVALID_ARG, "params size %zu, expected %zu", nbytes, want);
    HIPCHK(hipSetDevice(c->device));
    CHECK(alloc_bases(c, coset));
    HIPCHK(hipMemcpyAsync(c->bases[coset], bases_mont, want, hipMemcpyHostToDevice, c->stream));
    CHECK(sync_and_collect(c));
    c->loaded[coset] = true;
    return ZGR_OK;
}

int zgr_params_generate(zgr_ctx* c, const uint8_t tau[32]) {
    if (!c || !tau) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    uint32_t *table = nullptr, *tau_d = nullptr, *pc = nullptr;
    DALLOC(table, 32 * 256 * 64);
    DALLOC(tau_d, 32);
    DALLOC(pc, (3 + c->h) * 32);
    auto body = [&]() -> int {
        HIPCHK(hipMemcpyAsync(tau_d, tau, 32, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
        LAUNCH(c, "params_table", k_fixed_base_table, blocks(32 * 256, 64), 64, 0, table);
        for (int cs = 0; cs < c->coset_n; cs++) {
            CHECK(alloc_bases(c, cs));
            LAUNCH(c, "params_consts", k_params_consts, 1, 1, 0, tau_d, c->consts, cs, c->log_n, c->h, pc);
            LAUNCH(c, "params_gen", k_params_gen, blocks(c->n, 64), 64, 0, c->bases[cs], c->consts, pc, table,
                   c->log_n, c->h, c->flag);
        }
        int bad = 0;
        HIPCHK(hipMemcpyAsync(&bad, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        CHECK(sync_and_collect(c));
        if (bad) return fail(ZGR_ERR_INVALID_ARG, "tau hits the evaluation domain (or a base is infinity)");
        for (int cs = 0; cs < c->coset_n; cs++) c->loaded[cs] = true;
        return ZGR_OK;
    };
    int rc = body();
    hipStreamSynchronize(c->stream);
    hipFree(table);
    hipFree(tau_d);
    hipFree(pc);
    return rc;
}

int zgr_params_download(zgr_ctx* c, int coset, void* out, size_t nbytes) {
    if (!c || !out) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, coset));
    size_t want = (size_t)c->nlines * c->n * 64;
    if (nbytes != want) return fail(ZGR_ERR_INVALID_ARG, "params size %zu, expected %zu", nbytes, want);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(out, c->bases[coset], want, hipMemcpyDeviceToHost, c->stream));
    return sync_and_collect(c);
}

int zgr_multiple_multiexp(zgr_ctx* c, int coset, const uint8_t* ri_mont, size_t n, uint8_t* out_points) {
    if (!c || !ri_mont || !out_points) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, coset));
    if (n != c->n) return fail(ZGR_ERR_INVALID_ARG, "scalar count %zu, expected %zu", n, c->n);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->y, ri_mont, n * 32, hipMemcpyHostToDevice, c->stream));
    CHECK(msm_lines(c, coset, c->y, 0));
    uint32_t total = (uint32_t)(c->nlines * c->nb);
    LAUNCH(c, "msm_leaves_out", k_msm_leaves_out, blocks(total, 64), 64, 0, c->heap, c->h, c->nlines, c->outpts);
    HIPCHK(hipMemcpyAsync(out_points, c->outpts, (size_t)total * 64, hipMemcpyDeviceToHost, c->stream));
    return sync_and_collect(c);
}

int zgr_amt_prove(zgr_ctx* c, int coset, const uint8_t* ri_mont, size_t n, uint8_t* commitment,
                  uint8_t* comm_tree, uint8_t* proof_tree, uint8_t* high) {
    if (!c || !ri_mont) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(ensure_params(c, coset));
    if (n != c->n) return fail(ZGR_ERR_INVALID_ARG, "scalar count %zu, expected %zu", n, c->n);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->y, ri_mont, n * 32, hipMemcpyHostToDevice, c->stream));
    CHECK(msm_lines(c, coset, c->y, 0));
    CHECK(amt_outputs(c, 0));
    const size_t slots = out_slots(c), T = tree_pts(c);
    std::vector<uint8_t> pts(slots * 64);
    HIPCHK(hipMemcpyAsync(pts.data(), c->outpts, pts.size(), hipMemcpyDeviceToHost, c->stream));
    CHECK(sync_and_collect(c));
    if (proof_tree) memcpy(proof_tree, pts.data(), T * 64);
    if (comm_tree) memcpy(comm_tree, pts.data() + T * 64, T * 64);
    if (commitment) memcpy(commitment, pts.data() + 2 * T * 64, 64);
    if (high) memcpy(high, pts.data() + (2 * T + 1) * 64, 64);
    return ZGR_OK;
}

int zgr_encode_device(zgr_ctx* c, const void* raw_dev, size_t len, const zgr_encoded* out) {
    if (!c || !out || (!raw_dev && len)) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(check_len(c, len));
    HIPCHK(hipSetDevice(c->device));
    CHECK(encode_on_device(c, (const uint8_t*)(raw_dev ? raw_dev : c->raw), len));
    return download_outputs(c, out);
}

int zgr_encode(zgr_ctx* c, const uint8_t* raw, size_t len, const zgr_encoded* out) {
    if (!c || !out || (!raw && len)) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    CHECK(check_len(c, len));
    HIPCHK(hipSetDevice(c->device));
    if (len) HIPCHK(hipMemcpyAsync(c->raw, raw, len, hipMemcpyHostToDevice, c->stream));
    CHECK(encode_on_device(c, c->raw, len));
    return download_outputs(c, out);
}

int zgr_encode_batch_device(zgr_ctx* c, int nblobs, const void* const* raw_dev, const size_t* lens,
                            const zgr_encoded* outs) {
    if (!c || nblobs < 0 || (nblobs && (!raw_dev || !lens || !outs)))
        return fail(ZGR_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < nblobs; i++) CHECK(check_len(c, lens[i]));
    for (int i = 0; i < nblobs; i++) CHECK(zgr_encode_device(c, raw_dev[i], lens[i], &outs[i]));
    return ZGR_OK;
}

int zgr_ntt(zgr_ctx* c, uint8_t* data_mont, int log_n, int batch, int inverse) {
    if (!c || !data_mont || batch < 0) return fail(ZGR_ERR_INVALID_ARG, "bad argument");
    if (log_n < 1 || log_n > c->log_n)
        return fail(ZGR_ERR_INVALID_ARG, "log_n %d outside [1, %d] of this context", log_n, c->log_n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    const size_t row = (size_t)32 << log_n;
    for (int r = 0; r < batch; r++) {
        HIPCHK(hipMemcpyAsync(c->pts, data_mont + r * row, row, hipMemcpyHostToDevice, c->stream));
        LAUNCH(c, "ntt", k_bitrev_copy, blocks((size_t)1 << log_n, 256), 256, 0, c->y, c->pts, log_n);
        CHECK(ntt_stages(c, c->y, log_n, false, inverse != 0));
        if (inverse)
            LAUNCH(c, "ntt", k_mul_ninv_sub, blocks((size_t)1 << log_n, 256), 256, 0, c->y, c->consts,
                   c->log_n - log_n, 1u << log_n);
        HIPCHK(hipMemcpyAsync(data_mont + r * row, c->y, row, hipMemcpyDeviceToHost, c->stream));
    }
    return sync_and_collect(c);
}

int zgr_coset_extend(zgr_ctx* c, const uint8_t* data_mont, int coset, uint8_t* out_mont) {
    if (!c || !data_mont || !out_mont) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (coset < 0 || coset >= c->coset_n) return fail(ZGR_ERR_INVALID_ARG, "coset %d out of range", coset);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    const uint32_t n = (uint32_t)c->n;
    HIPCHK(hipMemcpyAsync(c->coeff, data_mont, c->n * 32, hipMemcpyHostToDevice, c->stream));
    if (coset == 0) {
        HIPCHK(hipMemcpyAsync(c->y, c->coeff, c->n * 32, hipMemcpyDeviceToDevice, c->stream));
    } else {
        CHECK(ntt_stages(c, c->coeff, c->log_n, true, true));
        LAUNCH(c, "coset_scale", k_mul_table, blocks(n, 256), 256, 0, c->y, c->coeff, c->scale[coset], n);
        CHECK(ntt_stages(c, c->y, c->log_n, false, false));
    }
    HIPCHK(hipMemcpyAsync(out_mont, c->y, c->n * 32, hipMemcpyDeviceToHost, c->stream));
    return sync_and_collect(c);
}

int zgr_msm(zgr_ctx* c, const uint8_t* bases_mont, const uint8_t* scalars_mont, size_t n, uint8_t* out_point) {
    if (!c || !bases_mont || !scalars_mont || !out_point) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (n == 0 || n > c->n) return fail(ZGR_ERR_INVALID_ARG, "msm size %zu outside [1, %zu]", n, c->n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    int log_np = 0;
    while (((size_t)1 << log_np) < n) log_np++;
    const size_t np = (size_t)1 << log_np;
    const int log_bs = log_np < 10 ? log_np : 10;
    const int h = log_np - log_bs;
    uint32_t* dbases = nullptr;
    DALLOC(dbases, np * 64);
    auto body = [&]() -> int {
        // zero-padded scalars: padding points are never read by the bucket kernel
        HIPCHK(hipMemsetAsync(c->y, 0, np * 32, c->stream));
        HIPCHK(hipMemsetAsync(dbases, 0, np * 64, c->stream));
        HIPCHK(hipMemcpyAsync(c->y, scalars_mont, n * 32, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dbases, bases_mont, n * 64, hipMemcpyHostToDevice, c->stream));
        LAUNCH(c, "msm_digits", k_msm_digits, blocks(np, 256), 256, 0, c->y, log_np, log_bs, c->digits, 0);
        uint32_t bs = 1u << log_bs, nb = 1u << h;
        LAUNCH(c, "msm_sort", k_msm_sort, nb * MSM_W, bs < 256 ? bs : 256u, 0, c->digits, log_bs, c->entries,
               c->counts);
        size_t shm = accum_lds_bytes(bs);
        LAUNCH(c, "msm_accum", k_msm_accum, nb, ACC_THREADS, shm, dbases, c->entries, c->counts, c->heap, log_np,
               h, 1);
        for (int d = h - 1; d >= 0; d--) {
            uint32_t total = (uint32_t)MSM_W << d;
            LAUNCH(c, "msm_tree", k_msm_tree_level, blocks(total, 64), 64, 0, c->heap, h, d, total);
        }
        LAUNCH(c, "msm_root", k_msm_root_out, 1, 64, 0, c->heap, h, c->outpts);
        HIPCHK(hipMemcpyAsync(out_point, c->outpts, 64, hipMemcpyDeviceToHost, c->stream));
        return sync_and_collect(c);
    };
    int rc = body();
    hipStreamSynchronize(c->stream);
    hipFree(dbases);
    return rc;
}

int zgr_ecfft(zgr_ctx* c, const uint8_t* points, int log_n, int inverse, uint8_t* out) {
    if (!c || !points || !out) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (log_n < 1 || log_n > c->log_n)
        return fail(ZGR_ERR_INVALID_ARG, "log_n %d outside [1, %d] of this context", log_n, c->log_n);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)1 << log_n;
    uint32_t *din = nullptr, *dpts = nullptr;
    DALLOC(din, n * 64);
    int rc = dmalloc((void**)&dpts, n * 128);
    if (rc) {
        hipFree(din);
        return rc;
    }
    auto body = [&]() -> int {
        HIPCHK(hipMemcpyAsync(din, points, n * 64, hipMemcpyHostToDevice, c->stream));
        LAUNCH(c, "ecfft", k_ec_load_bitrev, blocks(n, 256), 256, 0, din, dpts, log_n);
        const uint32_t* tw = inverse ? c->tw_inv : c->tw;
        uint32_t half_n = (uint32_t)(n / 2);
        for (int s = 1; s <= log_n; s++)
            LAUNCH(c, "ecfft", k_ec_dit_stage, blocks(half_n, 64), 64, 0, dpts, tw, s, c->log_n - s, half_n);
        LAUNCH(c, "ecfft", k_ec_store, blocks(n, 64), 64, 0, dpts, c->consts, c->log_n - log_n, inverse ? 1 : 0,
               din, (uint32_t)n);
        HIPCHK(hipMemcpyAsync(out, din, n * 64, hipMemcpyDeviceToHost, c->stream));
        return sync_and_collect(c);
    };
    rc = body();
    hipStreamSynchronize(c->stream);
    hipFree(din);
    hipFree(dpts);
    return rc;
}

int zgr_merkle(zgr_ctx* c, const uint8_t* data, size_t n_elems, int levels, uint8_t* out_layers,
               uint8_t* file_root) {
    if (!c || !data || !file_root) return fail(ZGR_ERR_INVALID_ARG, "null argument");
    if (n_elems == 0 || n_elems % 8) return fail(ZGR_ERR_INVALID_ARG, "n_elems %zu unsupported", n_elems);
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    std::vector<uint32_t*> layers;
    std::vector<size_t> lens;
    auto body = [&]() -> int {
        uint32_t* dsrc = nullptr;
        DALLOC(dsrc, n_elems * 32);
        layers.push_back(dsrc);
        lens.push_back(n_elems);
        HIPCHK(hipMemcpyAsync(dsrc, data, n_elems * 32, hipMemcpyHostToDevice, c->stream));
        size_t len = n_elems / 8;
        uint32_t* lv = nullptr;
        DALLOC(lv, len * 32);
        layers.push_back(lv);
        lens.push_back(len);
        LAUNCH(c, "merkle_leaves", k_merkle_leaves_flat, blocks(len, 128), 128, 0, dsrc, lv, (uint32_t)len);
        while (len > (size_t)c->coset_n && !(len & 1)) {
            len /= 2;
            uint32_t* nx = nullptr;
            DALLOC(nx, len * 32);
            LAUNCH(c, "merkle_level", k_merkle_level, blocks(len, 128), 128, 0, layers.back(), nx, (uint32_t)len);
            layers.push_back(nx);
            lens.push_back(len);
        }
        int L = (int)layers.size() - 1;  // tree layers are layers[1..L]
        if (levels >= L) return fail(ZGR_ERR_INVALID_ARG, "tree has only %d layers", L);
        uint32_t* droot = nullptr;
        DALLOC(droot, 32);
        layers.push_back(droot);
        LAUNCH(c, "file_root", k_file_root, 1, 1, 0, layers[L], (int)lens[L], droot);
        size_t off = 0;
        if (out_layers) {
            for (int k = 0; k <= levels; k++) {
                HIPCHK(hipMemcpyAsync(out_layers + off, layers[L - k], lens[L - k] * 32, hipMemcpyDeviceToHost,
                                      c->stream));
                off += lens[L - k] * 32;
            }
        }
        HIPCHK(hipMemcpyAsync(file_root, droot, 32, hipMemcpyDeviceToHost, c->stream));
        return sync_and_collect(c);
    };
    int rc = body();
    hipStreamSynchronize(c->stream);
    for (auto* p : layers) hipFree(p);
    return rc;
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
