// Device kernels of the encoding path.  Included once, by zgr.hip.
//
// Domain / data layout of one blob (N = 2^log_n = ROW * COL field elements):
//   pts    N Fr (mont), column-major: pts[c * ROW + r] = raw element (r, c)
//          (change_matrix_direction(LOG_COL, LOG_ROW), crates/amt/src/utils.rs:90-107)
//   coeff  N Fr, iNTT(pts) in bit-reversed order (DIF output)
//   y      N Fr, evaluations of one coset in natural order
//   The MSM scalar vector of a coset is ri[i] = y[bitrev(i)] (index_reverse,
//   crates/amt/src/utils.rs:70-88), gathered on the fly by k_msm_digits.
#pragma once
#include "common.h"
#include "curve29.h"
#include "keccak.h"

namespace zgr {

constexpr int MSM_C = 8;          // window width (bits)
constexpr int MSM_W = 32;         // windows: 254-bit scalars with signed 8-bit digits
constexpr int ACC_THREADS = 256;  // 32 windows x 8 lanes
constexpr int LANES_PER_WIN = 8;

// Fixed-base MSM: at most 2^FB_GL_MAX chunks share one bucket set.
constexpr int FB_GL_MAX = 2;  // (groups of 8 measured slower: their waves outlast the rest)

// log2 of the chunks that share one bucket set on line L of the AMT outputs: quotient
// line L < h is read only at tree level L + 1 (sums over groups of 2^(h-1-L) chunks),
// the basis line h at every level down to the chunks, the high line h + 1 at the root.
__host__ __device__ inline int fb_group_log(int L, int h) {
    int g = L < h ? h - 1 - L : (L == h ? 0 : h);
    return g < FB_GL_MAX ? g : FB_GL_MAX;
}

// Spread consecutive workgroup ids over the 8 XCDs so that each XCD works on a
// contiguous range of logical ids (and so on one L2-resident set of chunks).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t g, uint32_t total) {
    const uint32_t X = 8;
    uint32_t per = total / X;
    if (per == 0 || g >= per * X) return g;  // tail stays as is
    return (g % X) * per + g / X;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// ---------------------------------------------------------------------------
// constants
// ---------------------------------------------------------------------------
// TWO_ADIC_ROOT_OF_UNITY of Fr = 5^((r-1)/2^28), canonical
__device__ __forceinline__ fr g28_canon() {
    fr g;
    const uint32_t v[8] = {0x725b19f0u, 0x9bd61b6eu, 0x41112ed4u, 0x402d111eu,
                           0x8ef62abcu, 0x00e0a7ebu, 0xa58a7e85u, 0x2a3c09f0u};
#pragma unroll
    for (int i = 0; i < 8; i++) g.v[i] = v[i];
    return g;
}

__device__ __forceinline__ fr fr_from_u64(uint64_t x) {
    fr a = f_zero<FrP>();
    a.v[0] = (uint32_t)x;
    a.v[1] = (uint32_t)(x >> 32);
    return f_to_mont(a);
}

// omega, omega^-1, N^-1 and the coset factors (generate.rs:128-137) of a 2^log_n domain
__global__ void k_setup_consts(uint32_t* out, int log_n, int coset_n) {
    if (threadIdx.x || blockIdx.x) return;
    fr g = f_to_mont(g28_canon());
    fr w = g;
    for (int i = log_n; i < 28; i++) w = f_sqr(w);
    f_store(out + 8 * K_OMEGA, w);
    f_store(out + 8 * K_OMEGA_INV, f_inv(w));
    f_store(out + 8 * K_NINV, f_inv(fr_from_u64(1ull << log_n)));
    for (int c = 0; c < coset_n; c++) {
        uint32_t pw = bitrev_bits((uint32_t)c, 28 - log_n);
        f_store(out + 8 * (K_CF + c), f_pow_u64(g, pw));
    }
}

// tw[k] = base^k for k < count (base = consts[idx])
__global__ void k_powers(uint32_t* tw, const uint32_t* consts, int idx, uint32_t count) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    fr b = f_load<FrP>(consts + 8 * idx);
    fr r = f_one<FrP>();
    for (int i = 31; i >= 0; i--) {
        r = f_sqr(r);
        if ((k >> i) & 1) r = f_mul(r, b);
    }
    f_store(tw + 8 * (size_t)k, r);
}

// scale[p] = N^-1 * cf^bitrev(p) (2^256 form; the context converts it to the 2^261 form
// the NTT passes multiply by): coefficient bitrev(p) sits at DIF position p
__global__ void k_coset_scale_table(uint32_t* scale, const uint32_t* consts, int coset, int log_n) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (1u << log_n)) return;
    fr cf = f_load<FrP>(consts + 8 * (K_CF + coset));
    uint32_t e = bitrev_bits(p, log_n);
    fr r = f_one<FrP>();
    for (int i = log_n - 1; i >= 0; i--) {
        r = f_sqr(r);
        if ((e >> i) & 1) r = f_mul(r, cf);
    }
    f_store(scale + 8 * (size_t)p, f_mul(r, f_load<FrP>(consts + 8 * K_NINV)));
}

// ---------------------------------------------------------------------------
// packing: RawData::try_from + RawBlob::from (raw_data.rs:55-73, raw_blob.rs:33-40,
// utils.rs:10-15), written column-major (utils.rs:90-107)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint8_t raw_byte(const uint8_t* raw, size_t len, size_t max_raw, size_t q) {
    if (q < len) return raw[q];
    if (q >= max_raw) return (uint8_t)((uint32_t)len >> (8 * (q - max_raw)));
    return 0;
}

__global__ void k_pack(const uint8_t* raw, size_t len, int log_col, int log_row, uint32_t* pts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << (log_col + log_row);
    if (i >= n) return;
    const size_t max_raw = 31 * (size_t)n - 4;
    size_t off = 31 * (size_t)i;
    fr a;
#pragma unroll
    for (int l = 0; l < 8; l++) {
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int q = 4 * l + k;
            if (q < 31) v |= (uint32_t)raw_byte(raw, len, max_raw, off + q) << (8 * k);
        }
        a.v[l] = v;
    }
    uint32_t r = i >> log_col, c = i & ((1u << log_col) - 1);
    f_store(pts + 8 * (size_t)((c << log_row) | r), f_to_mont(a));
}

// ---------------------------------------------------------------------------
// MSM (multiple_multiexp semantics, prove_gpu.rs:77-84)
// ---------------------------------------------------------------------------
// Signed 8-bit digits of every scalar: digit[(b * 32 + w) * bs + k] for scalar
// ri[b * bs + k] = y[bitrev(b * bs + k)] (converted out of Montgomery form).
// from_y != 0: gather from y through the bit reversal; else read scalars in order.
__global__ void k_msm_digits(const uint32_t* y, int log_n, int log_bs, int8_t* digits, int from_y) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << log_n)) return;
    uint32_t src = from_y ? bitrev_bits(i, log_n) : i;
    fr s = f_from_mont(f_load<FrP>(y + 8 * (size_t)src));
    uint32_t b = i >> log_bs, k = i & ((1u << log_bs) - 1);
    int8_t* out = digits + ((size_t)b * MSM_W << log_bs) + k;
    int carry = 0;
#pragma unroll
    for (int w = 0; w < MSM_W; w++) {
        int d = (int)((s.v[w >> 2] >> (8 * (w & 3))) & 0xffu) + carry;
        carry = d >= 128;
        d -= carry << 8;  // d in [-128, 127]
        out[(size_t)w << log_bs] = (int8_t)d;
    }
}

// Counting sort of one (chunk, window) by |digit| into buckets 1..128.
// entries: k | sign << 15 | bucket << 16, ascending bucket; count = nonzero digits.
__global__ void k_msm_sort(const int8_t* digits, int log_bs, uint32_t* entries, uint32_t* counts) {
    __shared__ uint32_t hist[130];
    __shared__ uint32_t start[130];
    const uint32_t bw = blockIdx.x;  // b * 32 + w
    const uint32_t bs = 1u << log_bs;
    const int8_t* d = digits + ((size_t)bw << log_bs);
    for (uint32_t k = threadIdx.x; k < 130; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < bs; k += blockDim.x) {
        int v = d[k];
        if (v) atomicAdd(&hist[v < 0 ? -v : v], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int k = 1; k <= 128; k++) {
            start[k] = acc;
            acc += hist[k];
        }
        counts[bw] = acc;
    }
    __syncthreads();
    uint32_t* ent = entries + ((size_t)bw << log_bs);
    for (uint32_t k = threadIdx.x; k < bs; k += blockDim.x) {
        int v = d[k];
        if (!v) continue;
        uint32_t mag = v < 0 ? -v : v;
        uint32_t pos = atomicAdd(&start[mag], 1u);
        ent[pos] = k | ((uint32_t)(v < 0) << 15) | (mag << 16);
    }
}

// Uniform small scalar multiple: k * p for k < 2^bits (double-and-add without
// lane-dependent control flow).
__device__ __forceinline__ g1_xyzz xyzz_mul_small_uniform(const g1_xyzz& p, uint32_t k, int bits) {
    g1_xyzz r = xyzz_inf();
    for (int b = bits - 1; b >= 0; b--) {
        r = xyzz_dbl(r);
        g1_xyzz s = xyzz_add(r, p);
        r = xyzz_select((k >> b) & 1, s, r);
    }
    return r;
}

// Bucket accumulation of one (line L, chunk b): per window w the 8 lanes split the
// window's sorted entries evenly and each computes sum_e bucket(e) * P_e with the
// running-sum trick, one uniform XYZZ addition per step:
//   consume an entry:   R += +-P_e
//   close a bucket:     T += R
// giving T = sum_{j=lo..hi} R_j, so the lane's share is T + (lo - 1) * R.
// Point arithmetic runs in the reduced-radix field (field29.h); the chunk's bases
// are converted once while being staged in LDS (72 B per point).
// The per-window sums S_w land in heap leaf nb + b of tree (L, w).
constexpr int LDS_PT = 18;  // u32 per staged affine point (2 x 9 limbs)
constexpr int LDS_XP = 36;  // u32 per XYZZ point in the lane reduction

__global__ __launch_bounds__(ACC_THREADS, 2) void k_msm_accum(const uint32_t* bases, const uint32_t* entries,
                                                           const uint32_t* counts, uint32_t* heap,
                                                           int log_n, int h, int nlines) {
    extern __shared__ uint32_t lds[];
    const uint32_t nb = 1u << h;
    const int log_bs = log_n - h;
    const uint32_t bs = 1u << log_bs;
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t b = logical / nlines, L = logical % nlines;
    const int tid = threadIdx.x;
    const int w = tid >> 3, s = tid & 7;
    uint32_t* leaf = heap + ((((size_t)L * MSM_W + w) * 2 * nb) + nb + b) * X29_WORDS;

    // zero chunk: every window empty -> all sums are infinity
    const int nz = __syncthreads_or(tid < MSM_W && counts[b * MSM_W + tid] != 0);
    if (!nz) {
        if (s == 0) x29_store(leaf, x29_inf());
        return;
    }

    // stage the chunk's bases in LDS, converted to the reduced-radix form
    const uint32_t* src = bases + (((size_t)L << log_n) + ((size_t)b << log_bs)) * 16;
    for (uint32_t k = tid; k < bs; k += ACC_THREADS) {
        fq x = f_load<FqP>(src + 16 * (size_t)k), y = f_load<FqP>(src + 16 * (size_t)k + 8);
        fq29 qx = q_from_fq(x), qy = q_from_fq(y);
#pragma unroll
        for (int i = 0; i < 9; i++) {
            lds[k * LDS_PT + i] = qx.l[i];
            lds[k * LDS_PT + 9 + i] = qy.l[i];
        }
    }
    __syncthreads();

    const int cnt = (int)counts[b * MSM_W + w];
    const uint32_t* ent = entries + (((size_t)b * MSM_W + w) << log_bs);
    const int begin = (s * cnt) >> 3, end = ((s + 1) * cnt) >> 3;
    int hi = 0, lo = 1, steps = 0;
    if (end > begin) {
        hi = (int)(ent[end - 1] >> 16);
        lo = (int)(ent[begin] >> 16);
        steps = (end - begin) + (hi - lo + 1);
    }
    const int maxsteps = wave_max(steps);
    x29 R = x29_inf(), T = x29_inf();
    int e = end - 1, kc = hi;
    for (int st = 0; st < maxsteps; st++) {
        const bool active = st < steps;
        const bool have = e >= begin;
        const uint32_t ev = have ? ent[e] : 0u;
        const bool consume = active && have && (int)(ev >> 16) == kc;
        const bool trans = active && !consume;
        a29 P;
        const uint32_t* lp = lds + (ev & 0x7fffu) * LDS_PT;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            P.x.l[i] = lp[i];
            P.y.l[i] = lp[9 + i];
        }
        if (ev & 0x8000u) P.y = q_neg(P.y);
        // R is always one of the two operands (A on consume or idle, B on a bucket
        // close), so only A, B and T stay live across the addition
        x29 A = x29_select(trans, T, R);
        x29 B = x29_select(trans, R, x29_from_aff(P));
        x29 C = x29_add(A, B);
        R = x29_select(consume, C, x29_select(trans, B, A));
        T = x29_select(trans, C, T);
        e -= consume ? 1 : 0;
        kc -= trans ? 1 : 0;
    }
    x29 W = x29_add(T, x29_mul_small_uniform(R, (uint32_t)(lo - 1), 7));

    // sum the 8 lanes of each window
    __syncthreads();
    x29_store(lds + tid * LDS_XP, W);
    __syncthreads();
    for (int off = 1; off < LANES_PER_WIN; off <<= 1) {
        if ((s & (2 * off - 1)) == 0) {
            W = x29_add(W, x29_load(lds + (tid + off) * LDS_XP));
            x29_store(lds + tid * LDS_XP, W);
        }
        __syncthreads();
    }
    if (s == 0) x29_store(leaf, W);
}

// One level of the per-(line, window) binary trees over the chunks:
// heap[i] = heap[2i] + heap[2i+1] for i in [2^d, 2^(d+1)).  plan: 4 bits per line, the
// log2 chunk-group size gl of its bucket sets, which already produced its level h - gl
// (left as it is); 0 = every line from the chunks.
__global__ __launch_bounds__(64) void k_msm_tree_level(uint32_t* heap, int h, int d, uint32_t total, int nw,
                                                       uint64_t plan) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t nb = 1u << h;
    uint32_t lw = t >> d;
    if (d >= h - (int)((plan >> (4 * (lw / (uint32_t)nw))) & 15u)) return;
    uint32_t i = (1u << d) + (t & ((1u << d) - 1));
    uint32_t* base = heap + (size_t)lw * 2 * nb * X29_WORDS;
    x29 a = x29_load(base + (size_t)(2 * i) * X29_WORDS);
    x29 c = x29_load(base + (size_t)(2 * i + 1) * X29_WORDS);
    x29_store(base + (size_t)i * X29_WORDS, x29_add(a, c));
}

// sum_w 2^(8w) S_w over the nw windows of heap node `node` of line `line` (heap trees
// (line, w) of 2nb nodes each); nw == 1: the node itself
__device__ __forceinline__ x29 horner_windows(const uint32_t* heap, uint32_t nb, int nw, uint32_t line,
                                              uint32_t node) {
    x29 acc = x29_load(heap + (((size_t)line * nw + nw - 1) * 2 * nb + node) * X29_WORDS);
    for (int w = nw - 2; w >= 0; w--) {
        for (int k = 0; k < MSM_C; k++) acc = x29_dbl(acc);
        acc = x29_add(acc, x29_load(heap + (((size_t)line * nw + w) * 2 * nb + node) * X29_WORDS));
    }
    return acc;
}

// Window recombination sum_w 2^(8w) S_w for every output node, then canonical affine.
// Output slots (T = 2nb - 2):  [0, T) proof_tree  (line d-1, node i = slot + 2, d = log2 i)
//                              [T, 2T) comm_tree  (line h, node i = slot - T + 2)
//                              2T commitment (line h, node 1), 2T + 1 high (line h + 1, node 1)
__global__ __launch_bounds__(64) void k_msm_horner(const uint32_t* heap, int h, int nw, uint32_t* out,
                                                   uint32_t nslots) {
    const uint32_t nb = 1u << h;
    const uint32_t T = 2 * nb - 2;
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots || slot >= 2 * T + 2) return;
    uint32_t line, node;
    if (slot < T) {
        node = slot + 2;
        line = 31 - __builtin_clz(node) - 1;
    } else if (slot < 2 * T) {
        node = slot - T + 2;
        line = h;
    } else if (slot == 2 * T) {
        node = 1;
        line = h;
    } else {
        node = 1;
        line = h + 1;
    }
    x29_store_canon(out + (size_t)slot * 16, horner_windows(heap, nb, nw, line, node));
}

// multiple_multiexp output: out[L * nb + b] = sum_w 2^(8w) heap leaf (L, w, nb + b)
__global__ __launch_bounds__(64) void k_msm_leaves_out(const uint32_t* heap, int h, int nw, int nlines,
                                                       uint32_t* out) {
    const uint32_t nb = 1u << h;
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlines * nb) return;
    uint32_t L = t >> h, b = t & (nb - 1);
    x29_store_canon(out + (size_t)t * 16, horner_windows(heap, nb, nw, L, nb + b));
}

// single MSM result: Horner over the nw windows of tree node 1 of line 0
__global__ __launch_bounds__(64) void k_msm_root_out(const uint32_t* heap, int h, int nw, uint32_t* out) {
    if (threadIdx.x || blockIdx.x) return;
    x29_store_canon(out, horner_windows(heap, 1u << h, nw, 0, 1));
}

// ---------------------------------------------------------------------------
// Fixed-base MSM over the resident prove bases
// ---------------------------------------------------------------------------
// The prove bases do not change from blob to blob, so every base P is kept with its
// shifted copies P_j = 2^(c j) P, j < W (W c-bit windows cover a scalar).  A scalar with
// signed c-bit digits s = sum_j d_j 2^(c j) contributes sum_j |d_j| (sign(d_j) P_j): a
// chunk MSM (line L, chunk b) becomes ONE bucket sum over W * bs (point, digit) pairs
// into 2^(c-1) buckets, with no window recombination (no doublings, one tree per line).
// The digits, and so the sorted entries, are shared by every line.
//
// Two window widths:
//  * narrow, c = 12 (2048 buckets, 22 digits): one wave sums one (line, group of up to
//    4 chunks); about 24.6k additions per 1024-point chunk instead of 32 windows x
//    (1024 + 128) = 36.9k for windowed Pippenger;
//  * wide, c = 16 (32768 buckets, 16 digits): for the lines whose AMT outputs are sums
//    over groups of at least 16 chunks (quotient lines 0..h-5 and the high line).  A
//    bucket set per group then costs 2 x 32768 additions against 16 x 1024 x G entries,
//    and 16 digits instead of 22 cut the entries by 27 %.  Their waves take bucket
//    segments (ranges of the sorted list that never split a (bucket, group) run), so a
//    wave's length does not grow with the group.
//
// Table layout (16 u32 per point: x, y as 256-bit integers holding the 2^261-
// Montgomery value, canonical): point (slot, b, j, k) at ((slot * nb + b) * W + j) * bs
// + k, i.e. a chunk's W * bs points are contiguous.  The narrow table holds every line
// (slot = L), the wide table the wide lines (slot = their order).
constexpr int FB_C = 12;
constexpr int FB_W = 22;  // ceil(255 / 12): the top digit (bits 252..253 + carry) never carries out
constexpr int FB_BUCKETS = 1 << (FB_C - 1);  // |digit| in 1..2048
constexpr int FB_LANES = 64;                 // one wave per work item
constexpr int FB_MAX_LOG_BS = 10;
constexpr int FW_C = 16;
constexpr int FW_W = 16;  // 16 x 16 bits: the top digit (bits 240..253 + carry) never carries out
constexpr int FW_W_LOG = 4;
constexpr int FW_BUCKETS = 1 << (FW_C - 1);  // |digit| in 1..32768
constexpr int FW_BLK_LOG = 2;                // chunks per block of the wide sort: 4
constexpr int FW_MIN_GL = 4;                 // a line goes wide when its groups hold >= 16 chunks
constexpr int FW_MAX_LINES = 16;
constexpr int FW_RSEG = 2048;                // buckets per reduction wave
constexpr int FW_RSEGS = FW_BUCKETS / FW_RSEG;
constexpr int FW_SORT_T = 1024;  // threads per block of the wide sort (one 4-chunk block each)
constexpr int FW_TILE = 128;     // buckets per block of the offset scan
constexpr int FW_PARTS = 4;      // threads per bucket there (each sums a quarter of the blocks)
constexpr int FW_TILES = FW_BUCKETS / FW_TILE;
// Sorted entry (both widths, 8 bytes): x = point index within the wave's table region
// (narrow: (g * FB_W + j) * bs + k for chunk g of the group; wide: (b * FW_W + j) * bs +
// k) | sign << 31; y = |digit| << 16 | the block of the wide sort (0 in narrow lists).
// An entry's key (its bucket within its group) is y with the low bits below the group's
// block granularity cleared: every key is >= 2^16, never 0.
constexpr uint32_t FB_SIGN_BIT = 0x80000000u;
constexpr int FB_BSTART = FB_BUCKETS + 2;  // bucket starts 1..FB_BUCKETS, then the count
// Narrow entry as staged in k_fb_scatter's LDS: index (< 4 * 22 * 1024 < 2^18) | sign
// << 18 | |digit| << 19.
constexpr int FB_SIGN_SHIFT = 18;
constexpr int FB_MAG_SHIFT = 19;
constexpr uint32_t FB_IDX_MASK = (1u << FB_SIGN_SHIFT) - 1;

// per-level narrow sorted entries: level gl groups 2^gl consecutive chunks
struct FbLevels {
    const uint2* ent[FB_GL_MAX + 1];     // (nb >> gl) groups x 2^gl * FB_W * bs entries
    const uint32_t* cnt[FB_GL_MAX + 1];  // entries per group
    const uint32_t* bst[FB_GL_MAX + 1];  // bucket starts per group (FB_BSTART each)
};

// The wide lines of one launch: the shared sorted list, its bucket offsets, and where
// each wide line w keeps its segments, bucket sums and reduction partials.
struct FbWide {
    const uint2* ent;       // entries sorted by (|digit|, sort block)
    const uint32_t* off;    // off[blk * FW_BUCKETS + m - 1]: first entry of (m, blk)
    const uint32_t* tot;    // tot[0]: number of entries
    const uint32_t* seg;    // line w's segment boundaries at seg[segbase[w] + w + s], s <= nseg
    const uint32_t* table;  // wide shifted copies, slot w
    uint32_t* bkt;          // bucket sums: (gbase[w] + S) * FW_BUCKETS + m - 1
    uint32_t* hand;         // per wide sums wave: 2 x 64 lane partials
    uint32_t* part;         // reduction partials: (gbase[w] + S) * FW_RSEGS + q
    uint32_t nblk;          // blocks of the wide sort: nb >> FW_BLK_LOG
    int nw;
    // Per-line fields, indexed by a wave's line w: 32-bit each, so that every scalar load
    // of one has a dword-aligned base (a byte array here let the compiler fold a byte
    // offset into the base of a dword load, whose low address bits the scalar unit drops).
    uint32_t segbase[FW_MAX_LINES + 1];  // prefix of the lines' segment counts
    uint32_t gbase[FW_MAX_LINES + 1];    // prefix of the lines' group counts
    uint32_t line[FW_MAX_LINES];         // AMT line of wide line w
    uint32_t sh[FW_MAX_LINES];           // log2(sort blocks per group) = gl - FW_BLK_LOG
};

// Work items of k_fb_sums / k_fb_reduce (x; y = the wave's narrow bucket region or wide
// hand-off slot): narrow L | gl << 8 | S << 10; wide sums FB_WIDE | w | s << 8; wide
// reduction FB_WIDE | w | q << 4 | S << 8.
constexpr uint32_t FB_WIDE = 0x80000000u;

// Wave id: without a work list every wave has the same work, and consecutive ids (the
// lines of one chunk, which read the same sorted entries) are kept on one XCD; a work
// list is sorted longest first, so its entries are dealt round-robin to the XCDs as
// dispatched, giving each XCD the same mix of wave lengths.
__device__ __forceinline__ uint32_t fb_wave_id(const uint2* work) {
    return work ? blockIdx.x : xcd_remap(blockIdx.x, gridDim.x);
}

// wave -> (item, region): from the work list, or (no list) every line at level 0 with
// the wave's own region
__device__ __forceinline__ uint2 fb_item(const uint2* work, uint32_t wid, int lines) {
    if (work) return work[wid];
    return make_uint2((wid % (uint32_t)lines) | ((wid / (uint32_t)lines) << 10), wid);
}

__device__ __forceinline__ fq29 fb_load_coord(const uint32_t* s) {
    const uint4* p = (const uint4*)s;
    uint4 a = p[0], b = p[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return q_unpack(w);
}

__device__ __forceinline__ void fb_store29(uint32_t* d, const fq29& v) {  // 256-bit integer, 8 u32
    uint32_t w[8];
    q_pack(v, w);
    uint4* o = (uint4*)d;
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// Build the shifted copies of bases first .. first + count - 1 of line src_line into
// table slot dst_slot, one thread per base: P_0 = P, P_(j+1) = 2^C P_j, each stored
// affine.  The chain runs in XYZZ; the W - 1 projective copies and the running products
// of their ZZZ go to `scratch` (count * (W - 1) * 40 u32), so that one inversion per base
// (Montgomery's trick over its copies) converts all of them.  ZZ^3 = ZZZ^2 holds along
// the chain (ZZ = Z^2, ZZZ = Z^3), so 1/ZZ = (ZZ / ZZZ)^2.
template <int C, int W>
__global__ __launch_bounds__(64) void k_fb_build(const uint32_t* bases, uint32_t* table, uint32_t* scratch,
                                                 int log_n, int h, uint32_t src_line, uint32_t dst_slot,
                                                 uint32_t first, uint32_t count) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t i = first + t;
    const int log_bs = log_n - h;
    const uint32_t b = i >> log_bs, k = i & ((1u << log_bs) - 1);
    uint32_t* dst = table + ((((((size_t)dst_slot << h) + b) * W) << log_bs) + k) * 16;
    const size_t stride = (size_t)16 << log_bs;  // next copy of the same point
    auto slot = [&](int j) { return scratch + ((size_t)(j - 1) * count + t) * 40; };  // X, Y, ZZ, ZZZ, prefix
    const g1_aff a = aff_load(bases + (((size_t)src_line << log_n) + i) * 16);
    a29 P;
    P.x = q_canon(q_from_fq(a.x));
    P.y = q_canon(q_from_fq(a.y));
    fb_store29(dst, P.x);
    fb_store29(dst + 8, P.y);
    x29 q = x29_from_aff(P);
    fq29 acc = q_const(Q29::ONE);
    for (int j = 1; j < W; j++) {
        for (int s = 0; s < C; s++) q = x29_dbl(q);  // never infinity: 2^(C j) is not a multiple of r
        acc = q_mul(acc, q.ZZZ);
        uint32_t* sl = slot(j);
        fb_store29(sl, q.X);
        fb_store29(sl + 8, q.Y);
        fb_store29(sl + 16, q.ZZ);
        fb_store29(sl + 24, q.ZZZ);
        fb_store29(sl + 32, acc);  // prod_{1..j} ZZZ
    }
    fq29 inv = q_inv(acc);  // 1 / prod_{1..W-1} ZZZ
    for (int j = W - 1; j >= 1; j--) {
        const uint32_t* sl = slot(j);
        const fq29 X = fb_load_coord(sl), Y = fb_load_coord(sl + 8), ZZ = fb_load_coord(sl + 16),
                   ZZZ = fb_load_coord(sl + 24);
        const fq29 before = j > 1 ? fb_load_coord(slot(j - 1) + 32) : q_const(Q29::ONE);
        const fq29 izzz = q_mul(inv, before);  // 1 / ZZZ_j
        inv = q_mul(inv, ZZZ);                 // 1 / prod_{1..j-1} ZZZ
        const fq29 izz = q_sqr(q_mul(ZZ, izzz));
        fb_store29(dst + j * stride, q_canon(q_mul(X, izz)));
        fb_store29(dst + j * stride + 8, q_canon(q_mul(Y, izzz)));
    }
}

// Signed C-bit digits of scalar i (ri[i] = y[bitrev(i)] with from_y, else y[i]),
// converted out of Montgomery form: s = sum_w d_w 2^(C w), d_w in [-2^(C-1) + 1, 2^(C-1)].
template <int C, int W>
__device__ __forceinline__ void fb_digits(const uint32_t* y, int log_n, int from_y, uint32_t i, int dig[W]) {
    fr s = f_from_mont(f_load<FrP>(y + 8 * (size_t)(from_y ? bitrev_bits(i, log_n) : i)));
    int carry = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
        const int bit = C * w, wi = bit >> 5, sh = bit & 31;
        uint32_t v = s.v[wi] >> sh;
        if (sh > 32 - C && wi + 1 < 8) v |= s.v[wi + 1] << (32 - sh);
        int d = (int)(v & ((1u << C) - 1)) + carry;  // [0, 2^C]
        carry = d > (1 << (C - 1));
        dig[w] = d - (carry << C);
    }
}

// Counting sort of the narrow digits by |digit|, in two passes over each chunk b (one
// thread per scalar, blockDim == bs <= 1024):
//   k_fb_count    bucket sizes -> bstart[b * FB_BSTART + m] = first entry of bucket m
//                 (m = 1..FB_BUCKETS), [.. + FB_BUCKETS + 1] = count, counts[b] = count;
//   k_fb_scatter  entries at level 0 (the chunk's list, ascending |digit|) and at every
//                 chunk-group level gl <= gl_used: group S = b >> gl lists bucket m of
//                 its chunks one after the other (chunk g's indices offset by g * FB_W *
//                 bs, its table region), so the group's bucket start is the sum of its
//                 chunks' bucket starts.
//                 The chunk's list is assembled in LDS, FB_STAGE slots at a time (a
//                 window of the sorted order; the slots stay in registers meanwhile),
//                 and each window is copied out: at level 0 as one contiguous block, at
//                 a group level one bucket run at a time, so the global writes are runs
//                 of consecutive words rather than single words at random slots.  The
//                 window keeps the block's LDS near 47 KB, so that the scatter of one
//                 coset still finds room on CUs that run another coset's bucket kernels.
__global__ __launch_bounds__(1024) void k_fb_count(const uint32_t* y, int log_n, int log_bs, int from_y,
                                                   uint32_t* counts, uint32_t* bstart) {
    __shared__ uint32_t hist[FB_BUCKETS + 1];  // by |digit|; [0] unused
    __shared__ uint32_t wsum[16];
    const uint32_t b = blockIdx.x, k = threadIdx.x, T = blockDim.x;
    for (uint32_t q = k; q <= FB_BUCKETS; q += T) hist[q] = 0;
    __syncthreads();
    int dig[FB_W];
    fb_digits<FB_C, FB_W>(y, log_n, from_y, (b << log_bs) + k, dig);
#pragma unroll
    for (int w = 0; w < FB_W; w++)
        if (dig[w]) atomicAdd(&hist[dig[w] < 0 ? -dig[w] : dig[w]], 1u);
    __syncthreads();
    // exclusive prefix over the buckets: each thread owns FB_BUCKETS / T consecutive bins
    const uint32_t per = FB_BUCKETS / T;
    uint32_t loc = 0;
    for (uint32_t q = 0; q < per; q++) loc += hist[1 + k * per + q];
    uint32_t inc = loc;
    const int lane = (int)(k & 63);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
    }
    if (lane == 63 || k == T - 1) wsum[k >> 6] = inc;
    __syncthreads();
    if (k == 0) {
        uint32_t acc = 0;
        for (uint32_t w = 0; w < (T + 63) / 64; w++) {
            uint32_t t = wsum[w];
            wsum[w] = acc;
            acc += t;
        }
        counts[b] = acc;
        bstart[(size_t)b * FB_BSTART + FB_BUCKETS + 1] = acc;
    }
    __syncthreads();
    uint32_t run = wsum[k >> 6] + inc - loc;
    for (uint32_t q = 0; q < per; q++) {
        bstart[(size_t)b * FB_BSTART + 1 + k * per + q] = run;
        run += hist[1 + k * per + q];
    }
}

constexpr int FB_STAGE = (FB_W << FB_MAX_LOG_BS) / 4;  // k_fb_scatter's LDS window (5632 slots)
constexpr uint32_t FB_NO_SLOT = 0xffffffffu;

// level-gl outputs of k_fb_scatter (level 0: ent only; cnt / bst come from k_fb_count)
struct FbOut {
    uint2* ent[FB_GL_MAX + 1];
    uint32_t* cnt[FB_GL_MAX + 1];
    uint32_t* bst[FB_GL_MAX + 1];
};

// staged narrow entry -> sorted entry, its index moved by `offset` (chunk g of a group)
__device__ __forceinline__ uint2 fb_entry(uint32_t e, uint32_t offset) {
    return make_uint2(((e & FB_IDX_MASK) + offset) | (((e >> FB_SIGN_SHIFT) & 1u) << 31),
                      (e >> FB_MAG_SHIFT) << 16);
}

__global__ __launch_bounds__(1024) void k_fb_scatter(const uint32_t* y, int log_n, int log_bs, int from_y,
                                                     const uint32_t* bstart, int gl_used, FbOut out) {
    __shared__ uint32_t run[FB_BUCKETS + 1];                // next level-0 slot of bucket m
    __shared__ uint32_t base[FB_GL_MAX][FB_BUCKETS + 1];    // level-gl slot - level-0 slot
    __shared__ uint32_t staged[FB_STAGE];                   // one window of the sorted list
    const uint32_t b = blockIdx.x, k = threadIdx.x, T = blockDim.x;
    const size_t cap = (size_t)FB_W << log_bs;  // entries (and table points) of one chunk
    const uint32_t* b0 = bstart + (size_t)b * FB_BSTART;
    for (uint32_t m = 1 + k; m <= FB_BUCKETS; m += T) {
        run[m] = b0[m];
        for (int gl = 1; gl <= gl_used; gl++) {
            const uint32_t G = 1u << gl, S = b >> gl, g = b & (G - 1);
            uint32_t start = 0, before = 0;  // group's bucket start; earlier chunks' part of it
            for (uint32_t q = 0; q < G; q++) {
                const uint32_t* bq = bstart + (size_t)(S * G + q) * FB_BSTART;
                start += bq[m];
                if (q < g) before += bq[m + 1] - bq[m];
            }
            base[gl - 1][m] = start + before - b0[m];
            if (g == 0) {
                out.bst[gl][(size_t)S * FB_BSTART + m] = start;
                if (m == FB_BUCKETS) {
                    uint32_t total = 0;
                    for (uint32_t q = 0; q < G; q++) total += bstart[(size_t)(S * G + q) * FB_BSTART + FB_BUCKETS + 1];
                    out.bst[gl][(size_t)S * FB_BSTART + FB_BUCKETS + 1] = total;
                    out.cnt[gl][S] = total;
                }
            }
        }
    }
    __syncthreads();
    int dig[FB_W];
    fb_digits<FB_C, FB_W>(y, log_n, from_y, (b << log_bs) + k, dig);
    uint32_t pos[FB_W];  // level-0 slot of each digit's entry, or FB_NO_SLOT
#pragma unroll
    for (int w = 0; w < FB_W; w++) {
        const int d = dig[w];
        pos[w] = d ? atomicAdd(&run[d < 0 ? -d : d], 1u) : FB_NO_SLOT;  // < count <= cap
    }
    const uint32_t count = b0[FB_BUCKETS + 1];
    for (uint32_t lo = 0; lo < count; lo += FB_STAGE) {  // block-uniform
        __syncthreads();  // the previous window is copied out (first pass: every slot taken)
#pragma unroll
        for (int w = 0; w < FB_W; w++) {
            const uint32_t q = pos[w] - lo;  // wraps below lo; FB_NO_SLOT - lo >= 2^31
            if (q < FB_STAGE) {
                const int d = dig[w];
                const uint32_t mag = d < 0 ? -d : d;
                staged[q] = (((uint32_t)w << log_bs) | k) | ((uint32_t)(d < 0) << FB_SIGN_SHIFT) |
                            (mag << FB_MAG_SHIFT);
            }
        }
        __syncthreads();
        const uint32_t hi = min(count, lo + (uint32_t)FB_STAGE);
        for (uint32_t i = lo + k; i < hi; i += T) out.ent[0][(size_t)b * cap + i] = fb_entry(staged[i - lo], 0);
        for (int gl = 1; gl <= gl_used; gl++) {
            const uint32_t S = b >> gl, g = b & ((1u << gl) - 1);
            uint2* dst = out.ent[gl] + ((size_t)S << gl) * cap;
            const uint32_t* bs_gl = base[gl - 1];
            for (uint32_t i = lo + k; i < hi; i += T) {
                const uint32_t e = staged[i - lo];
                // the index absorbs g * cap: j * bs + k + g * cap < 2^gl * cap < 2^18
                dst[bs_gl[e >> FB_MAG_SHIFT] + i] = fb_entry(e, g * (uint32_t)cap);
            }
        }
    }
}

// ---- the wide sort: entries by (|digit|, block of 4 chunks) over the whole coset ----
// k_fw_count     cnt[blk * FW_BUCKETS + m - 1] = entries of block blk in bucket m
// k_fw_tile_sums / k_fw_tile_scan / k_fw_offsets
//                exclusive prefix of the counts in (m, blk) order -> off (same layout),
//                tot[0] = number of entries
// k_fw_scatter   entries to their slots, one LDS cursor per bucket
// One block per sort block (FW_SORT_T threads, 4 * bs scalars), the bucket counters or
// cursors in LDS (128 KB).
__device__ __forceinline__ uint32_t fw_sort_first(uint32_t blk, int log_bs) {
    return blk << (FW_BLK_LOG + log_bs);
}

__global__ __launch_bounds__(FW_SORT_T) void k_fw_count(const uint32_t* y, int log_n, int log_bs, int from_y,
                                                        uint32_t* cnt) {
    extern __shared__ uint32_t hist[];  // FW_BUCKETS
    const uint32_t blk = blockIdx.x, t = threadIdx.x;
    for (uint32_t q = t; q < FW_BUCKETS; q += FW_SORT_T) hist[q] = 0;
    __syncthreads();
    const uint32_t first = fw_sort_first(blk, log_bs), ns = 1u << (FW_BLK_LOG + log_bs);
    for (uint32_t u = t; u < ns; u += FW_SORT_T) {
        int dig[FW_W];
        fb_digits<FW_C, FW_W>(y, log_n, from_y, first + u, dig);
#pragma unroll
        for (int j = 0; j < FW_W; j++)
            if (dig[j]) atomicAdd(&hist[(dig[j] < 0 ? -dig[j] : dig[j]) - 1], 1u);
    }
    __syncthreads();
    for (uint32_t q = t; q < FW_BUCKETS; q += FW_SORT_T) cnt[(size_t)blk * FW_BUCKETS + q] = hist[q];
}

// part pp of bucket m - 1 = q's total: the blocks [pp nblk / FW_PARTS, (pp + 1) nblk / FW_PARTS)
__device__ __forceinline__ uint32_t fw_part_total(const uint32_t* cnt, uint32_t nblk, uint32_t q, uint32_t pp) {
    const uint32_t b0 = pp * nblk / FW_PARTS, b1 = (pp + 1) * nblk / FW_PARTS;
    uint32_t s = 0;
    for (uint32_t blk = b0; blk < b1; blk++) s += cnt[(size_t)blk * FW_BUCKETS + q];
    return s;
}

// Inclusive prefix over a tile's FW_TILE buckets of their totals, from FW_PARTS partial
// totals per bucket (thread t holds part t / FW_TILE of bucket t % FW_TILE).  Every
// thread of the block calls it and gets its bucket's prefix; *total = the bucket's total.
__device__ __forceinline__ uint32_t fw_tile_prefix(uint32_t partial, uint32_t (*part)[FW_TILE], uint32_t* wtot,
                                                   uint32_t* total) {
    const uint32_t t = threadIdx.x, tq = t % FW_TILE, lane = t & 63;
    part[t / FW_TILE][tq] = partial;
    __syncthreads();
    uint32_t v = 0;
#pragma unroll
    for (int pp = 0; pp < FW_PARTS; pp++) v += part[pp][tq];
    *total = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o);
        if (lane >= (uint32_t)o) v += u;
    }
    if (t < FW_TILE && lane == 63) wtot[t >> 6] = v;
    __syncthreads();
    if (tq >= 64) v += wtot[0];
    return v;
}

__global__ __launch_bounds__(FW_TILE * FW_PARTS) void k_fw_tile_sums(const uint32_t* cnt, uint32_t nblk,
                                                                     uint32_t* tile_tot) {
    __shared__ uint32_t part[FW_PARTS][FW_TILE];
    __shared__ uint32_t wtot[2];
    const uint32_t tq = threadIdx.x % FW_TILE, pp = threadIdx.x / FW_TILE;
    uint32_t total;
    const uint32_t v = fw_tile_prefix(fw_part_total(cnt, nblk, blockIdx.x * FW_TILE + tq, pp), part, wtot, &total);
    if (threadIdx.x == FW_TILE - 1) tile_tot[blockIdx.x] = v;
}

__global__ __launch_bounds__(FW_TILES) void k_fw_tile_scan(const uint32_t* tile_tot, uint32_t* tile_base,
                                                           uint32_t* tot) {
    __shared__ uint32_t wtot[FW_TILES / 64];
    const uint32_t t = threadIdx.x, lane = t & 63;
    const uint32_t x = tile_tot[t];
    uint32_t v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o);
        if (lane >= (uint32_t)o) v += u;
    }
    if (lane == 63) wtot[t >> 6] = v;
    __syncthreads();
    for (uint32_t w = 0; w < (t >> 6); w++) v += wtot[w];
    tile_base[t] = v - x;
    if (t == FW_TILES - 1) tot[0] = v;
}

// thread (part pp, bucket q) writes the offsets of q's blocks in part pp: the bucket's
// exclusive start, plus the counts of the earlier parts, then a running sum
__global__ __launch_bounds__(FW_TILE * FW_PARTS) void k_fw_offsets(const uint32_t* cnt, uint32_t nblk,
                                                                   const uint32_t* tile_base, uint32_t* off) {
    __shared__ uint32_t part[FW_PARTS][FW_TILE];
    __shared__ uint32_t wtot[2];
    const uint32_t tq = threadIdx.x % FW_TILE, pp = threadIdx.x / FW_TILE;
    const uint32_t q = blockIdx.x * FW_TILE + tq;
    uint32_t total;
    const uint32_t incl = fw_tile_prefix(fw_part_total(cnt, nblk, q, pp), part, wtot, &total);
    uint32_t r = tile_base[blockIdx.x] + incl - total;
    for (uint32_t k = 0; k < pp; k++) r += part[k][tq];
    const uint32_t b0 = pp * nblk / FW_PARTS, b1 = (pp + 1) * nblk / FW_PARTS;
    for (uint32_t blk = b0; blk < b1; blk++) {
        const size_t a = (size_t)blk * FW_BUCKETS + q;
        const uint32_t c = cnt[a];
        off[a] = r;
        r += c;
    }
}

__global__ __launch_bounds__(FW_SORT_T) void k_fw_scatter(const uint32_t* y, int log_n, int log_bs, int from_y,
                                                          const uint32_t* off, uint2* ent) {
    extern __shared__ uint32_t cursor[];  // FW_BUCKETS
    const uint32_t blk = blockIdx.x, t = threadIdx.x;
    for (uint32_t q = t; q < FW_BUCKETS; q += FW_SORT_T) cursor[q] = off[(size_t)blk * FW_BUCKETS + q];
    __syncthreads();
    const uint32_t first = fw_sort_first(blk, log_bs), ns = 1u << (FW_BLK_LOG + log_bs);
    for (uint32_t u = t; u < ns; u += FW_SORT_T) {
        const uint32_t i = first + u, b = i >> log_bs, k = i & ((1u << log_bs) - 1);
        int dig[FW_W];
        fb_digits<FW_C, FW_W>(y, log_n, from_y, i, dig);
#pragma unroll
        for (int j = 0; j < FW_W; j++) {
            const int d = dig[j];
            if (!d) continue;
            const uint32_t mag = d < 0 ? -d : d;
            const uint32_t pos = atomicAdd(&cursor[mag - 1], 1u);
            ent[pos] = make_uint2(((((b << FW_W_LOG) | j) << log_bs) | k) | ((uint32_t)(d < 0) << 31),
                                  (mag << 16) | blk);
        }
    }
}

// first entry of (bucket m0 + 1, block blk) in the (m, blk) order; blk == nblk is the
// next bucket's first block, and past the last bucket the entry count
__device__ __forceinline__ uint32_t fw_pos(const FbWide& wd, uint32_t m0, uint32_t blk) {
    if (blk >= wd.nblk) {
        m0++;
        blk = 0;
    }
    return m0 >= (uint32_t)FW_BUCKETS ? wd.tot[0] : wd.off[(size_t)blk * FW_BUCKETS + m0];
}

// Segment boundaries of the wide lines' sums waves: boundary s of line w is the first
// (bucket, group) run starting at or after entry s * tot / nseg_w, so no run is split
// between waves.  One thread per boundary.
__global__ void k_fw_segments(FbWide wd, uint32_t* seg) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    int w = 0;
    while (w < wd.nw && t >= wd.segbase[w + 1] + w + 1) w++;
    if (w >= wd.nw) return;
    const uint32_t s = t - wd.segbase[w] - w, nseg = wd.segbase[w + 1] - wd.segbase[w];
    const uint32_t sh = wd.sh[w], ng = wd.nblk >> sh;  // groups of the line
    const uint32_t target = (uint32_t)(((uint64_t)s * wd.tot[0]) / nseg);
    // smallest key index i in [0, FW_BUCKETS * ng] with start(i) >= target
    uint32_t lo = 0, hi = (uint32_t)FW_BUCKETS * ng;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (fw_pos(wd, mid / ng, (mid % ng) << sh) >= target) hi = mid;
        else lo = mid + 1;
    }
    seg[t] = lo == (uint32_t)FW_BUCKETS * ng ? wd.tot[0] : fw_pos(wd, lo / ng, (lo % ng) << sh);
}

// Bucket sums, one wave per work item, two kinds of item in one launch:
//  * narrow (line L, chunk group S of level gl): the group's sorted list; its 2^gl
//    chunks share one bucket set, region y of `bkt`;
//  * wide (wide line w, segment s): entries [seg_s, seg_(s+1)) of the shared wide list,
//    whose (bucket, group) runs go to the line's bucket sets in FbWide::bkt.
// Lane s takes an even share of the wave's entries and adds each run's points with
// mixed additions (R += +-P_e).  A run belongs to the lane holding its first entry,
// which writes the run's sum to its bucket slot (empty buckets stay unwritten); a lane
// whose share starts inside a run hands that partial sum to the owner through global
// memory (the wave's first / last partials), so LDS does not bound occupancy.
// Splitting the work this way keeps every step the same addition on every lane.  X stays
// unreduced in [0, 8p) (x29_madd_unchecked).  The reduction (k_fb_reduce) is a kernel
// of its own so that the waves on a CU run one loop body at a time: the two bodies
// together overflow the instruction cache.
constexpr uint32_t FB_NO_SHARE = 0xffffffffu;  // lane with no entries
// XYZZ slots per narrow region in `bkt`: the buckets, then each lane's first- and
// last-bucket partials
constexpr int FB_REGION = FB_BUCKETS + 2 * FB_LANES;

__global__ __launch_bounds__(FB_LANES, 2) void k_fb_sums(const uint32_t* table, FbLevels lv, FbWide wd,
                                                         const uint2* work, uint32_t* bkt, int log_n, int h,
                                                         int lines) {
    __shared__ uint32_t head_k[FB_LANES];               // run continued from an earlier lane, 0 if none
    __shared__ uint32_t one_bucket[FB_LANES];           // the share lies inside one run
    const int log_bs = log_n - h;
    const uint2 it = fb_item(work, fb_wave_id(work), lines);
    const int s = threadIdx.x;
    const uint2* ent;
    const uint32_t* tab;
    uint32_t* bk;       // bucket slots: key m << 16 | blk -> bk + (((blk >> sh) << 15) + m - 1) * X29_WORDS
    uint32_t* first_p;  // lanes' continued-first partials, then their last partials
    uint32_t kmask, sh;
    int cnt;
    if (it.x & FB_WIDE) {
        const uint32_t w = it.x & 0xffu, sg = (it.x >> 8) & 0x7fffffu;
        const uint32_t* b0 = wd.seg + wd.segbase[w] + w + sg;
        const uint32_t lo = b0[0];
        cnt = (int)(b0[1] - lo);
        ent = wd.ent + lo;
        tab = wd.table + ((size_t)w << (h + FW_W_LOG + log_bs)) * 16;  // slot w: nb * FW_W * bs points
        sh = wd.sh[w];
        kmask = 0xffff0000u | (0xffffu & ~((1u << sh) - 1));
        bk = wd.bkt + (size_t)wd.gbase[w] * FW_BUCKETS * X29_WORDS;
        first_p = wd.hand + (size_t)it.y * 2 * FB_LANES * X29_WORDS;
    } else {
        const uint32_t L = it.x & 0xffu, gl = (it.x >> 8) & 3u, S = it.x >> 10;
        cnt = (int)lv.cnt[gl][S];
        const size_t cap = (size_t)FB_W << log_bs;  // entries (and table points) of one chunk
        ent = lv.ent[gl] + ((size_t)S << gl) * cap;
        tab = table + ((((size_t)L << h) + ((size_t)S << gl)) * cap) * 16;
        sh = 0;
        kmask = 0xffff0000u;
        bk = bkt + (size_t)it.y * FB_REGION * X29_WORDS;
        first_p = bk + FB_BUCKETS * X29_WORDS;
    }
    if (cnt == 0) return;  // zero group: k_fb_reduce writes infinity
    uint32_t* last_p = first_p + FB_LANES * X29_WORDS;
    auto slot = [&](uint32_t key) {
        return bk + (((size_t)((key & 0xffffu) >> sh) << (FW_C - 1)) + (key >> 16) - 1) * X29_WORDS;
    };

    // even share [begin, end) of the wave's entries
    const int begin = (s * cnt) / FB_LANES, end = ((s + 1) * cnt) / FB_LANES;
    const int steps = end - begin;
    const uint32_t k_first = steps > 0 ? ent[begin].y & kmask : 0u;
    const uint32_t k_last = steps > 0 ? ent[end - 1].y & kmask : 0u;
    const bool cont = steps > 0 && begin > 0 && (ent[begin - 1].y & kmask) == k_first;
    const int maxsteps = wave_max(steps);
    x29 R = x29_inf();
    uint32_t cur = 0;   // run key of R (0: none yet)
    bool head = true;   // R still holds the share's first run
    // software pipeline: the entry two steps ahead and the point one step ahead are in
    // flight while a step computes.  Past the end of its share a lane re-reads the
    // list's last entry (an index clamp, not a select on the loaded value, so the load
    // need not complete before the next one is issued); its results are discarded.
    auto entry_at = [&](int st) -> uint2 { return ent[min(begin + st, cnt - 1)]; };
    uint2 ev_next = entry_at(0), ev_after = entry_at(1);
    const uint4* pn = (const uint4*)(tab + (size_t)(ev_next.x & ~FB_SIGN_BIT) * 16);
    uint4 p0 = pn[0], p1 = pn[1], p2 = pn[2], p3 = pn[3];
    for (int st = 0; st < maxsteps; st++) {
        const bool active = st < steps;
        const uint2 ev = ev_next;
        const uint32_t m = ev.y & kmask;
        const uint32_t wx[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
        const uint32_t wy[8] = {p2.x, p2.y, p2.z, p2.w, p3.x, p3.y, p3.z, p3.w};
        a29 P;
        P.x = q_unpack(wx);
        P.y = q_unpack(wy);
        // issue the next step's loads before this step's arithmetic
        ev_next = ev_after;
        ev_after = entry_at(st + 2);
        pn = (const uint4*)(tab + (size_t)(ev_next.x & ~FB_SIGN_BIT) * 16);
        p0 = pn[0];
        p1 = pn[1];
        p2 = pn[2];
        p3 = pn[3];
        P.y = q_cneg_canon(P.y, ev.x >> 31);  // table points are canonical
        const bool same = m == cur;
        bool rare;
        x29 M = x29_madd_unchecked(R, P, &rare);
        // a point added to itself or its negative, or to a bucket that cancelled to
        // infinity: the exact formula, on a wave-uniform branch that is almost never taken
        if (__builtin_expect(__any(active && same && rare), 0))
            if (active && same && rare) M = x29_madd(x29_norm(R), P);
        if (active && !same && cur != 0) {  // run cur is complete in this share
            if (head && cont)
                x29_store(first_p + s * X29_WORDS, R);
            else
                x29_store(slot(cur), R);
            head = false;
        }
        R = x29_select(same, M, x29_from_aff(P));
        cur = m;
        if (active && st == steps - 1) x29_store(last_p + s * X29_WORDS, R);
    }
    const bool single = steps > 0 && head;  // never left the first run
    head_k[s] = steps == 0 ? FB_NO_SHARE : (cont ? k_first : 0u);
    one_bucket[s] = single;
    __syncthreads();
    // the owner of the share's last run adds the partials of the lanes that continue it
    if (steps > 0 && !(single && cont)) {
        x29 W = x29_load(last_p + s * X29_WORDS);
        for (int t = s + 1; t < FB_LANES; t++) {
            const uint32_t hk = head_k[t];
            if (hk == FB_NO_SHARE) continue;
            if (hk != k_last) break;
            W = x29_add(W, x29_load((one_bucket[t] ? last_p : first_p) + t * X29_WORDS));
            if (!one_bucket[t]) break;
        }
        x29_store(slot(k_last), W);
    }
}

// Bucket reduction, RL lanes per item (RL = 64: one item per wave; RL = 32: two per
// wave): 2048 consecutive buckets m of one bucket set (a narrow group's whole set;
// segment q of a wide group's), lane s of the item folding buckets (P s, P s + P] of
// them, P = 2048 / RL, from the top with the running sum (R += B_m; T += R).  Its share
// of sum_m m B_m is T + (first bucket - 1 + P s) R; the RL shares are summed through LDS
// into the group's heap node of line L (narrow) or the segment's partial (wide, summed by
// k_fw_combine).  Empty buckets count as infinity, and a wave whose buckets are all
// empty (zero groups of sparse scalars) does no work.  With RL = 32 the per-lane offset
// multiplication and tree are spread over twice the buckets: 13 % fewer field
// multiplications per bucket set, for dense scalars; RL = 64 halves an item's latency,
// for sparse ones, whose few non-zero groups leave most of the device idle.  `items`
// bounds the list (a last odd item leaves the second half idle).
// RB < 2048 (wide items only): items of RB buckets, FW_BUCKETS / RB partials per group
// (item FB_WIDE | w | q << 4 | S << 10), for a single MSM whose few groups would
// otherwise give too few waves to fill the device.
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

template <int RL, int RB = FB_BUCKETS>
__global__ __launch_bounds__(FB_LANES, 2) void k_fb_reduce(FbLevels lv, FbWide wd, const uint2* work,
                                                           uint32_t items, const uint32_t* bkt, uint32_t* heap,
                                                           int h, int lines) {
    constexpr int FB_PER_LANE = RB / RL;
    constexpr int PL_LOG = ilog2c(FB_PER_LANE);
    constexpr int RSEGS = FW_BUCKETS / RB;  // partials per wide group
    constexpr int Q_BITS = ilog2c(RSEGS) > 4 ? ilog2c(RSEGS) : 4;
    static_assert(FB_PER_LANE <= 64 && (1 << PL_LOG) == FB_PER_LANE, "lane share");
    __shared__ uint32_t red[FB_LANES * X29_WORDS];
    const uint32_t nb = 1u << h;
    const int half = (int)(threadIdx.x / RL), s = (int)(threadIdx.x % RL);
    const uint32_t idx = (FB_LANES / RL) * fb_wave_id(work) + (uint32_t)half;
    const bool live = idx < items;
    const uint2 it = live ? fb_item(work, idx, lines) : make_uint2(0u, 0u);
    uint32_t* outp = nullptr;
    const uint32_t* bk = bkt;  // the lane's first bucket
    uint32_t mul = 0;          // the lane's bucket offset / FB_PER_LANE
    uint64_t filled = 0;       // bit i: bucket i of the lane's FB_PER_LANE has entries
    const bool wide = (it.x & FB_WIDE) != 0;
    if (live && wide) {
        const uint32_t w = it.x & 0xfu, q = (it.x >> 4) & ((1u << Q_BITS) - 1u),
                       S = (it.x >> (4 + Q_BITS)) & (0x7fffffffu >> (4 + Q_BITS));
        const uint32_t g = wd.gbase[w] + S, sh = wd.sh[w];
        mul = q * (RB / FB_PER_LANE) + (uint32_t)s;
        const uint32_t m0 = mul * FB_PER_LANE;  // bucket m0 + 1 is the lane's first
        bk = wd.bkt + ((size_t)g * FW_BUCKETS + m0) * X29_WORDS;
        outp = wd.part + ((size_t)g * RSEGS + q) * X29_WORDS;
        for (int i = 0; i < FB_PER_LANE; i++)
            filled |= (uint64_t)(fw_pos(wd, m0 + i, (S + 1) << sh) > fw_pos(wd, m0 + i, S << sh)) << i;
    } else if (RB == FB_BUCKETS && live) {
        const uint32_t L = it.x & 0xffu, gl = (it.x >> 8) & 3u, S = it.x >> 10;
        outp = heap + ((size_t)L * 2 * nb + (nb >> gl) + S) * X29_WORDS;
        if (lv.cnt[gl][S]) {  // a zero group leaves `filled` empty: its sum is infinity
            const uint32_t* bst = lv.bst[gl] + (size_t)S * FB_BSTART + s * FB_PER_LANE + 1;  // bucket P s + 1
            bk = bkt + ((size_t)it.y * FB_REGION + s * FB_PER_LANE) * X29_WORDS;
            mul = (uint32_t)s;
            for (int i = 0; i < FB_PER_LANE; i++) filled |= (uint64_t)(bst[i + 1] > bst[i]) << i;
        }
    }
    // a wave whose buckets are all empty (zero groups of sparse scalars) skips the work:
    // its sums are infinity
    x29 W = x29_inf();
    if (__any(filled != 0)) {  // uniform over the wave (one wave per block)
        x29 Rr = x29_inf(), T = x29_inf();
        for (int i = FB_PER_LANE - 1; i >= 0; i--) {
            x29 B = x29_inf();
            if ((filled >> i) & 1u) B = x29_load(bk + (size_t)i * X29_WORDS);
            Rr = x29_add_lazy(Rr, B);
            T = x29_add_lazy(T, Rr);
        }
        // T + P mul Rr = (mul Rr) doubled log2 P times, plus T (mul < RSEGS RL; the bit
        // count is uniform over the wave, the multiplier per lane)
        x29 off = x29_mul_small_uniform(Rr, mul, ilog2c(RL) + (__any(wide) ? ilog2c(RSEGS) : 0));
        for (int k = 0; k < PL_LOG; k++) off = x29_dbl(off);
        W = x29_add_lazy(T, off);

        x29_store(red + threadIdx.x * X29_WORDS, W);
        __syncthreads();
        for (int o = 1; o < RL; o <<= 1) {
            if ((s & (2 * o - 1)) == 0) {
                W = x29_add_lazy(W, x29_load(red + (threadIdx.x + o) * X29_WORDS));
                x29_store(red + threadIdx.x * X29_WORDS, W);
            }
            __syncthreads();
        }
    }
    if (s == 0 && live) x29_store(outp, x29_norm(W));
}

// Group results of the wide lines: heap node (nb >> gl) + S of line L = the sum of the
// group's RSEGS reduction partials.  One wave per group, in group order.
template <int RSEGS = FW_RSEGS>
__global__ __launch_bounds__(FB_LANES) void k_fw_combine(FbWide wd, uint32_t* heap, int h) {
    static_assert(RSEGS <= FB_LANES, "one partial per lane");
    __shared__ uint32_t red[RSEGS * X29_WORDS];
    const uint32_t g = blockIdx.x, s = threadIdx.x;
    int w = 0;
    while (w + 1 < wd.nw && g >= wd.gbase[w + 1]) w++;
    const uint32_t S = g - wd.gbase[w], gl = wd.sh[w] + FW_BLK_LOG, nb = 1u << h;
    if (s < (uint32_t)RSEGS)
        x29_store(red + s * X29_WORDS, x29_load(wd.part + ((size_t)g * RSEGS + s) * X29_WORDS));
    __syncthreads();
    for (uint32_t o = 1; o < (uint32_t)RSEGS; o <<= 1) {
        if (s < (uint32_t)RSEGS && (s & (2 * o - 1)) == 0) {
            const x29 W = x29_add(x29_load(red + s * X29_WORDS), x29_load(red + (s + o) * X29_WORDS));
            x29_store(red + s * X29_WORDS, W);
        }
        __syncthreads();
    }
    if (s == 0)
        x29_store(heap + ((size_t)wd.line[w] * 2 * nb + (nb >> gl) + S) * X29_WORDS, x29_load(red));
}

// ---------------------------------------------------------------------------
// G1 EC-FFT (Pairing::fast_fft / fast_ifft, crates/amt/src/ec_algebra.rs:34-95):
// radix-2 DIT over XYZZ points, twiddle products by 254-bit scalar multiplication.
// ---------------------------------------------------------------------------
// canonical affine (zeros = infinity) -> XYZZ mont, at bit-reversed position
// EC-FFT points live as reduced-radix XYZZ (X29_WORDS u32 each) between the stages.
// canonical affine (64 zero bytes = infinity) -> x29
__device__ __forceinline__ x29 ec29_from_canon(const uint32_t* s) {
    fq x = f_load<FqP>(s), y = f_load<FqP>(s + 8);
    const bool inf = f_is_zero(x) && f_is_zero(y);
    a29 a;
    a.x = q_from_fq(f_to_mont(x));
    a.y = q_from_fq(f_to_mont(y));
    return x29_select(inf, x29_inf(), x29_from_aff(a));
}

__global__ void k_ec_load_bitrev(const uint32_t* in, uint32_t* pts, int log_n) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (1u << log_n)) return;
    x29_store(pts + X29_WORDS * (size_t)p, ec29_from_canon(in + 16 * (size_t)bitrev_bits(p, log_n)));
}

// k * P for a canonical scalar k < r by the GLV split of the BN254 endomorphism
// phi(x, y) = (beta x, y) = lambda (x, y): k = k1 + lambda k2 (mod r) with |k1|, |k2| <
// 2^127, and k P = k1 P + k2 phi(P).  Both halves share signed 3-bit (Booth) windows,
// digits in [-4, 4], over the affine multiples P, 2P, 3P, 4P (phi of a multiple is one
// multiplication of its x; a negative digit negates y): 43 x (3 doublings + 2 mixed
// additions), the additions selected away on zero digits, where the plain 254-bit scalar
// needs 127 x (2 doublings + 1 addition).  The multiples are made affine with one
// inversion and sit in LDS, thread-interleaved: a block of EC_THREADS threads needs
// EC_LDS_WORDS u32.
constexpr int EC_THREADS = 64;
constexpr int EC_TAB_POINTS = 4;
constexpr int EC_AFF_WORDS = 18;  // x, y: 9 limbs each
constexpr int EC_LDS_WORDS = EC_TAB_POINTS * EC_AFF_WORDS * EC_THREADS;

struct Glv {
    // lambda = 0xb3c4d79d41a917585bfc41088d8daaa78b17ea66b99c90dd (a cube root of unity mod r),
    // beta = 0x59e26bcea0d48bacd4f263f1acdb5c4f5763473177fffffe (mod p), in the 29-bit form
    static constexpr uint32_t BETA29[9] = {0x0a337995u, 0x158d1d23u, 0x189c9b98u, 0x12fa4e45u, 0x185faadcu,
                                           0x0176f16du, 0x0eed93bau, 0x14291140u, 0x000c0afeu};
    // a reduced basis of {(x, y) : x + lambda y = 0 mod r}: (a1, b1), (a2, b2) with b1 < 0,
    // determinant r; g1 = round(2^256 b2 / r), g2 = round(2^256 (-b1) / r)
    static constexpr uint32_t G1[3] = {0xc7e0b3d7u, 0xd91d232eu, 0x00000002u};
    static constexpr uint32_t G2[5] = {0x391eb18eu, 0x7a7bd9d4u, 0xa773d2cfu, 0x4ccef014u, 0x00000002u};
    static constexpr uint32_t A1[2] = {0x94d213e3u, 0x89d32568u};
    static constexpr uint32_t A2[4] = {0x1221250bu, 0x0be4e154u, 0xeeb859fdu, 0x6f4d8248u};
    static constexpr uint32_t NB1[4] = {0x7d4f1128u, 0x8211bbebu, 0xeeb859fcu, 0x6f4d8248u};  // -b1
    static constexpr uint32_t B2[2] = {0x94d213e3u, 0x89d32568u};
};

// out[0 .. NA + NB) = a * b (schoolbook with a carry per row)
template <int NA, int NB>
__device__ __forceinline__ void u32_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {
#pragma unroll
    for (int i = 0; i < NA + NB; i++) out[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const uint64_t t = (uint64_t)a[i] * b[j] + out[i + j] + c;
            out[i + j] = (uint32_t)t;
            c = t >> 32;
        }
        out[i + NB] = (uint32_t)c;
    }
}

// c = (k g + 2^255) >> 256 for the 8-limb k and an NG-limb g; the low NC limbs of c
template <int NG, int NC>
__device__ __forceinline__ void glv_round(const uint32_t* k, const uint32_t* g, uint32_t* c) {
    uint32_t t[8 + NG];
    u32_mul<8, NG>(k, g, t);
    uint64_t x = (uint64_t)t[7] + 0x80000000u;  // + 2^255
    uint32_t carry = (uint32_t)(x >> 32);
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const uint64_t y = (uint64_t)t[8 + i] + carry;
        c[i] = (uint32_t)y;
        carry = (uint32_t)(y >> 32);
    }
}

// (x mod 2^256, a two's-complement value of magnitude < 2^127) -> magnitude, sign
__device__ __forceinline__ bool glv_abs(const uint32_t* x, uint32_t* m) {
    const bool neg = x[7] >> 31;
    uint32_t c = neg ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t y = (uint64_t)(neg ? ~x[i] : x[i]) + c;
        m[i] = (uint32_t)y;
        c = (uint32_t)(y >> 32);
    }
    return neg;
}

// k (canonical, < r) -> |k1|, |k2| (< 2^127, 4 limbs each) and their signs, with
// k1 = k - c1 a1 - c2 a2 and k2 = c1 (-b1) - c2 b2 for c1 = round(k b2 / r),
// c2 = round(k (-b1) / r) (c1 < 2^64, c2 < 2^127)
__device__ __forceinline__ void glv_split(const fr& k, uint32_t* m1, bool* n1, uint32_t* m2, bool* n2) {
    uint32_t c1[2], c2[4];
    glv_round<3, 2>(k.v, Glv::G1, c1);
    glv_round<5, 4>(k.v, Glv::G2, c2);
    uint32_t p1[4], p2[8], q1[6], q2[6];
    u32_mul<2, 2>(c1, Glv::A1, p1);
    u32_mul<4, 4>(c2, Glv::A2, p2);
    u32_mul<2, 4>(c1, Glv::NB1, q1);
    u32_mul<4, 2>(c2, Glv::B2, q2);
    uint32_t x[8], y[8];
    int64_t bx = 0, by = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {  // x = k - p1 - p2, y = q1 - q2 (mod 2^256)
        const int64_t tx = (int64_t)k.v[i] - (i < 4 ? p1[i] : 0u) - p2[i] + bx;
        x[i] = (uint32_t)tx;
        bx = tx >> 32;
        const int64_t ty = (int64_t)(i < 6 ? q1[i] : 0u) - (i < 6 ? q2[i] : 0u) + by;
        y[i] = (uint32_t)ty;
        by = ty >> 32;
    }
    *n1 = glv_abs(x, m1);
    *n2 = glv_abs(y, m2);
}

__device__ __forceinline__ void ec_tab_store(uint32_t* tab, int k, const fq29& x, const fq29& y) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
        tab[(k * EC_AFF_WORDS + i) * EC_THREADS + threadIdx.x] = x.l[i];
        tab[(k * EC_AFF_WORDS + 9 + i) * EC_THREADS + threadIdx.x] = y.l[i];
    }
}

__device__ __forceinline__ a29 ec_tab_load(const uint32_t* tab, int k) {
    a29 a;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        a.x.l[i] = tab[(k * EC_AFF_WORDS + i) * EC_THREADS + threadIdx.x];
        a.y.l[i] = tab[(k * EC_AFF_WORDS + 9 + i) * EC_THREADS + threadIdx.x];
    }
    return a;
}

// Booth digit of a window from its nibble (b_{3i+2} b_{3i+1} b_{3i} b_{3i-1}):
// -4 b_{3i+2} + 2 b_{3i+1} + b_{3i} + b_{3i-1}
__device__ __forceinline__ int ec_booth(uint32_t nib) {
    return (int)(((nib >> 1) & 2u) + ((nib >> 1) & 1u) + (nib & 1u)) - 4 * (int)(nib >> 3);
}

// s = m << 31 over 160 bits: window 42's nibble (b_128 .. b_125) is the top of s[4], and
// each window after it comes up by a shift of 3
__device__ __forceinline__ void ec_booth_init(const uint32_t (&m)[4], uint32_t (&s)[5]) {
    s[0] = m[0] << 31;
    s[1] = (m[1] << 31) | (m[0] >> 1);
    s[2] = (m[2] << 31) | (m[1] >> 1);
    s[3] = (m[3] << 31) | (m[2] >> 1);
    s[4] = m[3] >> 1;
}

__device__ __forceinline__ uint32_t ec_booth_next(uint32_t (&s)[5]) {
    const uint32_t nib = s[4] >> 28;
#pragma unroll
    for (int i = 4; i > 0; i--) s[i] = (s[i] << 3) | (s[i - 1] >> 29);
    s[0] <<= 3;
    return nib;
}

__device__ __forceinline__ x29 x29_mul_scalar(const x29& p, const fr& k, uint32_t* tab) {
    uint32_t m1[4], m2[4];
    bool n1, n2;
    glv_split(k, m1, &n1, m2, &n2);
    // P, 2P, 3P, 4P, made affine with one inversion of the product of their ZZ ZZZ
    const x29 p2 = x29_dbl(p), p3 = x29_add(p2, p), p4 = x29_dbl(p2);
    const fq29 d0 = q_mul(p.ZZ, p.ZZZ), d1 = q_mul(p2.ZZ, p2.ZZZ), d2 = q_mul(p3.ZZ, p3.ZZZ),
               d3 = q_mul(p4.ZZ, p4.ZZZ);
    const fq29 c1 = q_mul(d0, d1), c2 = q_mul(c1, d2);
    fq29 inv = q_inv(q_mul(c2, d3));  // 0 when P is infinity (then the result is replaced)
    const fq29 i3 = q_mul(inv, c2);
    inv = q_mul(inv, d3);
    const fq29 i2 = q_mul(inv, c1);
    inv = q_mul(inv, d2);
    const fq29 i1 = q_mul(inv, d0), i0 = q_mul(inv, d1);
    // 1 / (ZZ ZZZ) = i: x = X ZZZ i, y = Y ZZ i
    ec_tab_store(tab, 0, q_mul(q_mul(p.X, p.ZZZ), i0), q_mul(q_mul(p.Y, p.ZZ), i0));
    ec_tab_store(tab, 1, q_mul(q_mul(p2.X, p2.ZZZ), i1), q_mul(q_mul(p2.Y, p2.ZZ), i1));
    ec_tab_store(tab, 2, q_mul(q_mul(p3.X, p3.ZZZ), i2), q_mul(q_mul(p3.Y, p3.ZZ), i2));
    ec_tab_store(tab, 3, q_mul(q_mul(p4.X, p4.ZZZ), i3), q_mul(q_mul(p4.Y, p4.ZZ), i3));
    // r = |k1| P + (+-)|k2| phi(P), the second half negated when the signs differ; the
    // sign of k1 is applied at the end
    const bool flip = n1 != n2;
    const fq29 beta = q_const(Glv::BETA29);
    uint32_t s1[5], s2[5];
    ec_booth_init(m1, s1);
    ec_booth_init(m2, s2);
    x29 r = x29_inf();
    for (int w = 0; w < 43; w++) {
        r = x29_dbl(x29_dbl(x29_dbl(r)));
        const int e1 = ec_booth(ec_booth_next(s1)), e2 = ec_booth(ec_booth_next(s2));
        a29 t1 = ec_tab_load(tab, e1 ? (e1 < 0 ? -e1 : e1) - 1 : 0);
        if (e1 < 0) t1.y = q_neg(t1.y);
        r = x29_select(e1 != 0, x29_madd(r, t1), r);
        a29 t2 = ec_tab_load(tab, e2 ? (e2 < 0 ? -e2 : e2) - 1 : 0);
        t2.x = q_mul(t2.x, beta);
        if ((e2 < 0) != flip) t2.y = q_neg(t2.y);
        r = x29_select(e2 != 0, x29_madd(r, t2), r);
    }
    if (n1) r.Y = q_sub(q_zero(), r.Y);
    return x29_select(x29_is_inf(p), x29_inf(), r);
}

// one radix-2 DIT stage: (u, v) -> (u + w v, u - w v), w = tw[j << tw_shift]
__global__ __launch_bounds__(EC_THREADS) void k_ec_dit_stage(uint32_t* pts, const uint32_t* tw, int s, int tw_shift,
                                                             uint32_t half_n) {
    __shared__ uint32_t tab[EC_LDS_WORDS];
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= half_n) return;
    uint32_t half = 1u << (s - 1);
    uint32_t j = t & (half - 1);
    uint32_t i0 = ((t >> (s - 1)) << s) | j, i1 = i0 + half;
    x29 v = x29_load(pts + X29_WORDS * (size_t)i1);
    // stage 1's twiddle is w^0 = 1, and a wave whose v are all infinity skips the
    // multiplication (from_pp's quotient transforms take inputs that are infinity past the
    // first N / 2^d + 1, so in their stages below d every v is infinity).  The test is per
    // wave: a wave with some finite v multiplies on every lane (x29_mul_scalar returns
    // infinity for infinity), since running the multiplication on part of the lanes
    // measured twice as slow as on all of them.
    if (s > 1 && __any(!x29_is_inf(v)))
        v = x29_mul_scalar(v, f_from_mont(f_load<FrP>(tw + 8 * (size_t)(j << tw_shift))), tab);
    const x29 u = x29_load(pts + X29_WORDS * (size_t)i0);
    x29 nv = v;
    nv.Y = q_sub(q_zero(), v.Y);
    x29_store(pts + X29_WORDS * (size_t)i0, x29_add(u, v));
    x29_store(pts + X29_WORDS * (size_t)i1, x29_add(u, nv));
}

// size_inv scaling of the inverse transform (N_ctx^-1 * 2^shift), then canonical affine out
__global__ __launch_bounds__(EC_THREADS) void k_ec_store(const uint32_t* pts, const uint32_t* consts, int scale_shift,
                                                         int scale, uint32_t* out, uint32_t n) {
    __shared__ uint32_t tab[EC_LDS_WORDS];
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    x29 v = x29_load(pts + X29_WORDS * (size_t)p);
    if (scale) {
        fr f = f_load<FrP>(consts + 8 * K_NINV);
        for (int i = 0; i < scale_shift; i++) f = f_dbl(f);
        v = x29_mul_scalar(v, f_from_mont(f), tab);
    }
    x29_store_canon(out + 16 * (size_t)p, v);
}

// ---------------------------------------------------------------------------
// Merkle (EncodedBlobMerkle::build, crates/encoder/src/merkle/blob.rs:15-34)
// ---------------------------------------------------------------------------
// One leaf per thread: 8 consecutive row-major elements of coset c, converted out of
// Montgomery form (scalar_to_h256, crates/encoder/src/utils.rs:17-20), written to the
// encoded-data output and hashed (keccak_chunked(.., 8), utils.rs:44-57).
__global__ void k_merkle_leaves(const uint32_t* y, int log_col, int log_row, int coset, uint32_t* data_out,
                                uint32_t* leaves) {
    const uint32_t gpr = 1u << (log_col - 3);  // leaf groups per row
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t rows = 1u << log_row;
    if (t >= rows * gpr) return;
    uint32_t r = t & (rows - 1), g = t >> log_row;  // consecutive threads walk down a column
    const size_t n = (size_t)1 << (log_col + log_row);
    uint64_t lanes[32];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint32_t col = 8 * g + k;
        fr v = f_from_mont(f_load<FrP>(y + 8 * (((size_t)col << log_row) | r)));
#pragma unroll
        for (int q = 0; q < 4; q++) lanes[4 * k + q] = (uint64_t)v.v[2 * q] | ((uint64_t)v.v[2 * q + 1] << 32);
    }
    size_t elem = (size_t)coset * n + ((size_t)r << log_col) + 8 * g;
    if (data_out) {
        uint4* dst = (uint4*)(data_out + 8 * elem);
#pragma unroll
        for (int q = 0; q < 16; q++)
            dst[q] = make_uint4((uint32_t)lanes[2 * q], (uint32_t)(lanes[2 * q] >> 32), (uint32_t)lanes[2 * q + 1],
                                (uint32_t)(lanes[2 * q + 1] >> 32));
    }
    uint64_t hsh[4];
    keccak256_256(lanes, hsh);
    uint64_t* lo = (uint64_t*)(leaves + 8 * (elem >> 3));
#pragma unroll
    for (int q = 0; q < 4; q++) lo[q] = hsh[q];
}

// leaves from 32-byte elements already in memory (primitive)
__global__ void k_merkle_leaves_flat(const uint32_t* data, uint32_t* leaves, uint32_t count) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint64_t lanes[32];
    const uint64_t* src = (const uint64_t*)(data + 64 * (size_t)t);
#pragma unroll
    for (int q = 0; q < 32; q++) lanes[q] = src[q];
    uint64_t hsh[4];
    keccak256_256(lanes, hsh);
    uint64_t* lo = (uint64_t*)(leaves + 8 * (size_t)t);
#pragma unroll
    for (int q = 0; q < 4; q++) lo[q] = hsh[q];
}

// parent[i] = keccak256(child[2i] || child[2i+1])  (keccak_chunked(.., 2))
__global__ void k_merkle_level(const uint32_t* child, uint32_t* parent, uint32_t count) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint64_t in[8];
    const uint64_t* src = (const uint64_t*)(child + 16 * (size_t)t);
#pragma unroll
    for (int q = 0; q < 8; q++) in[q] = src[q];
    uint64_t hsh[4];
    keccak256_64(in, hsh);
    uint64_t* dst = (uint64_t*)(parent + 8 * (size_t)t);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = hsh[q];
}

// compute_file_root (crates/encoder/src/encoder/blob.rs:64-72) over the top layer
__global__ void k_file_root(const uint32_t* roots, int count, uint32_t* out) {
    if (threadIdx.x || blockIdx.x) return;
    const uint64_t* r = (const uint64_t*)roots;
    uint64_t res[4];
    if (count == 1) {
        for (int q = 0; q < 4; q++) res[q] = r[q];
    } else {
        uint64_t in[8];
        for (int q = 0; q < 8; q++) in[q] = r[q];
        keccak256_64(in, res);
        if (count == 3) {
            for (int q = 0; q < 4; q++) {
                in[q] = res[q];
                in[4 + q] = r[8 + q];
            }
            keccak256_64(in, res);
        }
    }
    uint64_t* o = (uint64_t*)out;
    for (int q = 0; q < 4; q++) o[q] = res[q];
}

// ---------------------------------------------------------------------------
// params from tau: closed form of AMTParams::from_pp (generate.rs:139-232)
// ---------------------------------------------------------------------------
// fixed-base table: table[w * 256 + j] = [j * 2^(8w)] G1 (affine mont), j > 0
__global__ __launch_bounds__(64) void k_fixed_base_table(uint32_t* table) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 32 * 256) return;
    uint32_t w = t >> 8, j = t & 255;
    g1_aff G;
    G.x = f_one<FqP>();  // generator (1, 2)
    G.y = f_dbl(G.x);
    g1_xyzz acc = xyzz_inf();
    for (int b = 7; b >= 0; b--) {
        acc = xyzz_dbl(acc);
        if ((j >> b) & 1) acc = xyzz_madd(acc, G);
    }
    for (uint32_t k = 0; k < 8 * w; k++) acc = xyzz_dbl(acc);
    bool inf;
    g1_aff a = xyzz_to_aff(acc, &inf);
    if (inf) {
        a.x = f_zero<FqP>();
        a.y = f_zero<FqP>();
    }
    aff_store(table + 16 * (size_t)t, a);
}

// per-coset scalars: pc[0] = tau' = tau / cf, pc[1] = tau'^N - 1, pc[2] = tau^(3N),
// pc[3 + d - 1] = tau'^(N / 2^d) for d = 1..h
__global__ void k_params_consts(const uint32_t* tau_canon, const uint32_t* consts, int coset, int log_n, int h,
                                uint32_t* pc) {
    if (threadIdx.x || blockIdx.x) return;
    fr tau = f_to_mont(f_load<FrP>(tau_canon));
    fr cf = f_load<FrP>(consts + 8 * (K_CF + coset));
    fr tp = f_mul(tau, f_inv(cf));
    f_store(pc, tp);
    fr tpN = tp;
    for (int i = 0; i < log_n; i++) tpN = f_sqr(tpN);
    f_store(pc + 8, f_sub(tpN, f_one<FrP>()));
    // tau^(3N) (power_tau.rs:66-71: high_start = 2^(depth+2) - 2^depth)
    fr tN = tau;
    for (int i = 0; i < log_n; i++) tN = f_sqr(tN);
    f_store(pc + 16, f_mul(f_sqr(tN), tN));
    for (int d = 1; d <= h; d++) {
        fr x = tp;
        for (int i = 0; i < log_n - d; i++) x = f_sqr(x);
        f_store(pc + 8 * (3 + d - 1), x);
    }
}

// One thread per domain index t: the 2^h + ... = h + 2 line scalars
//   A_t = w^t / (N (tau' - w^t)),  basis = A_t (tau'^N - 1),  high = tau^(3N) basis,
//   quot_d = A_t (tau'^m - w^(t m)), m = N / 2^d,
// each turned into [s] G1 and stored at the bit-reversed position (enact, generate.rs:122-126).
__global__ __launch_bounds__(64) void k_params_gen(uint32_t* bases, const uint32_t* consts, const uint32_t* pc, const uint32_t* table,
                             int log_n, int h, int* bad) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_n;
    if (t >= n) return;
    fr w = f_load<FrP>(consts + 8 * K_OMEGA);
    fr wt = f_one<FrP>();
    for (int i = log_n - 1; i >= 0; i--) {
        wt = f_sqr(wt);
        if ((t >> i) & 1) wt = f_mul(wt, w);
    }
    fr tp = f_load<FrP>(pc);
    fr den = f_sub(tp, wt);
    if (f_is_zero(den)) *bad = 1;
    fr A = f_mul(f_mul(f_inv(den), wt), f_load<FrP>(consts + 8 * K_NINV));
    const uint32_t pos = bitrev_bits(t, log_n);
    fr sb = f_mul(A, f_load<FrP>(pc + 8));
    for (int line = 0; line < h + 2; line++) {
        fr sc;
        if (line == h) {
            sc = sb;
        } else if (line == h + 1) {
            sc = f_mul(sb, f_load<FrP>(pc + 16));
        } else {
            int d = line + 1;
            fr wtm = wt;  // w^(t m) = (w^t)^(2^(log_n - d))
            for (int i = 0; i < log_n - d; i++) wtm = f_sqr(wtm);
            sc = f_mul(A, f_sub(f_load<FrP>(pc + 8 * (3 + d - 1)), wtm));
        }
        g1_xyzz P = fixed_base_mul(table, f_from_mont(sc));
        bool inf;
        g1_aff a = xyzz_to_aff(P, &inf);
        if (inf) *bad = 1;
        aff_store(bases + (((size_t)line << log_n) + pos) * 16, a);
    }
}

// ---- one MSM of arbitrary bases (zgr_msm) through the wide sort and bucket kernels ----
// Each scalar is split by GLV (glv_split: k = +-k1 +- lambda k2, |k1|, |k2| < 2^127), so
// the MSM is one of 2n points -- P_i, then phi(P_i) -- with 127-bit scalars: 8 signed
// 16-bit windows.  Each window of each point is a virtual scalar of one digit: sort block
// blk holds window blk >> sh over points [(blk mod 2^sh) ns, + ns), so a window's entries
// form one group of the wide sort (bucket runs by (|digit|, window)) and an entry's point
// index is its table row: no shifted copies.  The bucket kernels give one sum S_w per
// window, and k_pm_horner returns sum_w 2^(16 w) S_w: one chain of 112 doublings, the
// latency that halving the windows by GLV halves.
constexpr int PM_W = 8;           // 16-bit windows of a 127-bit half
constexpr int PM_SPLIT_WORDS = 9;  // per scalar: |k1| (4 words), |k2| (4 words), signs

// |k1|, |k2| and their signs of every scalar (Montgomery in), one thread per scalar
__global__ __launch_bounds__(256) void k_pm_split(const uint32_t* y, uint32_t n, uint32_t* halves) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t m1[4], m2[4];
    bool n1, n2;
    glv_split(f_from_mont(f_load<FrP>(y + 8 * (size_t)i)), m1, &n1, m2, &n2);
    uint32_t* o = halves + (size_t)PM_SPLIT_WORDS * i;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[k] = m1[k];
        o[4 + k] = m2[k];
    }
    o[8] = (uint32_t)n1 | ((uint32_t)n2 << 1);
}

// signed digit of window w of virtual point j (j < n: |k1| of scalar j, else |k2| of
// scalar j - n) in the Booth form -2^15 b_{16w+15} + sum_{k<15} 2^k b_{16w+k} + b_{16w-1},
// |digit| <= 2^15, with the half's sign folded in; no carry chain across the windows
__device__ __forceinline__ int pm_digit(const uint32_t* halves, uint32_t n, uint32_t j, uint32_t w) {
    const uint32_t half = j >= n ? 1u : 0u, i = j - half * n;
    const uint32_t* o = halves + (size_t)PM_SPLIT_WORDS * i + 4 * half;
    const uint32_t word = o[w >> 1], v = (w & 1) ? word >> 16 : word & 0xffffu;
    uint32_t low = 0;  // bit 16 w - 1
    if (w) low = (w & 1) ? (word >> 15) & 1u : o[(w >> 1) - 1] >> 31;
    int d = (int)(v & 0x7fffu) - (int)(v & 0x8000u) + (int)low;
    const bool neg = (o[8 - 4 * half] >> half) & 1u;
    return neg ? -d : d;
}

__global__ __launch_bounds__(FW_SORT_T) void k_fwp_count(const uint32_t* halves, uint32_t n, int sh, uint32_t ns,
                                                         uint32_t* cnt) {
    extern __shared__ uint32_t hist[];  // FW_BUCKETS
    const uint32_t blk = blockIdx.x, t = threadIdx.x;
    for (uint32_t q = t; q < FW_BUCKETS; q += FW_SORT_T) hist[q] = 0;
    __syncthreads();
    const uint32_t win = blk >> sh, first = (blk & ((1u << sh) - 1)) * ns;
    for (uint32_t u = t; u < ns; u += FW_SORT_T) {
        const int d = pm_digit(halves, n, first + u, win);
        if (d) atomicAdd(&hist[(d < 0 ? -d : d) - 1], 1u);
    }
    __syncthreads();
    for (uint32_t q = t; q < FW_BUCKETS; q += FW_SORT_T) cnt[(size_t)blk * FW_BUCKETS + q] = hist[q];
}

__global__ __launch_bounds__(FW_SORT_T) void k_fwp_scatter(const uint32_t* halves, uint32_t n, int sh, uint32_t ns,
                                                           const uint32_t* off, uint2* ent) {
    extern __shared__ uint32_t cursor[];  // FW_BUCKETS
    const uint32_t blk = blockIdx.x, t = threadIdx.x;
    for (uint32_t q = t; q < FW_BUCKETS; q += FW_SORT_T) cursor[q] = off[(size_t)blk * FW_BUCKETS + q];
    __syncthreads();
    const uint32_t win = blk >> sh, first = (blk & ((1u << sh) - 1)) * ns;
    for (uint32_t u = t; u < ns; u += FW_SORT_T) {
        const int d = pm_digit(halves, n, first + u, win);
        if (!d) continue;
        const uint32_t mag = d < 0 ? -d : d;
        const uint32_t pos = atomicAdd(&cursor[mag - 1], 1u);
        ent[pos] = make_uint2((first + u) | ((uint32_t)(d < 0) << 31), (mag << 16) | blk);
    }
}

// A base of all zeros is the point at infinity (the oracle's and the reference's reading
// of (0, 0)); the bucket kernels add affine points only, so its scalar is cleared first
// and the base is never added.  One thread per point, scalars Montgomery (8 words).
__global__ __launch_bounds__(256) void k_msm_clear_inf(const uint32_t* bases, uint32_t n, uint32_t* y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const g1_aff a = aff_load(bases + 16 * (size_t)i);
    if (f_is_zero(a.x) && f_is_zero(a.y)) {
#pragma unroll
        for (int k = 0; k < 8; k++) y[8 * (size_t)i + k] = 0;
    }
}

// bases (x_mont || y_mont, 2^256 form) -> the table rows of the bucket kernels (x || y as
// 256-bit integers of the 2^261 form): row i = P_i, row n + i = phi(P_i) = (beta x, y)
__global__ __launch_bounds__(256) void k_pm_table(const uint32_t* bases, uint32_t n, uint32_t* table) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const g1_aff a = aff_load(bases + 16 * (size_t)i);
    const fq29 x = q_from_fq(a.x), y = q_canon(q_from_fq(a.y));
    fb_store29(table + 16 * (size_t)i, q_canon(x));
    fb_store29(table + 16 * (size_t)i + 8, y);
    fb_store29(table + 16 * ((size_t)n + i), q_canon(q_mul(x, q_const(Glv::BETA29))));
    fb_store29(table + 16 * ((size_t)n + i) + 8, y);
}

// The Horner below is one dependent chain: 112 doublings and 7 additions on one point,
// whose cost is the instructions one wave issues, not the lanes it uses.  So the
// independent multiplications of each formula level run side by side on lanes 0..3 of
// the wave (operands picked per lane, the products broadcast back with readlane): an
// XYZZ doubling is 3 multiplication steps instead of 9, an addition 4 instead of 14.

// lane 0..3 picks a0..a3 (lanes above 3 repeat lane 3's pick)
__device__ __forceinline__ fq29 q_pick4(int lane, const fq29& a0, const fq29& a1, const fq29& a2,
                                        const fq29& a3) {
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = lane == 0 ? a0.l[k] : lane == 1 ? a1.l[k] : lane == 2 ? a2.l[k] : a3.l[k];
    return r;
}

// lane src's value, uniform over the wave
__device__ __forceinline__ fq29 q_lane(const fq29& a, int src) {
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = (uint32_t)__builtin_amdgcn_readlane((int)a.l[k], src);
    return r;
}

// dbl-2008-s-1 (as x29_dbl) in three steps: V = U^2, X^2 | W = U V, S = X V, M^2, V ZZ |
// M (S - X3), W Y, W ZZZ, with U = 2Y, M = 3X^2, X3 = M^2 - 2S, Y3 = M (S - X3) - W Y
__device__ __forceinline__ x29 x29_dbl_lanes(const x29& p, int lane) {
    const fq29 U = q_dbl(p.Y);
    fq29 t = q_mul(q_pick4(lane, U, p.X, U, p.X), q_pick4(lane, U, p.X, U, p.X));
    const fq29 V = q_lane(t, 0), X2 = q_lane(t, 1);
    const fq29 M = q_add(q_dbl(X2), X2);
    t = q_mul(q_pick4(lane, U, p.X, M, V), q_pick4(lane, V, V, M, p.ZZ));
    const fq29 W = q_lane(t, 0), S = q_lane(t, 1);
    x29 r;
    r.X = q_sub(q_lane(t, 2), q_dbl(S));
    r.ZZ = q_lane(t, 3);
    const fq29 SX = q_sub(S, r.X);
    t = q_mul(q_pick4(lane, M, W, W, W), q_pick4(lane, SX, p.Y, p.ZZZ, p.ZZZ));
    r.Y = q_sub(q_lane(t, 0), q_lane(t, 1));
    r.ZZZ = q_lane(t, 2);
    return r;
}

// add-2008-s (as x29_add, every exceptional case handled) in four steps:
// U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1 | PP = P^2, ZZ1 ZZ2, ZZZ1 ZZZ2, R^2 |
// PPP = P PP, Q = U1 PP, ZZ3 = ZZ1 ZZ2 PP | ZZZ3 = ZZZ1 ZZZ2 PPP, S1 PPP, R (Q - X3)
// with P = U2 - U1, R = S2 - S1, X3 = R^2 - PPP - 2Q, Y3 = R (Q - X3) - S1 PPP
__device__ __forceinline__ x29 x29_add_lanes(const x29& p, const x29& q, int lane) {
    if (x29_is_inf(p)) return q;
    if (x29_is_inf(q)) return p;
    fq29 t = q_mul(q_pick4(lane, p.X, q.X, p.Y, q.Y), q_pick4(lane, q.ZZ, p.ZZ, q.ZZZ, p.ZZZ));
    const fq29 U1 = q_lane(t, 0), S1 = q_lane(t, 2);
    const fq29 P = q_sub(q_lane(t, 1), U1), R = q_sub(q_lane(t, 3), S1);
    if (q_is_zero(P)) return q_is_zero(R) ? x29_dbl_lanes(p, lane) : x29_inf();
    t = q_mul(q_pick4(lane, P, p.ZZ, p.ZZZ, R), q_pick4(lane, P, q.ZZ, q.ZZZ, R));
    const fq29 PP = q_lane(t, 0), ZZ12 = q_lane(t, 1), ZZZ12 = q_lane(t, 2), R2 = q_lane(t, 3);
    t = q_mul(q_pick4(lane, P, U1, ZZ12, ZZ12), PP);
    const fq29 PPP = q_lane(t, 0), Q = q_lane(t, 1);
    x29 r;
    r.ZZ = q_lane(t, 2);
    r.X = q_sub(q_sub(R2, PPP), q_dbl(Q));
    const fq29 QX = q_sub(Q, r.X);
    t = q_mul(q_pick4(lane, ZZZ12, S1, R, R), q_pick4(lane, PPP, PPP, QX, QX));
    r.ZZZ = q_lane(t, 0);
    r.Y = q_sub(q_lane(t, 2), q_lane(t, 1));
    return r;
}

// sum_w 2^(16 w) S_w from the window sums at heap nodes 1 .. PM_W, canonical affine
__global__ __launch_bounds__(64) void k_pm_horner(const uint32_t* heap, uint32_t* out) {
    const int lane = (int)threadIdx.x;
    x29 r = x29_load(heap + PM_W * X29_WORDS);
    for (int w = PM_W - 2; w >= 0; w--) {
        for (int k = 0; k < FW_C; k++) r = x29_dbl_lanes(r, lane);
        r = x29_add_lanes(r, x29_load(heap + (1 + w) * X29_WORDS), lane);
    }
    if (lane == 0) x29_store_canon(out, r);
}

}  // namespace zgr
