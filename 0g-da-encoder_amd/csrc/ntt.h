// Fr NTT passes (the radix-2 FFTs of ark-poly's Radix2EvaluationDomain that
// to_coset_blob runs, crates/amt/src/blob/encode.rs:142-157).  Included once, by zgr.hip.
//
// One kernel, k_ntt_pass, runs a sub-transform of up to 2^NTT_MAX_LOG points per
// workgroup entirely in LDS.  Transforms of up to 2^11 points are one pass; larger ones
// (up to 2^22) are the four-step decomposition N = N1 N2 (N1 = 2^n1 over the high index
// bits, N2 = 2^n2 over the low ones), two passes over HBM instead of log2 N:
//
//   natural -> bit-reversed (DIF order, the iNTT of the encoder):
//     pass A  column j < N2 (elements j + N2 t): N1-point DFT over t, written in
//             bit-reversed order, each output a multiplied by w_N^(j a);
//     pass B  block p < N1 (elements N2 p + i): N2-point DFT over i, bit-reversed.
//     The result sits at bitrev_n(k) = bitrev_n2(k >> n1) + N2 bitrev_n1(k mod N1).
//   bit-reversed -> natural (DIT order, the coset evaluations):
//     pass 1  block p: N2-point DFT of its bit-reversed elements, natural order, each
//             output j multiplied by w_N^(bitrev_n1(p) j);
//     pass 2  column j: N1-point DFT over the bit-reversed blocks, natural order.
//
// Inside a workgroup the data are always transformed by radix-2 DIT stages (bit-reversed
// LDS order -> natural LDS order); a natural input is stored to LDS bit-reversed, and a
// bit-reversed output is read from LDS bit-reversed.
//
// Arithmetic: data keep arkworks' 2^256 Montgomery form; in LDS they are 9 limbs of 29
// bits (fr29.h), and every constant multiplied in (twiddles, coset factors, 1/N) is in
// the 2^261 form, so r29_mul keeps the data's form.  Sums are reduced lazily: a stage
// adds at most 2r to an element's bound (u + w v and u - w v + 2r with w v < 2r), so
// after 11 stages the values stay below 24 r < 2^259, inside r29_mul's input range
// when multiplied by a canonical constant; r29_reduce makes each output canonical.
#pragma once
#include "fr29.h"

namespace zgr {

constexpr int NTT_MAX_LOG = 11;                 // points of one in-LDS sub-transform
constexpr int NTT_STAGE_TW = 1 << NTT_MAX_LOG;  // entries of the stage-twiddle table
constexpr int NTT_MAX_TWO_PASS = 2 * NTT_MAX_LOG;

// dynamic LDS of k_ntt_pass for a 2^log_m sub-transform: the data, 9 limb planes.  The
// stage twiddles are read from the global table (2 x 73 KB per context, cache-resident):
// holding them in LDS as well doubled the footprint and halved the workgroups per CU.
__host__ __device__ inline size_t ntt_lds_bytes(int log_m) { return (size_t)9 * 4 << log_m; }

// one pass of a batched transform (see the file comment)
struct NttPass {
    const uint32_t* src;    // Fr, 2^256 Montgomery, 8 u32 per element
    uint32_t* dst;          // may equal src
    const uint32_t* pw_lo;  // post twiddles w_ctx^x = pw_hi[x >> pw_s] pw_lo[x mod 2^pw_s] (k_ntt_tables),
    const uint32_t* pw_hi;  //   2^261 form, the pass's direction
    const uint32_t* stw;    // stage twiddles (k_ntt_tables), 9 x NTT_STAGE_TW, same direction
    const uint32_t* pre;    // null, or a factor per input element (2^261 form, row-local index)
    const uint32_t* scale;  // null, or one factor (2^261 form) for every output
    uint64_t row_len;       // elements from one row (blockIdx.y) to the next
    uint32_t gstride;       // elements between consecutive sub-transforms of a row
    uint32_t estride;       // elements between consecutive points of a sub-transform
    int log_m;              // sub-transform size
    int log_g;              // log2 of the sub-transforms per row (gridDim.x)
    int log_ctx;            // the context domain of pw_lo / pw_hi
    int pw_s;               // bits of the exponent that pw_lo covers
    int in_rev;             // input point e goes to LDS position bitrev(e)
    int out_rev;            // output point e is LDS position bitrev(e)
    int post;               // 0: none; 1: x w^(g a); 2: x w^(bitrev(g) a); w = w_(2^log_post), a = LDS index
    int log_post;
};

__device__ __forceinline__ fq29 ntt_lds_load(const uint32_t* base, uint32_t stride, uint32_t i) {
    fq29 v;
#pragma unroll
    for (int l = 0; l < 9; l++) v.l[l] = base[l * stride + i];
    return v;
}

__device__ __forceinline__ void ntt_lds_store(uint32_t* base, uint32_t stride, uint32_t i, const fq29& v) {
#pragma unroll
    for (int l = 0; l < 9; l++) base[l * stride + i] = v.l[l];
}

// w_ctx^x (x < N_ctx) as the product of two entries of small tables that stay in cache:
// a direct gather from an N/2-entry table brings a 128 B line per 32 B twiddle when a
// column's exponents step by its index.  Result below 2r.
__device__ __forceinline__ fq29 ntt_wpow(const NttPass& a, uint64_t x) {
    const fq29 hi = fr29_unpack(a.pw_hi + 8 * (size_t)(x >> a.pw_s));
    const fq29 lo = fr29_unpack(a.pw_lo + 8 * (size_t)(x & ((1ull << a.pw_s) - 1)));
    return r29_mul(hi, lo);
}

__global__ __launch_bounds__(1024) void k_ntt_pass(NttPass a) {
    extern __shared__ uint32_t lds[];
    const int m = a.log_m;
    const uint32_t M = 1u << m, H = M >> 1, T = blockDim.x, tid = threadIdx.x;
    uint32_t* dat = lds;  // 9 x M
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);  // columns that share cache lines stay on one XCD
    const size_t row0 = (size_t)blockIdx.y * a.row_len;
    const size_t base = (size_t)g * a.gstride;  // row-local index of point 0

    for (uint32_t e = tid; e < M; e += T) {
        const size_t idx = base + (size_t)e * a.estride;
        fq29 v = fr29_unpack(a.src + 8 * (row0 + idx));
        if (a.pre) v = r29_mul(v, fr29_unpack(a.pre + 8 * idx));
        ntt_lds_store(dat, M, a.in_rev ? bitrev_bits(e, m) : e, v);
    }
    __syncthreads();
    for (int s = 1; s <= m; s++) {
        const uint32_t half = 1u << (s - 1);
        for (uint32_t t = tid; t < H; t += T) {
            const uint32_t j = t & (half - 1);
            const uint32_t i0 = ((t >> (s - 1)) << s) | j, i1 = i0 + half;
            const fq29 u = ntt_lds_load(dat, M, i0);
            fq29 v = ntt_lds_load(dat, M, i1);
            if (s > 1) v = r29_mul(v, ntt_lds_load(a.stw, NTT_STAGE_TW, half - 1 + j));  // w_(2^s)^j; v < 2r
            ntt_lds_store(dat, M, i0, r29_add(u, v));
            ntt_lds_store(dat, M, i1, r29_sub_lazy(u, v));
        }
        __syncthreads();
    }
    fq29 sc;
    if (a.scale) sc = fr29_unpack(a.scale);
    const uint32_t G = a.post == 2 ? bitrev_bits(g, a.log_g) : g;
    for (uint32_t e = tid; e < M; e += T) {
        const uint32_t p = a.out_rev ? bitrev_bits(e, m) : e;
        fq29 v = ntt_lds_load(dat, M, p);
        if (a.post) {
            const uint64_t x = ((uint64_t)G * p) & ((1ull << a.log_post) - 1);
            v = r29_mul(v, ntt_wpow(a, x << (a.log_ctx - a.log_post)));  // v < 24r, w < 2r
        }
        if (a.scale) v = r29_mul(v, sc);
        uint32_t w[8];
        fr29_pack(r29_reduce(v), w);
        uint4* o = (uint4*)(a.dst + 8 * (row0 + base + (size_t)e * a.estride));
        o[0] = make_uint4(w[0], w[1], w[2], w[3]);
        o[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// out[row][p] = in[row][bitrev_n(p)] (x scale when given): a bit-reversed transform of
// more than 2^11 points back to natural order
__global__ void k_ntt_unrev(const uint32_t* in, uint32_t* out, int n, const uint32_t* scale) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >> n) return;
    const size_t row0 = (size_t)blockIdx.y << n;
    fq29 v = fr29_unpack(in + 8 * (row0 + bitrev_bits(p, n)));
    if (scale) v = r29_reduce(r29_mul(v, fr29_unpack(scale)));
    uint32_t w[8];
    fr29_pack(v, w);
    uint4* o = (uint4*)(out + 8 * (row0 + p));
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// a value in the 2^256 form -> the same field element in the 2^261 form, canonical
__device__ __forceinline__ fq29 ntt_to261(const uint32_t* p) {
    return r29_reduce(r29_mul(fr29_unpack(p), q_const(R29M::C_TO)));
}

// bits of a post-twiddle exponent that the low table covers
__host__ __device__ inline int ntt_pw_bits(int log_n) { return (log_n + 1) / 2; }

// w^x (x < N, 2^256 form) from the half table tw[k] = w^k, k < N/2 (w^(N/2) = -1), in the
// 2^261 form, canonical
__device__ __forceinline__ fq29 ntt_pow261(const uint32_t* tw, uint32_t x, int log_n) {
    const uint32_t half = 1u << (log_n - 1);
    const fq29 v = ntt_to261(tw + 8 * (size_t)(x & (half - 1)));
    return (x & half) ? r29_neg_canon(v) : v;
}

__device__ __forceinline__ void ntt_store_packed(uint32_t* dst, const fq29& v) {
    uint32_t w[8];
    fr29_pack(v, w);
#pragma unroll
    for (int i = 0; i < 8; i++) dst[i] = w[i];
}

// The 2^261-form tables of a context (from its 2^256-form twiddles tw / tw_inv), packed
// 8 u32 per entry unless noted, S = ntt_pw_bits(log_n):
//   pw_lo[k], pw_lo_inv[k]   w^(+-k), k < 2^S
//   pw_hi[k], pw_hi_inv[k]   w^(+-k 2^S), k < 2^(log_n - S)
//   stw, stw_inv             stage twiddles: entry 2^(s-1) - 1 + j = w_(2^s)^(+-j), j < 2^(s-1),
//                            s <= min(11, log_n), as 9 limb planes of NTT_STAGE_TW words
//   ninv29[k]                2^k / N for k <= log_n (1 / size of a 2^(log_n - k) transform)
// One thread per k < N/2.
__global__ void k_ntt_tables(const uint32_t* tw, const uint32_t* tw_inv, const uint32_t* consts, int log_n,
                             uint32_t* pw_lo, uint32_t* pw_lo_inv, uint32_t* pw_hi, uint32_t* pw_hi_inv,
                             uint32_t* stw, uint32_t* stw_inv, uint32_t* ninv29) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t half = 1u << (log_n - 1);
    if (k >= half) return;
    uint32_t w[8];
    const int S = ntt_pw_bits(log_n);
    if (k < (1u << S)) {
        ntt_store_packed(pw_lo + 8 * (size_t)k, ntt_pow261(tw, k, log_n));
        ntt_store_packed(pw_lo_inv + 8 * (size_t)k, ntt_pow261(tw_inv, k, log_n));
    }
    if (k < (1u << (log_n - S))) {
        ntt_store_packed(pw_hi + 8 * (size_t)k, ntt_pow261(tw, k << S, log_n));
        ntt_store_packed(pw_hi_inv + 8 * (size_t)k, ntt_pow261(tw_inv, k << S, log_n));
    }
    // 2^smax - 1 stage entries (up to N - 1) over N / 2 threads: entries k and k + N / 2
    const int smax = log_n < NTT_MAX_LOG ? log_n : NTT_MAX_LOG;
    for (uint32_t e = k; e + 1 < (1u << smax); e += half) {
        const int s = 32 - __builtin_clz(e + 1);  // stage: 2^(s-1) <= e + 1 < 2^s
        const uint32_t j = e + 1 - (1u << (s - 1));
        const size_t at = (size_t)j << (log_n - s);  // w_(2^s)^j = w^(j N / 2^s)
        const fq29 f = ntt_to261(tw + 8 * at), b = ntt_to261(tw_inv + 8 * at);
#pragma unroll
        for (int l = 0; l < 9; l++) {
            stw[l * NTT_STAGE_TW + e] = f.l[l];
            stw_inv[l * NTT_STAGE_TW + e] = b.l[l];
        }
    }
    if ((int)k <= log_n) {
        fr f = f_load<FrP>(consts + 8 * K_NINV);
        for (uint32_t i = 0; i < k; i++) f = f_dbl(f);
        fr29_pack(ntt_to261(f.v), w);
#pragma unroll
        for (int i = 0; i < 8; i++) ninv29[8 * (size_t)k + i] = w[i];
    }
}

// a table of Fr in the 2^256 form -> the 2^261 form, in place
__global__ void k_to261_inplace(uint32_t* t, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t w[8];
    fr29_pack(ntt_to261(t + 8 * (size_t)k), w);
#pragma unroll
    for (int i = 0; i < 8; i++) t[8 * (size_t)k + i] = w[i];
}

}  // namespace zgr
