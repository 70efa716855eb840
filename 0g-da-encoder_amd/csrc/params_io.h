// Device kernels of the public-parameter pipeline beyond the prove bases: the powers of
// tau in G1 and G2 (PowerTau::setup_inner, crates/amt/src/power_tau.rs:54-80), the G2
// vanishes of each coset (gen_vanishes, crates/amt/src/prove_params/generate.rs:234-263),
// and the conversion of points to arkworks' serialized forms for the parameter files
// (ark-serialize 0.4 CanonicalSerialize of G1Affine / G2Affine).  Compiled once, in
// zgr_params.hip; zgr.hip launches these kernels through params_kernels.h.
#pragma once
#include "common.h"
#include "g2.h"
#include "params_kernels.h"
#include "ppot.h"

namespace zgr {

// ---------------------------------------------------------------------------
// fixed-base multiplication in G2
// ---------------------------------------------------------------------------

// table[w * 256 + j] = [j 2^(8 w)] G2 (affine Montgomery), j > 0
__global__ __launch_bounds__(64) void k_g2_fixed_base_table(uint32_t* table) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 32 * 256) return;
    const uint32_t w = t >> 8, j = t & 255;
    const g2_aff G = g2_generator();
    g2_xyzz acc = g2_inf();
    for (int b = 7; b >= 0; b--) {
        acc = g2_dbl(acc);
        if ((j >> b) & 1) acc = g2_madd(acc, G);
    }
    for (uint32_t k = 0; k < 8 * w; k++) acc = g2_dbl(acc);
    bool inf;
    g2_aff a = g2_to_aff(acc, &inf);
    if (inf) a = {q2_zero(), q2_zero()};
    g2_store(table + G2_WORDS * (size_t)t, a);
}

// [s] G2 for a canonical scalar s: one mixed addition per non-zero byte
__device__ __forceinline__ g2_xyzz g2_fixed_base_mul(const uint32_t* table, const fr& s) {
    g2_xyzz acc = g2_inf();
    for (int w = 0; w < 32; w++) {
        const uint32_t d = (s.v[w >> 2] >> (8 * (w & 3))) & 0xffu;
        if (d) acc = g2_madd(acc, g2_load(table + G2_WORDS * (size_t)(w * 256 + d)));
    }
    return acc;
}

__device__ __forceinline__ void g1_store_canon(uint32_t* o, const g1_aff& a, bool inf) {
    const fq x = f_from_mont(a.x), y = f_from_mont(a.y);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        o[k] = inf ? 0u : x.v[k];
        o[8 + k] = inf ? 0u : y.v[k];
    }
}

__device__ __forceinline__ void g2_store_canon(uint32_t* o, const g2_aff& a, bool inf) {
    const fq v[4] = {f_from_mont(a.x.c0), f_from_mont(a.x.c1), f_from_mont(a.y.c0), f_from_mont(a.y.c1)};
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int k = 0; k < 8; k++) o[8 * c + k] = inf ? 0u : v[c].v[k];
}

// canonical affine G2 (all zeros = infinity) -> affine Montgomery
__device__ __forceinline__ g2_aff g2_load_canon(const uint32_t* p, bool* inf) {
    const g2_aff c = g2_load(p);
    *inf = q2_is_zero(c.x) && q2_is_zero(c.y);
    return {{f_to_mont(c.x.c0), f_to_mont(c.x.c1)}, {f_to_mont(c.y.c0), f_to_mont(c.y.c1)}};
}

// ---------------------------------------------------------------------------
// powers of tau
// ---------------------------------------------------------------------------
// One thread per i < n: [tau^(start + i)] G1 into out1 and [tau^(start + i)] G2 into out2
// (canonical affine, 16 / 32 u32; either may be null).  Infinity (tau = 0) sets *bad.
__global__ __launch_bounds__(64) void k_power_tau(const uint32_t* tau_canon, uint64_t start, uint32_t n,
                                                  const uint32_t* table1, const uint32_t* table2, uint32_t* out1,
                                                  uint32_t* out2, int* bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fr s = f_from_mont(f_pow_u64(f_to_mont(f_load<FrP>(tau_canon)), start + i));
    bool inf;
    if (out1) {
        const g1_aff a = xyzz_to_aff(fixed_base_mul(table1, s), &inf);
        if (inf) *bad = 1;
        g1_store_canon(out1 + 16 * (size_t)i, a, inf);
    }
    if (out2) {
        const g2_aff a = g2_to_aff(g2_fixed_base_mul(table2, s), &inf);
        if (inf) *bad = 1;
        g2_store_canon(out2 + G2_WORDS * (size_t)i, a, inf);
    }
}

// ---------------------------------------------------------------------------
// vanishes (gen_vanishes + enact): level d = 1..h holds 2^d points at point offset
// 2^d - 2; stored index k is the DFT output t = bitrev_d(k).  Affine Montgomery.
// ---------------------------------------------------------------------------
// Closed form with tau known (tau' = tau / cf_c, m = N / 2^d, D = 2^d):
//   vanish_t = sum_i coeff_i w_D^(-i t) = [(tau'^N - 1) / (tau'^m - w_D^t)] G2
// for coeff_0 = [tau'^(N - m)] G2, coeff_i = [tau'^((i - 1) m)] G2.  pc: k_params_consts'
// per-coset scalars; one thread per stored point.
__global__ __launch_bounds__(64) void k_vanish_closed(const uint32_t* pc, const uint32_t* consts, int log_n, int h,
                                                      const uint32_t* table2, uint32_t* out, int* bad) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx + 2 >= (2u << h)) return;
    const int d = 31 - __builtin_clz(idx + 2);
    const uint32_t t = bitrev_bits(idx + 2 - (1u << d), d);
    const fr wt = f_pow_u64(f_load<FrP>(consts + 8 * K_OMEGA), (uint64_t)t << (log_n - d));  // w_D^t
    const fr den = f_sub(f_load<FrP>(pc + 8 * (3 + d - 1)), wt);
    if (f_is_zero(den)) *bad = 1;
    const fr v = f_mul(f_load<FrP>(pc + 8), f_inv(den));
    bool inf;
    g2_aff a = g2_to_aff(g2_fixed_base_mul(table2, f_from_mont(v)), &inf);
    if (inf) *bad = 1;
    g2_store(out + G2_WORDS * (size_t)idx, a);
}


__device__ __forceinline__ void g2x_store(uint32_t* p, const g2_xyzz& v) {
    const fq f[8] = {v.X.c0, v.X.c1, v.Y.c0, v.Y.c1, v.ZZ.c0, v.ZZ.c1, v.ZZZ.c0, v.ZZZ.c1};
#pragma unroll
    for (int c = 0; c < 8; c++) f_store(p + 8 * c, f[c]);
}

__device__ __forceinline__ g2_xyzz g2x_load(const uint32_t* p) {
    g2_xyzz v;
    v.X = {f_load<FqP>(p), f_load<FqP>(p + 8)};
    v.Y = {f_load<FqP>(p + 16), f_load<FqP>(p + 24)};
    v.ZZ = {f_load<FqP>(p + 32), f_load<FqP>(p + 40)};
    v.ZZZ = {f_load<FqP>(p + 48), f_load<FqP>(p + 56)};
    return v;
}

// From the powers, as the reference computes it (no tau): level d's coefficients
// coeff_0 = g2pp'[N - m], coeff_i = g2pp'[(i - 1) m] (i = 1..D-1), with from_pp's coset
// scaling g2pp'[k] = w^k g2pp[k] (w = 1 / cf_c, generate.rs:152-168; scale = 0 for coset
// 0), loaded bit-reversed (XYZZ) for the DIT stages.  g2pp canonical affine.
// The h levels are independent transforms of 2^d points (d = 1..h), so every launch
// below covers all of them at once: level d's points sit at work[2^d - 2, 2^(d+1) - 2),
// the layout of the stored vanishes.  One launch per stage then costs one G2 scalar
// multiplication's latency for all levels together (the levels' transforms are too small
// to fill the device, so a launch per level and stage cost that latency each time).

// point t of all levels -> (level d, index i in it): t + 2 in [2^d, 2^(d+1))
__device__ __forceinline__ uint32_t van_level(uint32_t t, uint32_t* i) {
    const uint32_t d = 31u - (uint32_t)__builtin_clz(t + 2u);
    *i = t - ((1u << d) - 2u);
    return d;
}

__global__ __launch_bounds__(64) void k_vanish_load(const uint32_t* g2pp, const uint32_t* w_mont, int scale, int log_n,
                                                    int h, uint32_t* work) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (2u << h) - 2u) return;
    uint32_t i;
    const uint32_t d = van_level(t, &i);
    const uint32_t m = 1u << (log_n - d);
    const uint32_t src = i ? (i - 1) * m : (1u << log_n) - m;
    bool inf;
    const g2_aff a = g2_load_canon(g2pp + G2_WORDS * (size_t)src, &inf);
    g2_xyzz P = inf ? g2_inf() : g2_from_aff(a);
    if (scale) P = g2_mul(P, f_from_mont(f_pow_u64(f_load<FrP>(w_mont), src)));
    g2x_store(work + G2X_WORDS * (size_t)((1u << d) - 2u + bitrev_bits(i, (int)d)), P);
}

// one radix-2 DIT stage s of the inverse-root DFT (fft_domain with group_gen and
// group_gen_inv swapped, generate.rs:254-260) in every level d >= s: (u, v) ->
// (u + w v, u - w v), w = tw_inv[j << (log_n - s)] = w_(2^s)^-j; one thread per
// butterfly.  Level d's 2^(d-1) butterflies are threads [2^(d-1) - 2^(s-1), 2^d - 2^(s-1)).
// Stage 1's twiddle is w^0 = 1: no multiplication.
__global__ __launch_bounds__(64) void k_vanish_stage(uint32_t* work, const uint32_t* tw_inv, int s, int log_n, int h) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t h0 = 1u << (s - 1);
    if (t >= (1u << h) - h0) return;
    const uint32_t d = 32u - (uint32_t)__builtin_clz(t + h0);  // floor(log2(t + h0)) + 1
    const uint32_t td = t - ((1u << (d - 1)) - h0);
    uint32_t* lv = work + G2X_WORDS * (size_t)((1u << d) - 2u);
    const uint32_t j = td & (h0 - 1);
    const uint32_t i0 = ((td >> (s - 1)) << s) | j, i1 = i0 + h0;
    const g2_xyzz u = g2x_load(lv + G2X_WORDS * (size_t)i0);
    g2_xyzz v = g2x_load(lv + G2X_WORDS * (size_t)i1);
    if (s > 1) v = g2_mul(v, f_from_mont(f_load<FrP>(tw_inv + 8 * ((size_t)j << (log_n - s)))));
    g2_xyzz nv = v;
    nv.Y = q2_neg(v.Y);
    g2x_store(lv + G2X_WORDS * (size_t)i0, g2_add(u, v));
    g2x_store(lv + G2X_WORDS * (size_t)i1, g2_add(u, nv));
}

// enact: DFT output i of level d -> affine Montgomery at stored index bitrev_d(i) of
// level d, for every level; an infinite output (a degenerate tau) sets *bad
__global__ __launch_bounds__(64) void k_vanish_store(const uint32_t* work, int h, uint32_t* out, int* bad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (2u << h) - 2u) return;
    uint32_t i;
    const uint32_t d = van_level(t, &i);
    bool inf;
    g2_aff a = g2_to_aff(g2x_load(work + G2X_WORDS * (size_t)t), &inf);
    if (inf) *bad = 1;
    g2_store(out + G2_WORDS * (size_t)((1u << d) - 2u + bitrev_bits(i, (int)d)), a);
}

// ---------------------------------------------------------------------------
// arkworks serialized forms (ark-serialize 0.4, short Weierstrass affine):
//   uncompressed  x || y, y carrying the flags in the top bits of its last byte
//   compressed    x carrying the flags
//   flags         0x40 point at infinity (coordinates zero), 0x80 when y > -y
//                 (Fq: y > (p - 1) / 2; Fq2: c1 decides unless zero, then c0)
// Coordinates canonical little-endian; Fq2 = c0 || c1 (the flags on c1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool fq_gt_half(const fq& a) {  // a > (p - 1) / 2 for canonical a
    const uint32_t H[8] = {0x6c3e7ea3u, 0x9e10460bu, 0xb438e546u, 0xcbc0b548u,
                           0x40c0ac2eu, 0xdc2822dbu, 0x7098d014u, 0x18322739u};
    for (int k = 7; k >= 0; k--)
        if (a.v[k] != H[k]) return a.v[k] > H[k];
    return false;
}

// n G1 points (16 u32 each, Montgomery when mont != 0, else canonical; all-zero words =
// infinity) -> ark form, 64 B (uncompressed) or 32 B (compressed) per point
__global__ void k_g1_ark(const uint32_t* pts, uint32_t n, int mont, int compressed, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fq x = f_load<FqP>(pts + 16 * (size_t)i), y = f_load<FqP>(pts + 16 * (size_t)i + 8);
    const bool inf = f_is_zero(x) && f_is_zero(y);
    if (mont) {
        x = f_from_mont(x);
        y = f_from_mont(y);
    }
    const uint32_t flag = inf ? 0x40000000u : (fq_gt_half(y) ? 0x80000000u : 0u);
    uint32_t* o = out + (compressed ? 8 : 16) * (size_t)i;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = inf ? 0u : x.v[k];
    if (compressed) {
        o[7] |= flag;
        return;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) o[8 + k] = inf ? 0u : y.v[k];
    o[15] |= flag;
}

// n G2 points (32 u32 each, Montgomery or canonical) -> ark form, 128 B / 64 B per point
__global__ void k_g2_ark(const uint32_t* pts, uint32_t n, int mont, int compressed, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    g2_aff a = g2_load(pts + G2_WORDS * (size_t)i);
    const bool inf = q2_is_zero(a.x) && q2_is_zero(a.y);
    if (mont) a = {{f_from_mont(a.x.c0), f_from_mont(a.x.c1)}, {f_from_mont(a.y.c0), f_from_mont(a.y.c1)}};
    const bool neg = f_is_zero(a.y.c1) ? fq_gt_half(a.y.c0) : fq_gt_half(a.y.c1);
    const uint32_t flag = inf ? 0x40000000u : (neg ? 0x80000000u : 0u);
    const fq c[4] = {a.x.c0, a.x.c1, a.y.c0, a.y.c1};
    const int nc = compressed ? 2 : 4;
    uint32_t* o = out + 8 * nc * (size_t)i;
    for (int q = 0; q < nc; q++)
#pragma unroll
        for (int k = 0; k < 8; k++) o[8 * q + k] = inf ? 0u : c[q].v[k];
    o[8 * nc - 1] |= flag;
}

// count Fq elements: canonical -> Montgomery (to_mont != 0) or back
__global__ void k_fq_convert(const uint32_t* in, uint32_t* out, uint64_t count, int to_mont) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const fq a = f_load<FqP>(in + 8 * i);
    f_store(out + 8 * i, to_mont ? f_to_mont(a) : f_from_mont(a));
}

}  // namespace zgr
