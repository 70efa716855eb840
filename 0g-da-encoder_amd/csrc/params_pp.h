// Device kernels of AMTParams::from_pp (crates/amt/src/prove_params/generate.rs:139-232):
// prove bases from powers of tau [tau^i] G1 and [tau^(H+i)] G1 without knowing tau, as
// a ceremony file provides them.  Included once, by zgr.hip, after kernels.h (the EC-FFT
// stages k_ec_dit_stage and x29_mul_scalar).
//
//   coset scaling     g1pp'[i] = w^i g1pp[i], w = 1 / cf_c            (generate.rs:152-170)
//   basis             iDFT(g1pp') = fast_ifft                          (gen_basis, :196-202)
//   quotients[d-1]    DFT(coeff) / N, coeff[i] = g1pp'[m - i] for i = 1..m, m = N / 2^d
//                                                                      (gen_quotients, :204-232)
//   high_basis        iDFT(high_g1pp')
//   enact             affine, index_reverse                            (:122-126)
// The transforms run as radix-2 DIT stages over reduced-radix XYZZ points: a
// bit-reversed load, log N stages of butterflies with one 254-bit scalar
// multiplication each (signed 3-bit windows over the GLV halves), then the store.
// Every transform ends with the factor 1 / N (the iDFT's size_inv, and gen_quotients'
// explicit division); the transforms are linear, so the factor is applied once to the
// two inputs (folded into the coset scaling) instead of to the 12 outputs.
#pragma once

namespace zgr {

// x29 point of a canonical affine input (64 zero bytes = infinity) times w^i / N, w^i
// from the Montgomery w by a 64-bit exponent (scale = 0, coset 0: w = 1, the factor
// 1 / N alone).
__global__ __launch_bounds__(EC_THREADS) void k_pp_scale(const uint32_t* in, const uint32_t* w_mont, int scale,
                                                         const uint32_t* consts, uint32_t n, uint32_t* out) {
    __shared__ uint32_t tab[EC_LDS_WORDS];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x29 p = ec29_from_canon(in + 16 * (size_t)i);
    fr k = f_load<FrP>(consts + 8 * K_NINV);
    if (scale) k = f_mul(k, f_pow_u64(f_load<FrP>(w_mont), i));
    p = x29_mul_scalar(p, f_from_mont(k), tab);
    x29_store(out + X29_WORDS * (size_t)i, p);
}

// Bit-reversed load of the transform input: pts[p] = src[q], q = bitrev(p); with m > 0
// the quotient coefficients instead: pts[p] = src[m - q] for 1 <= q <= m, else infinity.
__global__ void k_pp_load(const uint32_t* src, uint32_t* pts, int log_n, uint32_t m) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (1u << log_n)) return;
    const uint32_t q = bitrev_bits(p, log_n);
    uint32_t* d = pts + X29_WORDS * (size_t)p;
    if (m && (q == 0 || q > m)) {
        x29_store(d, x29_inf());
        return;
    }
    const uint32_t* s = src + X29_WORDS * (size_t)(m ? m - q : q);
#pragma unroll
    for (int k = 0; k < X29_WORDS; k++) d[k] = s[k];
}

// Transform output p -> param layout (x_mont || y_mont) at index bitrev(p) (enact's
// index_reverse), scaled by N^-1 when scale != 0 (from_pp's inputs already carry it, so
// it passes 0); infinity is stored as zeros.
__global__ __launch_bounds__(EC_THREADS) void k_pp_store(const uint32_t* pts, const uint32_t* consts, int scale,
                                                         int log_n, uint32_t* line) {
    __shared__ uint32_t tab[EC_LDS_WORDS];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (1u << log_n)) return;
    x29 v = x29_load(pts + X29_WORDS * (size_t)p);
    if (scale) v = x29_mul_scalar(v, f_from_mont(f_load<FrP>(consts + 8 * K_NINV)), tab);
    bool inf;
    fq29 x, y;
    x29_to_aff(v, x, y, &inf);
    g1_aff a;
    a.x = q_to_fq(x);
    a.y = q_to_fq(y);
    uint32_t* o = line + 16 * (size_t)bitrev_bits(p, log_n);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        o[k] = inf ? 0u : a.x.v[k];
        o[8 + k] = inf ? 0u : a.y.v[k];
    }
}

// w = 1 / cf_c = cf_c^(2^28 - 1) (cf_c is a power of the 2^28-th root), Montgomery
__global__ void k_pp_coset_w(const uint32_t* consts, int coset, uint32_t* w) {
    if (threadIdx.x || blockIdx.x) return;
    f_store(w, f_pow_u64(f_load<FrP>(consts + 8 * (K_CF + coset)), (1ull << 28) - 1));
}

}  // namespace zgr
