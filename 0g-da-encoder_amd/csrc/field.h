// BN254 prime-field arithmetic for gfx950: 8 x 32-bit limbs, Montgomery form with
// R = 2^256 (the same internal form arkworks ark-bn254 0.4 uses, so parameter files
// in the bamt layout, crates/amt/src/prove_params/fast_serde_bn254.rs:139-187, are
// consumed verbatim).  Multiplication is CIOS: v_mad_u64_u32 carry chains held in
// VGPRs.  Every value is kept fully reduced (< modulus), so zero tests and equality
// are plain limb compares.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zgr {

struct FrP {  // scalar field r
    static constexpr uint32_t M[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                      0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static constexpr uint32_t ONE[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                        0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    static constexpr uint32_t R2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                       0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
    static constexpr uint32_t INV = 0xefffffffu;
};

struct FqP {  // base field p
    static constexpr uint32_t M[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                      0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static constexpr uint32_t ONE[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                        0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    static constexpr uint32_t R2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                       0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
    static constexpr uint32_t INV = 0xe4866389u;
};

template <class P>
struct fe {
    uint32_t v[8];
};

using fr = fe<FrP>;
using fq = fe<FqP>;

template <class P>
__device__ __forceinline__ fe<P> f_zero() {
    fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
}

template <class P>
__device__ __forceinline__ fe<P> f_one() {
    fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = P::ONE[i];
    return r;
}

template <class P>
__device__ __forceinline__ bool f_is_zero(const fe<P>& a) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o |= a.v[i];
    return o == 0;
}

template <class P>
__device__ __forceinline__ bool f_eq(const fe<P>& a, const fe<P>& b) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o |= a.v[i] ^ b.v[i];
    return o == 0;
}

// r = (s >= M) ? s - M : s, for s < 2M
template <class P>
__device__ __forceinline__ fe<P> f_reduce_once(const uint32_t s[8]) {
    uint32_t d[8];
    uint32_t br = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        uint64_t x = (uint64_t)s[j] - P::M[j] - br;
        d[j] = (uint32_t)x;
        br = (uint32_t)(x >> 63);
    }
    fe<P> r;
#pragma unroll
    for (int j = 0; j < 8; j++) r.v[j] = br ? s[j] : d[j];
    return r;
}

template <class P>
__device__ __forceinline__ fe<P> f_add(const fe<P>& a, const fe<P>& b) {
    uint32_t s[8];
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        c = (uint64_t)a.v[j] + b.v[j] + (c >> 32);
        s[j] = (uint32_t)c;
    }
    return f_reduce_once<P>(s);  // a + b < 2M < 2^255: no carry out of limb 7
}

template <class P>
__device__ __forceinline__ fe<P> f_dbl(const fe<P>& a) {
    return f_add<P>(a, a);
}

template <class P>
__device__ __forceinline__ fe<P> f_sub(const fe<P>& a, const fe<P>& b) {
    uint32_t d[8];
    uint32_t br = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        uint64_t x = (uint64_t)a.v[j] - b.v[j] - br;
        d[j] = (uint32_t)x;
        br = (uint32_t)(x >> 63);
    }
    // if borrow, add M back
    uint32_t mask = 0u - br;
    fe<P> r;
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        c = (uint64_t)d[j] + (P::M[j] & mask) + (c >> 32);
        r.v[j] = (uint32_t)c;
    }
    return r;
}

template <class P>
__device__ __forceinline__ fe<P> f_neg(const fe<P>& a) {
    return f_sub<P>(f_zero<P>(), a);
}

// Montgomery product a * b * 2^-256 mod M (CIOS, 32-bit limbs)
template <class P>
__device__ __forceinline__ fe<P> f_mul(const fe<P>& a, const fe<P>& b) {
    uint32_t t[10];
#pragma unroll
    for (int j = 0; j < 10; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            c = (uint64_t)a.v[j] * b.v[i] + t[j] + (c >> 32);
            t[j] = (uint32_t)c;
        }
        uint64_t s = (uint64_t)t[8] + (c >> 32);
        t[8] = (uint32_t)s;
        t[9] = (uint32_t)(s >> 32);
        uint32_t m = t[0] * P::INV;
        c = (uint64_t)m * P::M[0] + t[0];
#pragma unroll
        for (int j = 1; j < 8; j++) {
            c = (uint64_t)m * P::M[j] + t[j] + (c >> 32);
            t[j - 1] = (uint32_t)c;
        }
        s = (uint64_t)t[8] + (c >> 32);
        t[7] = (uint32_t)s;
        t[8] = t[9] + (uint32_t)(s >> 32);
    }
    // M < 2^254 < R/4, so the CIOS result is < 2M and t[8] == 0
    return f_reduce_once<P>(t);
}

template <class P>
__device__ __forceinline__ fe<P> f_sqr(const fe<P>& a) {
    return f_mul<P>(a, a);
}

template <class P>
__device__ __forceinline__ fe<P> f_to_mont(const fe<P>& a) {
    fe<P> r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.v[i] = P::R2[i];
    return f_mul<P>(a, r2);
}

template <class P>
__device__ __forceinline__ fe<P> f_from_mont(const fe<P>& a) {
    fe<P> one = f_zero<P>();
    one.v[0] = 1;
    return f_mul<P>(a, one);
}

// a^e for a 64-bit exponent (square-and-multiply, MSB first)
template <class P>
__device__ __forceinline__ fe<P> f_pow_u64(const fe<P>& a, uint64_t e) {
    fe<P> r = f_one<P>();
    for (int i = 63; i >= 0; i--) {
        r = f_sqr<P>(r);
        if ((e >> i) & 1) r = f_mul<P>(r, a);
    }
    return r;
}

// a^(M-2) = a^-1 (Fermat); 0 maps to 0
template <class P>
__device__ __noinline__ fe<P> f_inv(const fe<P>& a) {
    uint32_t e[8];
    uint32_t br = 2;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        uint64_t x = (uint64_t)P::M[j] - br;
        e[j] = (uint32_t)x;
        br = (uint32_t)(x >> 63);
    }
    fe<P> r = f_one<P>();
    for (int j = 7; j >= 0; j--) {
        for (int b = 31; b >= 0; b--) {
            r = f_sqr<P>(r);
            if ((e[j] >> b) & 1) r = f_mul<P>(r, a);
        }
    }
    return r;
}

template <class P>
__device__ __forceinline__ fe<P> f_load(const uint32_t* p) {
    fe<P> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = p[i];
    return r;
}

template <class P>
__device__ __forceinline__ void f_store(uint32_t* p, const fe<P>& a) {
#pragma unroll
    for (int i = 0; i < 8; i++) p[i] = a.v[i];
}

}  // namespace zgr
