// BN254 scalar field Fr in the reduced-radix Montgomery form of field29.h: 9 limbs of
// 29 bits, R' = 2^261.  A 29 x 29-bit product is below 2^58, so every partial product
// of a multiplication is one v_mad_u64_u32 into a 64-bit column and the Montgomery
// reduction runs once per product instead of once per limb row.  Used by the NTT
// passes and by the recovery kernels.
//
// Values held in the 2^256 Montgomery form (arkworks' in-memory form) keep that form
// under r29_mul with an operand in the 2^261 form: (a 2^256)(b 2^261) / 2^261 =
// (a b) 2^256.  So data can stay in the 2^256 form with only its limbs repacked,
// while constants (twiddles, scale factors) are stored in the 2^261 form.
#pragma once
#include "field.h"
#include "field29.h"

namespace zgr {

// Fr in 29-bit limbs: products of two limbs stay below 2^58, so a 64-bit column takes
// 63 of them (7 terms of 9 products each) between carry passes (the recovery GEMM uses
// that headroom), and the Montgomery reduction runs once per output.
struct R29 {
    static constexpr uint32_t P[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
    static constexpr uint32_t INV = 0x0fffffffu;  // -r^-1 mod 2^29
    // r * 2^k for k = 5..0 (the final reduction of a value below 2^259)
    static constexpr uint32_t RK[6][9] = {
        {0x00000020u, 0x01f593f0u, 0x0b848a1fu, 0x1a121e6eu, 0x10ba5067u, 0x1b681815u, 0x14dc2822u, 0x0b84c680u, 0x060c89ceu},
        {0x00000010u, 0x10fac9f8u, 0x05c2450fu, 0x1d090f37u, 0x185d2833u, 0x0db40c0au, 0x0a6e1411u, 0x05c26340u, 0x030644e7u},
        {0x00000008u, 0x187d64fcu, 0x12e12287u, 0x1e84879bu, 0x0c2e9419u, 0x16da0605u, 0x05370a08u, 0x12e131a0u, 0x01832273u},
        {0x00000004u, 0x1c3eb27eu, 0x19709143u, 0x1f4243cdu, 0x16174a0cu, 0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u},
        {0x00000002u, 0x1e1f593fu, 0x1cb848a1u, 0x0fa121e6u, 0x0b0ba506u, 0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu},
        {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}};
};

constexpr uint32_t R29_MASK = 0x1fffffffu;

// 256-bit value -> 9 limbs of 29 bits
__device__ __forceinline__ void fr29_limbs(const fr& a, uint32_t* o) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int bit = 29 * k, w = bit >> 5, sh = bit & 31;
        uint32_t v = a.v[w] >> sh;
        if (sh > 3 && w + 1 < 8) v |= a.v[w + 1] << (32 - sh);
        o[k] = v & R29_MASK;
    }
}

// Fr in the 29-bit Montgomery form (R' = 2^261), the scalar-field twin of field29.h's Fq
// form: one 9 x 9 v_mad_u64_u32 product and one reduction per multiplication instead
// of the 8-limb CIOS chain.
struct R29M {
    static constexpr uint32_t ONE[9] = {0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu,
                                        0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};  // 2^261 mod r
    // 2^266 mod r: r29_mul(v, C_TO) turns a 2^256-Montgomery value into the 2^261 form
    static constexpr uint32_t C_TO[9] = {0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau,
                                         0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
    // 2^256 mod r: r29_mul(v, C_FROM) turns it back
    static constexpr uint32_t C_FROM[9] = {0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu,
                                           0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
};

// a b / 2^261 mod r for a, b below 4r (a b < r 2^261), result in [0, 2r)
__device__ __forceinline__ fq29 r29_mul(const fq29& a, const fq29& b) {
    uint64_t T[18];
#pragma unroll
    for (int k = 0; k < 18; k++) T[k] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)a.l[i] * b.l[j];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t m = ((uint32_t)T[i] * R29::INV) & R29_MASK;
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)m * R29::P[j];
        T[i + 1] += T[i] >> 29;
    }
    fq29 o;
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const uint64_t x = T[9 + k] + c;
        o.l[k] = (uint32_t)x & R29_MASK;
        c = x >> 29;
    }
    return o;
}

// a - b + 2r for a, b in [0, 2r): (0, 4r), for multiplications only
__device__ __forceinline__ fq29 r29_sub_lazy(const fq29& a, const fq29& b) {
    fq29 o;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int32_t x = (int32_t)a.l[k] + (int32_t)R29::RK[4][k] - (int32_t)b.l[k] + c;  // RK[4] = 2r
        o.l[k] = (uint32_t)x & R29_MASK;
        c = x >> 29;
    }
    return o;
}

// a^(r-2) = a^-1 (Fermat) for a in [0, 4r); the exponent is the same on every lane
__device__ __noinline__ fq29 r29_inv(const fq29& a) {
    uint32_t e[8];
    uint32_t br = 2;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t x = (uint64_t)FrP::M[j] - br;
        e[j] = (uint32_t)x;
        br = (uint32_t)(x >> 63);
    }
    fq29 r = q_const(R29M::ONE);
    for (int j = 7; j >= 0; j--)
        for (int b = 31; b >= 0; b--) {
            r = r29_mul(r, r);
            if ((e[j] >> b) & 1) r = r29_mul(r, a);
        }
    return r;
}

// 29-bit limbs of a value below 2^256 -> 8 little-endian u32
__device__ __forceinline__ void fr29_pack(const fq29& v, uint32_t* o) {
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const int bit = 32 * w, k = bit / 29, sh = bit % 29;
        uint64_t x = (uint64_t)v.l[k] >> sh;
        if (k + 1 < 9) x |= (uint64_t)v.l[k + 1] << (29 - sh);
        if (k + 2 < 9 && 29 - sh + 29 < 32) x |= (uint64_t)v.l[k + 2] << (58 - sh);
        o[w] = (uint32_t)x;
    }
}

__device__ __forceinline__ fq29 fr29_unpack(const uint32_t* p) {
    fq29 v;
    fr29_limbs(f_load<FrP>(p), v.l);
    return v;
}

// a + b with carries (normalised limbs); the caller keeps the sum below 2^261
__device__ __forceinline__ fq29 r29_add(const fq29& a, const fq29& b) {
    fq29 o;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const uint32_t x = a.l[k] + b.l[k] + c;
        o.l[k] = x & R29_MASK;
        c = x >> 29;
    }
    return o;
}

// a - m if a >= m, else a (normalised limbs, m a 9-limb constant)
__device__ __forceinline__ fq29 r29_csub(const fq29& a, const uint32_t (&m)[9]) {
    fq29 d;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int32_t x = (int32_t)a.l[k] - (int32_t)m[k] + c;
        d.l[k] = (uint32_t)x & R29_MASK;
        c = x >> 29;
    }
    const bool keep = c < 0;
    fq29 o;
#pragma unroll
    for (int k = 0; k < 9; k++) o.l[k] = keep ? a.l[k] : d.l[k];
    return o;
}

// r - a for 0 < a < r (a twiddle's negative)
__device__ __forceinline__ fq29 r29_neg_canon(const fq29& a) {
    fq29 o;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int32_t x = (int32_t)R29::P[k] - (int32_t)a.l[k] + c;
        o.l[k] = (uint32_t)x & R29_MASK;
        c = x >> 29;
    }
    return o;
}

// floor(2^40 / (r_8 + 1)), r_8 = the top 29-bit limb of r
constexpr uint64_t R29_QMAGIC = (1ull << 40) / (0x0030644eull + 1);

// a mod r, canonical, for any a below 2^259.  q = floor(a_8 M / 2^40) from the top limb
// satisfies q <= a / (r_8 + 1) < a / r and a / r - q < 2.0001, so a - q r lies in
// [0, 3r) and two conditional subtractions finish it.
__device__ __forceinline__ fq29 r29_reduce(const fq29& a) {
    const uint32_t q = (uint32_t)(((uint64_t)a.l[8] * R29_QMAGIC) >> 40);
    fq29 o;
    int64_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int64_t x = (int64_t)a.l[k] - (int64_t)q * R29::P[k] + c;
        o.l[k] = (uint32_t)x & R29_MASK;
        c = x >> 29;
    }
    o = r29_csub(o, R29::P);
    return r29_csub(o, R29::P);
}

}  // namespace zgr
