// BN254 base field in a reduced-radix Montgomery form for the MSM hot loops:
// 9 limbs of 29 bits in 32-bit VGPRs, R' = 2^261.
//
// Why: a 29 x 29-bit product is < 2^58, so a column of up to 18 products plus a
// carry stays below 2^64.  Every partial product is then exactly one v_mad_u64_u32
// accumulating into a 64-bit column register in place, with no carry-out handling
// and no zero-extension moves; carries are resolved once per column.  This is the
// gfx950-native shape of the multiplication (the 32-bit-limb CIOS form needs a
// 64-bit add and moves around every v_mad_u64_u32).
//
// Invariants: limbs normalised (< 2^29); values kept in [0, 2p).  Montgomery
// multiplication of two such values returns [0, 2p) without a final subtraction
// (4p < R').  Zero tests accept both 0 and p.
#pragma once
#include "field.h"

namespace zgr {

constexpr uint32_t M29 = 0x1fffffffu;

struct Q29 {
    static constexpr uint32_t P[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
    static constexpr uint32_t P2[9] = {0x10f9fa8eu, 0x0208c16du, 0x18e5469eu, 0x05aa45a1u, 0x0b0bb2f0u,
                                       0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu};
    static constexpr uint32_t P4[9] = {0x01f3f51cu, 0x041182dbu, 0x11ca8d3cu, 0x0b548b43u, 0x161765e0u,
                                       0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u};
    static constexpr uint32_t P6[9] = {0x12edefaau, 0x061a4448u, 0x0aafd3dau, 0x10fed0e5u, 0x012318d0u,
                                       0x11238484u, 0x03e94786u, 0x1628e538u, 0x012259d6u};
    static constexpr uint32_t P8[9] = {0x03e7ea38u, 0x082305b6u, 0x03951a78u, 0x16a91687u, 0x0c2ecbc0u,
                                       0x16da0605u, 0x05370a08u, 0x12e131a0u, 0x01832273u};
    // p_0^-1 mod 2^29 (p_0 = P[0] is odd): a normalised value's low limb l equals that of
    // k p (mod 2^29) iff (l * P0_INV) mod 2^29 == k
    static constexpr uint32_t P0_INV = 0x1b799c77u;
    static constexpr uint32_t ONE[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                                        0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
    // 2^5 * R' mod p: mont29(v, C_TO) = v * 2^5, i.e. 2^256-Montgomery -> 2^261-Montgomery
    static constexpr uint32_t C_TO[9] = {0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u,
                                         0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
    // 2^256 mod p: mont29(v, C_FROM) = v / 2^5, i.e. 2^261-Montgomery -> 2^256-Montgomery
    static constexpr uint32_t C_FROM[9] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u,
                                           0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
    static constexpr uint32_t INV = 0x04866389u;  // -p^-1 mod 2^29
};

struct fq29 {
    uint32_t l[9];
};

__device__ __forceinline__ fq29 q_const(const uint32_t (&c)[9]) {
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = c[k];
    return r;
}

__device__ __forceinline__ fq29 q_zero() {
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = 0;
    return r;
}

__device__ __forceinline__ fq29 q_one() { return q_const(Q29::ONE); }

// Montgomery product a * b / 2^261; result in [0, 2p) whenever a * b < p 2^261 (about
// 169 p^2: both inputs below 8p, or one below 10p and the other below 16p, qualify),
// since the result is below a b / 2^261 + p.  Inputs below 4p are the common case.
__device__ __forceinline__ fq29 q_mul(const fq29& a, const fq29& b) {
    uint64_t T[18];
#pragma unroll
    for (int k = 0; k < 18; k++) T[k] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)a.l[i] * b.l[j];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint32_t m = ((uint32_t)T[i] * Q29::INV) & M29;
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)m * Q29::P[j];
        T[i + 1] += T[i] >> 29;
    }
    fq29 r;
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        uint64_t x = T[9 + k] + c;
        r.l[k] = (uint32_t)x & M29;
        c = x >> 29;
    }
    return r;
}

// a^2 / 2^261: 45 products (cross terms with the doubled operand 2 a_i < 2^30, at most
// 4 of them plus one square per column before the reduction)
__device__ __forceinline__ fq29 q_sqr(const fq29& a) {
    uint64_t T[18];
#pragma unroll
    for (int k = 0; k < 18; k++) T[k] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        T[2 * i] += (uint64_t)a.l[i] * a.l[i];
        const uint32_t d = a.l[i] << 1;
#pragma unroll
        for (int j = i + 1; j < 9; j++) T[i + j] += (uint64_t)d * a.l[j];
    }
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint32_t m = ((uint32_t)T[i] * Q29::INV) & M29;
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)m * Q29::P[j];
        T[i + 1] += T[i] >> 29;
    }
    fq29 r;
    uint64_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        uint64_t x = T[9 + k] + c;
        r.l[k] = (uint32_t)x & M29;
        c = x >> 29;
    }
    return r;
}

// (a * b + c * d) / 2^261 with one Montgomery reduction, for a, b, c, d in [0, 4p):
// the sum is below 32 p^2, so the result is below 32 p^2 / 2^261 + p < 1.25 p; a
// column holds at most 27 products below 2^58 (< 2^63).
__device__ __forceinline__ fq29 q_mul2(const fq29& a, const fq29& b, const fq29& c, const fq29& d) {
    uint64_t T[18];
#pragma unroll
    for (int k = 0; k < 18; k++) T[k] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)a.l[i] * b.l[j] + (uint64_t)c.l[i] * d.l[j];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint32_t m = ((uint32_t)T[i] * Q29::INV) & M29;
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)m * Q29::P[j];
        T[i + 1] += T[i] >> 29;
    }
    fq29 r;
    uint64_t cy = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        uint64_t x = T[9 + k] + cy;
        r.l[k] = (uint32_t)x & M29;
        cy = x >> 29;
    }
    return r;
}

// s in [0, 4p) with normalised limbs -> [0, 2p)
__device__ __forceinline__ fq29 q_reduce_2p(const fq29& s) {
    fq29 d;
    int32_t br = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)s.l[k] - (int32_t)Q29::P2[k] - br;
        br = (x >> 31) & 1;
        d.l[k] = (uint32_t)x & M29;
    }
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = br ? s.l[k] : d.l[k];
    return r;
}

__device__ __forceinline__ fq29 q_add(const fq29& a, const fq29& b) {
    fq29 s;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        uint32_t x = a.l[k] + b.l[k] + c;
        s.l[k] = x & M29;
        c = x >> 29;
    }
    return q_reduce_2p(s);
}

__device__ __forceinline__ fq29 q_dbl(const fq29& a) { return q_add(a, a); }

// a - b + 2p in (0, 4p), then reduced to [0, 2p)
__device__ __forceinline__ fq29 q_sub(const fq29& a, const fq29& b) {
    fq29 s;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)a.l[k] + (int32_t)Q29::P2[k] - (int32_t)b.l[k] + c;
        s.l[k] = (uint32_t)x & M29;
        c = x >> 29;  // arithmetic: -1, 0 or 1
    }
    return q_reduce_2p(s);
}

__device__ __forceinline__ fq29 q_neg(const fq29& a) { return q_sub(q_zero(), a); }

// neg ? p - a : a for a in [0, p]; p - a needs no reduction, so the result stays in [0, p]
__device__ __forceinline__ fq29 q_cneg_canon(const fq29& a, bool neg) {
    fq29 r;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)Q29::P[k] - (int32_t)a.l[k] + c;
        r.l[k] = neg ? ((uint32_t)x & M29) : a.l[k];
        c = x >> 29;  // arithmetic: -1 or 0
    }
    return r;
}

// a - b + 2p in (0, 4p) without the final reduction: only for values that feed a
// multiplication (which accepts inputs below 4p) or q_is_zero4
__device__ __forceinline__ fq29 q_sub_lazy(const fq29& a, const fq29& b) {
    fq29 s;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)a.l[k] + (int32_t)Q29::P2[k] - (int32_t)b.l[k] + c;
        s.l[k] = (uint32_t)x & M29;
        c = x >> 29;
    }
    return s;
}

// a - b + 8p for a in [0, 2p), b in [0, 8p): (0, 10p) with normalised limbs, for
// values that only feed multiplications (see q_mul)
__device__ __forceinline__ fq29 q_sub_lazy8(const fq29& a, const fq29& b) {
    fq29 s;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)a.l[k] + (int32_t)Q29::P8[k] - (int32_t)b.l[k] + c;
        s.l[k] = (uint32_t)x & M29;
        c = x >> 29;
    }
    return s;
}

// r2 - t - 2 u + 6p for r2, t, u in [0, 2p): (0, 8p) in one carry pass (the X3 of an
// XYZZ addition, R^2 - PPP - 2Q, left unreduced)
__device__ __forceinline__ fq29 q_x3_lazy(const fq29& r2, const fq29& t, const fq29& u) {
    fq29 s;
    int32_t c = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)r2.l[k] + (int32_t)Q29::P6[k] - (int32_t)t.l[k] - 2 * (int32_t)u.l[k] + c;
        s.l[k] = (uint32_t)x & M29;
        c = x >> 29;  // arithmetic: -3 .. 1
    }
    return s;
}

// [0, 8p) -> [0, 2p)
__device__ __forceinline__ fq29 q_reduce_8p(const fq29& s) {
    fq29 d;
    int32_t br = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)s.l[k] - (int32_t)Q29::P4[k] - br;
        br = (x >> 31) & 1;
        d.l[k] = (uint32_t)x & M29;
    }
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = br ? s.l[k] : d.l[k];
    return q_reduce_2p(r);
}

// [0, 2p) -> [0, p)
__device__ __forceinline__ fq29 q_canon(const fq29& s) {
    fq29 d;
    int32_t br = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int32_t x = (int32_t)s.l[k] - (int32_t)Q29::P[k] - br;
        br = (x >> 31) & 1;
        d.l[k] = (uint32_t)x & M29;
    }
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.l[k] = br ? s.l[k] : d.l[k];
    return r;
}

// x in [0, 2p): x == 0 mod p iff x == 0 or x == p
__device__ __forceinline__ bool q_is_zero(const fq29& a) {
    uint32_t z = 0, e = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        z |= a.l[k];
        e |= a.l[k] ^ Q29::P[k];
    }
    return z == 0 || e == 0;
}

// x in [0, 4p): x == 0 mod p iff x is 0, p, 2p or 3p
__device__ __forceinline__ bool q_is_zero4(const fq29& a) { return q_is_zero(q_reduce_2p(a)); }

// a^(p-2) = a^-1 (Fermat; 0 maps to 0) for a in [0, 4p), result in [0, 2p).  The
// exponent bits are the same on every lane, so the branch is uniform.
__device__ __noinline__ fq29 q_inv(const fq29& a) {
    uint32_t e[8];
    uint32_t br = 2;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        uint64_t x = (uint64_t)FqP::M[j] - br;
        e[j] = (uint32_t)x;
        br = (uint32_t)(x >> 63);
    }
    fq29 r = q_const(Q29::ONE);
    for (int j = 7; j >= 0; j--)
        for (int b = 31; b >= 0; b--) {
            r = q_sqr(r);
            if ((e[j] >> b) & 1) r = q_mul(r, a);
        }
    return r;
}

// 8 x 32-bit limbs (any value < 2^256) -> 9 x 29-bit limbs of the same integer
__device__ __forceinline__ fq29 q_unpack(const uint32_t w[8]) {
    fq29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int bit = 29 * k, i = bit >> 5, s = bit & 31;
        uint32_t v = w[i] >> s;
        if (s > 3 && i + 1 < 8) v |= w[i + 1] << (32 - s);
        r.l[k] = v & M29;
    }
    return r;
}

// 9 x 29-bit limbs (value < 2^256) -> 8 x 32-bit limbs
__device__ __forceinline__ void q_pack(const fq29& a, uint32_t w[8]) {
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int bit = 29 * k, i = bit >> 5, s = bit & 31;
        w[i] |= a.l[k] << s;
        if (s > 3 && i + 1 < 8) w[i + 1] |= a.l[k] >> (32 - s);
    }
}

// 2^256-Montgomery fq (arkworks form, fully reduced) -> 2^261-Montgomery fq29
__device__ __forceinline__ fq29 q_from_fq(const fq& a) { return q_mul(q_unpack(a.v), q_const(Q29::C_TO)); }

// fq29 -> fully reduced 2^256-Montgomery fq
__device__ __forceinline__ fq q_to_fq(const fq29& a) {
    fq29 m = q_mul(a, q_const(Q29::C_FROM));  // [0, 2p)
    fq r;
    q_pack(m, r.v);
    return f_reduce_once<FqP>(r.v);
}

}  // namespace zgr
