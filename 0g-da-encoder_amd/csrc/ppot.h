// Ceremony import: the points of a perpetual-powers-of-tau file decoded and checked on
// the device.  crates/ppot2ark/src/main.rs:45-119 reads them through the powersoftau
// crate of kobigurk/phase2-bn254 (rev dd6b966, BatchedAccumulator::read_chunk with
// CheckForCorrectness::Yes; an absent dependency).  Its point encodings are those of
// pairing_ce's bn256 EncodedPoint:
//   G1 uncompressed  x || y            32 B big-endian each (64 B)
//   G1 compressed    x                 (32 B)
//   G2 uncompressed  x.c1 || x.c0 || y.c1 || y.c0   (128 B)
//   G2 compressed    x.c1 || x.c0      (64 B)
// with flags in the two top bits of the first byte: 0x40 marks the point at infinity
// (every other bit zero), and in a compressed point 0x80 says that y is the larger of
// y and -y (canonical integers; Fq2 ordered by c1, then c0).  An uncompressed point with
// 0x80 set is the wrong compression mode.  A checked read rejects coordinates >= p,
// points off the curve, G2 points outside the order-r subgroup, and the point at
// infinity (the accumulator never holds it).
//
// One lane per point: the big-endian words are byte-swapped into limbs, the square root
// of a compressed point's x^3 + b is an exponentiation (p = 3 mod 4; over Fq2 through
// the norm), and the G2 subgroup test uses the endomorphism psi (ppot_g2_in_subgroup).  Output points
// are canonical affine limbs, the library's memory form (G1 x || y, G2 x.c0 || x.c1 ||
// y.c0 || y.c1), and every lane writes a status (PpotStatus).
#pragma once
#include "g2.h"
#include "params_kernels.h"

namespace zgr {

enum PpotStatus : uint32_t {
    PPOT_OK = 0,
    PPOT_COMPRESSION_MODE = 1,  // 0x80 set on an uncompressed point
    PPOT_UNEXPECTED_INFO = 2,   // infinity flag with other bits set
    PPOT_INFINITY = 3,          // the point at infinity (rejected by a checked read)
    PPOT_COORDINATE = 4,        // a coordinate >= p
    PPOT_NOT_ON_CURVE = 5,      // off the curve, or x^3 + b has no square root
    PPOT_NOT_IN_SUBGROUP = 6,   // G2 point outside the order-r subgroup
};


struct PpotConst {
    // (p + 1) / 4: a^((p+1)/4) is a square root of a when one exists (p = 3 mod 4)
    static constexpr uint32_t SQRT_E[8] = {0xb61f3f52u, 0x4f082305u, 0x5a1c72a3u, 0x65e05aa4u,
                                           0xa0605617u, 0x6e14116du, 0xb84c680au, 0x0c19139cu};
    // (p - 1) / 2: a canonical y > 0 is the larger of y, p - y iff y > (p - 1) / 2
    static constexpr uint32_t HALF[8] = {0x6c3e7ea3u, 0x9e10460bu, 0xb438e546u, 0xcbc0b548u,
                                         0x40c0ac2eu, 0xdc2822dbu, 0x7098d014u, 0x18322739u};
    // b of the twist, 3 / (9 + u), canonical
    static constexpr uint32_t B2_0[8] = {0x24a138e5u, 0x3267e6dcu, 0x59dbefa3u, 0xb5b4c5e5u,
                                         0x1be06ac3u, 0x81be1899u, 0xceb8aaaeu, 0x2b149d40u};
    static constexpr uint32_t B2_1[8] = {0x85c315d2u, 0xe4a2bd06u, 0xe52d1852u, 0xa74fa084u,
                                         0xeed8fdf4u, 0xcd2cafadu, 0x3af0fed4u, 0x009713b0u};
    // psi(x, y) = (conj(x) xi^((p-1)/3), conj(y) xi^((p-1)/2)), xi = 9 + u: the untwist-
    // Frobenius-twist endomorphism of the twist, canonical
    static constexpr uint32_t PSI_X0[8] = {0x176f553du, 0x99e39557u, 0xc2c3330cu, 0xb78cc310u,
                                           0xf559b143u, 0x4c0bec3cu, 0x4f7911f7u, 0x2fb34798u};
    static constexpr uint32_t PSI_X1[8] = {0x640fcba2u, 0x1665d51cu, 0x0b7c9dceu, 0x32ae2a1du,
                                           0xd75a0794u, 0x4ba4cc8bu, 0x61ebae20u, 0x16c9e550u};
    static constexpr uint32_t PSI_Y0[8] = {0x71a0135au, 0xdc540146u, 0xa9c95998u, 0xdbaae0edu,
                                           0xb6e2f9b9u, 0xdc5ec698u, 0x489af5dcu, 0x063cf305u};
    static constexpr uint32_t PSI_Y1[8] = {0x2623b0e3u, 0x82d37f63u, 0x8fa25bd2u, 0x21807dc9u,
                                           0xec796f2bu, 0x0704b5a7u, 0xac41049au, 0x07c03cbcu};
    static constexpr uint64_t BN_X = 0x44e992b44a6909f1ull;  // the curve parameter x of BN254
};

// a > b for canonical limbs
__device__ __forceinline__ bool ppot_gt(const uint32_t (&a)[8], const uint32_t (&b)[8]) {
    bool gt = false, decided = false;
#pragma unroll
    for (int j = 7; j >= 0; j--) {
        if (!decided && a[j] != b[j]) {
            gt = a[j] > b[j];
            decided = true;
        }
    }
    return gt;
}

// 32 big-endian bytes (eight words as loaded) -> canonical limbs; `top` masks the first
// byte (the flag bits)
__device__ __forceinline__ void ppot_be_limbs(const uint32_t* w, uint32_t top, uint32_t (&l)[8]) {
#pragma unroll
    for (int j = 0; j < 8; j++) l[j] = __builtin_bswap32(w[7 - j]);
    l[7] &= (top << 24) | 0x00ffffffu;
}

__device__ __forceinline__ bool ppot_below_p(const uint32_t (&l)[8]) {
    uint32_t pm[8];
#pragma unroll
    for (int j = 0; j < 8; j++) pm[j] = FqP::M[j];
    return ppot_gt(pm, l);
}

__device__ __forceinline__ fq ppot_mont(const uint32_t (&l)[8]) {
    fq r;
#pragma unroll
    for (int j = 0; j < 8; j++) r.v[j] = l[j];
    return f_to_mont(r);
}

__device__ __forceinline__ void ppot_canon(const fq& a, uint32_t (&l)[8]) {
    const fq c = f_from_mont(a);
#pragma unroll
    for (int j = 0; j < 8; j++) l[j] = c.v[j];
}

// a^e for a 256-bit exponent given as canonical limbs (square-and-multiply, MSB first)
__device__ __noinline__ fq ppot_pow(const fq& a, const uint32_t (&e)[8]) {
    fq r = f_one<FqP>();
    for (int j = 7; j >= 0; j--)
        for (int b = 31; b >= 0; b--) {
            r = f_sqr(r);
            if ((e[j] >> b) & 1) r = f_mul(r, a);
        }
    return r;
}

// *s = a square root of a; false when a is not a square
__device__ __forceinline__ bool ppot_sqrt(const fq& a, fq* s) {
    uint32_t e[8];
#pragma unroll
    for (int j = 0; j < 8; j++) e[j] = PpotConst::SQRT_E[j];
    *s = ppot_pow(a, e);
    return f_eq(f_sqr(*s), a);
}

// square root in Fq2 = Fq[u] / (u^2 + 1) through the norm: for a = a0 + a1 u with
// a1 != 0, x = x0 + x1 u with x0^2 = (a0 +- sqrt(a0^2 + a1^2)) / 2 and x1 = a1 / (2 x0);
// for a1 = 0, sqrt(a0) or u sqrt(-a0) (-1 is not a square).  The result is checked.
__device__ __noinline__ bool ppot_sqrt2(const fq2& a, fq2* x) {
    const fq zero = f_zero<FqP>();
    if (f_is_zero(a.c1)) {
        fq s;
        if (ppot_sqrt(a.c0, &s)) {
            *x = {s, zero};
            return true;
        }
        if (!ppot_sqrt(f_neg(a.c0), &s)) return false;
        *x = {zero, s};
        return true;
    }
    fq t;
    if (!ppot_sqrt(f_add(f_sqr(a.c0), f_sqr(a.c1)), &t)) return false;
    const fq half = f_inv(f_dbl(f_one<FqP>()));
    fq d = f_mul(f_add(a.c0, t), half), x0;
    if (!ppot_sqrt(d, &x0)) {
        d = f_mul(f_sub(a.c0, t), half);
        if (!ppot_sqrt(d, &x0)) return false;
    }
    const fq x1 = f_mul(a.c1, f_inv(f_dbl(x0)));
    *x = {x0, x1};
    const fq2 x2 = q2_sqr(*x);
    return f_eq(x2.c0, a.c0) && f_eq(x2.c1, a.c1);
}

// flags of the first byte; returns PPOT_OK or the status of an infinity encoding
template <bool COMPRESSED>
__device__ __forceinline__ uint32_t ppot_flags(const uint32_t* w, int nwords, bool* greatest) {
    // nwords: the encoding's length in words
    const uint32_t b0 = w[0] & 0xffu;  // first byte of the encoding
    *greatest = false;
    if (!COMPRESSED && (b0 & 0x80u)) return PPOT_COMPRESSION_MODE;
    if (b0 & 0x40u) {
        uint32_t rest = w[0] & ~0xc0u;  // an infinity encoding clears both flag bits first
        for (int k = 1; k < nwords; k++) rest |= w[k];
        return rest ? PPOT_UNEXPECTED_INFO : PPOT_INFINITY;
    }
    *greatest = (b0 & 0x80u) != 0;
    return PPOT_OK;
}

// y or -y, whichever is the larger (greatest) or smaller canonical integer
__device__ __forceinline__ fq ppot_pick_y(const fq& y, bool greatest) {
    uint32_t c[8], h[8];
    ppot_canon(y, c);
#pragma unroll
    for (int j = 0; j < 8; j++) h[j] = PpotConst::HALF[j];
    return ppot_gt(c, h) == greatest ? y : f_neg(y);
}

__device__ __forceinline__ fq2 ppot_pick_y2(const fq2& y, bool greatest) {
    uint32_t c1[8], c0[8], h[8];
    ppot_canon(y.c1, c1);
    ppot_canon(y.c0, c0);
#pragma unroll
    for (int j = 0; j < 8; j++) h[j] = PpotConst::HALF[j];
    const bool c1_zero = f_is_zero(y.c1);
    const bool larger = c1_zero ? ppot_gt(c0, h) : ppot_gt(c1, h);
    return larger == greatest ? y : q2_neg(y);
}

// psi on XYZZ coordinates: conjugation is a field automorphism, so it maps X / ZZ and
// Y / ZZZ coordinate-wise
__device__ __forceinline__ g2_xyzz ppot_psi(const g2_xyzz& p, const fq2& cx, const fq2& cy) {
    auto conj = [](const fq2& a) { return fq2{a.c0, f_neg(a.c1)}; };
    return {q2_mul(conj(p.X), cx), q2_mul(conj(p.Y), cy), conj(p.ZZ), conj(p.ZZZ)};
}

__device__ __forceinline__ bool ppot_g2_eq(const g2_xyzz& a, const g2_xyzz& b) {
    const bool ia = g2_is_inf(a), ib = g2_is_inf(b);
    if (ia || ib) return ia == ib;
    const fq2 x1 = q2_mul(a.X, b.ZZ), x2 = q2_mul(b.X, a.ZZ);
    const fq2 y1 = q2_mul(a.Y, b.ZZZ), y2 = q2_mul(b.Y, a.ZZZ);
    return f_eq(x1.c0, x2.c0) && f_eq(x1.c1, x2.c1) && f_eq(y1.c0, y2.c0) && f_eq(y1.c1, y2.c1);
}

// Q in the order-r subgroup of the twist iff [x+1]Q + psi([x]Q) + psi^2([x]Q) =
// psi^3([2x]Q) (El Housni, Guillevic and Piellard, "Co-factor clearing and subgroup
// membership testing on pairing-friendly curves", 2022, for BN curves): a 63-bit scalar
// multiplication instead of [r]Q
__device__ __noinline__ bool ppot_g2_in_subgroup(const g2_aff& q) {
    const fq2 cx = {ppot_mont(PpotConst::PSI_X0), ppot_mont(PpotConst::PSI_X1)};
    const fq2 cy = {ppot_mont(PpotConst::PSI_Y0), ppot_mont(PpotConst::PSI_Y1)};
    g2_xyzz xq = g2_inf();
    for (int b = 62; b >= 0; b--) {
        xq = g2_dbl(xq);
        if ((PpotConst::BN_X >> b) & 1) xq = g2_madd(xq, q);
    }
    const g2_xyzz p1 = ppot_psi(xq, cx, cy), p2 = ppot_psi(p1, cx, cy);
    const g2_xyzz lhs = g2_add(g2_add(g2_madd(xq, q), p1), p2);
    const g2_xyzz rhs = ppot_psi(ppot_psi(ppot_psi(g2_dbl(xq), cx, cy), cx, cy), cx, cy);
    return ppot_g2_eq(lhs, rhs);
}

// G1: y^2 = x^3 + 3
template <bool COMPRESSED>
__device__ __forceinline__ void ppot_g1(const uint32_t* src, size_t count, uint32_t* out, uint32_t* status) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    constexpr int NW = COMPRESSED ? 8 : 16;
    const uint32_t* w = src + i * NW;
    bool greatest;
    uint32_t st = ppot_flags<COMPRESSED>(w, NW, &greatest);
    uint32_t lx[8], ly[8];
    fq x = f_zero<FqP>(), y = f_zero<FqP>();
    if (st == PPOT_OK) {
        ppot_be_limbs(w, 0x3fu, lx);
        if (!COMPRESSED) ppot_be_limbs(w + 8, 0xffu, ly);
        if (!ppot_below_p(lx) || (!COMPRESSED && !ppot_below_p(ly))) st = PPOT_COORDINATE;
    }
    if (st == PPOT_OK) {
        x = ppot_mont(lx);
        fq three = f_zero<FqP>();
        three.v[0] = 3;
        const fq rhs = f_add(f_mul(f_sqr(x), x), f_to_mont(three));
        if (COMPRESSED) {
            if (!ppot_sqrt(rhs, &y)) st = PPOT_NOT_ON_CURVE;
            else y = ppot_pick_y(y, greatest);
        } else {
            y = ppot_mont(ly);
            if (!f_eq(f_sqr(y), rhs)) st = PPOT_NOT_ON_CURVE;
        }
    }
    if (st != PPOT_OK) x = y = f_zero<FqP>();
    ppot_canon(x, lx);
    ppot_canon(y, ly);
    uint32_t* o = out + i * 16;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        o[j] = lx[j];
        o[8 + j] = ly[j];
    }
    status[i] = st;
}

// G2: y^2 = x^3 + 3 / (9 + u) over Fq2, then the subgroup test
template <bool COMPRESSED>
__device__ __forceinline__ void ppot_g2(const uint32_t* src, size_t count, uint32_t* out, uint32_t* status) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    constexpr int NW = COMPRESSED ? 16 : 32, NC = COMPRESSED ? 2 : 4;
    const uint32_t* w = src + i * NW;
    bool greatest;
    uint32_t st = ppot_flags<COMPRESSED>(w, NW, &greatest);
    uint32_t l[4][8];  // x.c1, x.c0, y.c1, y.c0 in file order
    if (st == PPOT_OK) {
#pragma unroll
        for (int k = 0; k < NC; k++) {
            ppot_be_limbs(w + 8 * k, k == 0 ? 0x3fu : 0xffu, l[k]);
            if (!ppot_below_p(l[k])) st = PPOT_COORDINATE;
        }
    }
    g2_aff P = {q2_zero(), q2_zero()};
    if (st == PPOT_OK) {
        P.x = {ppot_mont(l[1]), ppot_mont(l[0])};
        const fq2 b2 = {ppot_mont(PpotConst::B2_0), ppot_mont(PpotConst::B2_1)};
        const fq2 rhs = q2_add(q2_mul(q2_sqr(P.x), P.x), b2);
        if (COMPRESSED) {
            if (!ppot_sqrt2(rhs, &P.y)) st = PPOT_NOT_ON_CURVE;
            else P.y = ppot_pick_y2(P.y, greatest);
        } else {
            P.y = {ppot_mont(l[3]), ppot_mont(l[2])};
            const fq2 y2 = q2_sqr(P.y);
            if (!f_eq(y2.c0, rhs.c0) || !f_eq(y2.c1, rhs.c1)) st = PPOT_NOT_ON_CURVE;
        }
    }
    if (st == PPOT_OK && !ppot_g2_in_subgroup(P)) st = PPOT_NOT_IN_SUBGROUP;
    if (st != PPOT_OK) P = {q2_zero(), q2_zero()};
    uint32_t* o = out + i * 32;
    ppot_canon(P.x.c0, l[0]);
    ppot_canon(P.x.c1, l[1]);
    ppot_canon(P.y.c0, l[2]);
    ppot_canon(P.y.c1, l[3]);
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int j = 0; j < 8; j++) o[8 * k + j] = l[k][j];
    status[i] = st;
}

// one kernel per encoding; src holds count encodings back to back, out count points
__global__ __launch_bounds__(PPOT_G1_THREADS) void k_ppot_g1u(const uint32_t* src, size_t count, uint32_t* out,
                                                              uint32_t* status) {
    ppot_g1<false>(src, count, out, status);
}
__global__ __launch_bounds__(PPOT_G1_THREADS) void k_ppot_g1c(const uint32_t* src, size_t count, uint32_t* out,
                                                              uint32_t* status) {
    ppot_g1<true>(src, count, out, status);
}
__global__ __launch_bounds__(PPOT_G2_THREADS) void k_ppot_g2u(const uint32_t* src, size_t count, uint32_t* out,
                                                              uint32_t* status) {
    ppot_g2<false>(src, count, out, status);
}
__global__ __launch_bounds__(PPOT_G2_THREADS) void k_ppot_g2c(const uint32_t* src, size_t count, uint32_t* out,
                                                              uint32_t* status) {
    ppot_g2<true>(src, count, out, status);
}

}  // namespace zgr
