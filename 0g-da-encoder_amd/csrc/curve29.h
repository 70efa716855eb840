// G1 XYZZ arithmetic over the reduced-radix base field (field29.h) for the MSM
// bucket loop.  Same formulas and exceptional-case handling as curve.h
// (add-2008-s, dbl-2008-s-1, mdbl-2008-s-1); ZZ == 0 (mod p) means infinity.
#pragma once
#include "curve.h"
#include "field29.h"

namespace zgr {

struct a29 {  // affine
    fq29 x, y;
};

struct x29 {  // XYZZ
    fq29 X, Y, ZZ, ZZZ;
};

__device__ __forceinline__ x29 x29_inf() {
    x29 r;
    r.X = q_zero();
    r.Y = q_zero();
    r.ZZ = q_zero();
    r.ZZZ = q_zero();
    return r;
}

__device__ __forceinline__ bool x29_is_inf(const x29& p) { return q_is_zero(p.ZZ); }

__device__ __forceinline__ x29 x29_from_aff(const a29& a) {
    x29 r;
    r.X = a.x;
    r.Y = a.y;
    r.ZZ = q_one();
    r.ZZZ = q_one();
    return r;
}

__device__ __forceinline__ x29 x29_select(bool c, const x29& a, const x29& b) {
    x29 r;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        r.X.l[k] = c ? a.X.l[k] : b.X.l[k];
        r.Y.l[k] = c ? a.Y.l[k] : b.Y.l[k];
        r.ZZ.l[k] = c ? a.ZZ.l[k] : b.ZZ.l[k];
        r.ZZZ.l[k] = c ? a.ZZZ.l[k] : b.ZZZ.l[k];
    }
    return r;
}

__device__ __forceinline__ x29 x29_dbl(const x29& p) {
    fq29 U = q_dbl(p.Y);
    fq29 V = q_sqr(U);
    fq29 W = q_mul(U, V);
    fq29 S = q_mul(p.X, V);
    fq29 X2 = q_sqr(p.X);
    fq29 M = q_add(q_dbl(X2), X2);
    x29 r;
    r.X = q_sub(q_sqr(M), q_dbl(S));
    r.Y = q_mul2(M, q_sub_lazy(S, r.X), W, q_sub_lazy(q_zero(), p.Y));  // M (S - X3) - W Y
    r.ZZ = q_mul(V, p.ZZ);
    r.ZZZ = q_mul(W, p.ZZZ);
    return r;
}

// p + q.  Inputs may carry X unreduced in [0, 8p) (X only feeds multiplications here and
// in x29_dbl).  LAZY_X: the result's X is left in [0, 8p) as well (one carry pass for
// X3 = R^2 - PPP - 2Q instead of three reduced subtractions); x29_norm restores it.
template <bool LAZY_X = false>
__device__ __forceinline__ x29 x29_add_t(const x29& p, const x29& q) {
    const bool pinf = x29_is_inf(p), qinf = x29_is_inf(q);
    fq29 U1 = q_mul(p.X, q.ZZ);
    fq29 P = q_sub_lazy(q_mul(q.X, p.ZZ), U1);  // < 4p, feeds multiplications only
    fq29 S1 = q_mul(p.Y, q.ZZZ);
    fq29 R = q_sub_lazy(q_mul(q.Y, p.ZZZ), S1);
    // equal x: p == q (double) or p == -q (infinity); one test on the common path
    if (!pinf && !qinf && q_is_zero4(P)) return q_is_zero4(R) ? x29_dbl(p) : x29_inf();
    x29 r;
    // Z coordinates first so the inputs' ZZ / ZZZ die early
    fq29 PP = q_sqr(P);
    r.ZZ = q_mul(q_mul(p.ZZ, q.ZZ), PP);
    fq29 PPP = q_mul(P, PP);
    r.ZZZ = q_mul(q_mul(p.ZZZ, q.ZZZ), PPP);
    fq29 Q = q_mul(U1, PP);
    if (LAZY_X) {
        r.X = q_x3_lazy(q_sqr(R), PPP, Q);
        r.Y = q_mul2(R, q_sub_lazy8(Q, r.X), q_sub_lazy(q_zero(), S1), PPP);  // R (Q - X3) - S1 PPP
    } else {
        r.X = q_sub(q_sub(q_sqr(R), PPP), q_dbl(Q));
        r.Y = q_mul2(R, q_sub_lazy(Q, r.X), q_sub_lazy(q_zero(), S1), PPP);
    }
    r = x29_select(pinf, q, r);
    r = x29_select(qinf, p, r);
    return r;
}

__device__ __forceinline__ x29 x29_add(const x29& p, const x29& q) { return x29_add_t<false>(p, q); }
__device__ __forceinline__ x29 x29_add_lazy(const x29& p, const x29& q) { return x29_add_t<true>(p, q); }

// p + q for affine q (madd-2008-s, 8M + 2S) with every exceptional case handled; q is
// never infinity.  The bucket loop runs x29_madd_unchecked and comes here only for the
// lanes that one flags.
__device__ __forceinline__ x29 x29_madd(const x29& p, const a29& q) {
    const bool pinf = x29_is_inf(p);
    fq29 P = q_sub_lazy(q_mul(q.x, p.ZZ), p.X);
    fq29 R = q_sub_lazy(q_mul(q.y, p.ZZZ), p.Y);
    // equal x: p == q (double) or p == -q (infinity)
    if (!pinf && q_is_zero4(P)) return q_is_zero4(R) ? x29_dbl(p) : x29_inf();
    x29 r;
    fq29 PP = q_sqr(P);
    r.ZZ = q_mul(p.ZZ, PP);
    fq29 PPP = q_mul(P, PP);
    r.ZZZ = q_mul(p.ZZZ, PPP);
    fq29 Q = q_mul(p.X, PP);
    r.X = q_sub(q_sub(q_sqr(R), PPP), q_dbl(Q));
    r.Y = q_mul2(R, q_sub_lazy(Q, r.X), q_sub_lazy(q_zero(), p.Y), PPP);  // R (Q - X3) - Y1 PPP
    return x29_select(pinf, x29_from_aff(q), r);
}

// p + q for affine q by the generic formula alone (madd-2008-s without its exceptional
// cases), for the bucket loop, on lazy-X points: X in [0, 8p), Y, ZZ, ZZZ in [0, 2p), as
// this function returns them (X only ever feeds multiplications and q_sub_lazy8, which
// accept it; x29_norm brings a point back to the usual form).  *rare is set on every
// lane where x(p) may equal x(q) or p may be infinity: H = x(q) ZZ - X + 8p lies in
// (0, 10p) and is 0 mod p only when it is one of 0, p, .., 9p, which its low limb
// screens; the all-zero infinity (x29_inf) gives H = 8p exactly.  Those lanes (and, at
// random, about one in 2^26 others) must redo the addition exactly (x29_madd on
// x29_norm(p)); the value returned there is meaningless.
__device__ __forceinline__ x29 x29_madd_unchecked(const x29& p, const a29& q, bool* rare) {
    fq29 H = q_sub_lazy8(q_mul(q.x, p.ZZ), p.X);
    fq29 R = q_sub_lazy(q_mul(q.y, p.ZZZ), p.Y);
    // H mod 2^29 == (k p) mod 2^29 for some k in 0..9, in one product: k = H.l[0] / p_0
    *rare = ((H.l[0] * Q29::P0_INV) & M29) <= 9u;
    x29 r;
    fq29 PP = q_sqr(H);  // H^2 < 100 p^2: below 2p (q_mul's bound)
    r.ZZ = q_mul(p.ZZ, PP);
    fq29 PPP = q_mul(H, PP);
    r.ZZZ = q_mul(p.ZZZ, PPP);
    fq29 Q = q_mul(p.X, PP);
    r.X = q_x3_lazy(q_sqr(R), PPP, Q);
    // R (Q - X3) - Y1 PPP: products below 4p * 10p + 2p * 2p = 44 p^2
    r.Y = q_mul2(R, q_sub_lazy8(Q, r.X), q_sub_lazy(q_zero(), p.Y), PPP);
    return r;
}

// a lazy-X point (x29_madd_unchecked) in the usual form: every coordinate in [0, 2p)
__device__ __forceinline__ x29 x29_norm(const x29& p) {
    x29 r = p;
    r.X = q_reduce_8p(p.X);
    return r;
}

// memory layout of a stored XYZZ point: X, Y, ZZ, ZZZ, 9 limbs each (36 u32)
constexpr int X29_WORDS = 36;

__device__ __forceinline__ void x29_store(uint32_t* d, const x29& v) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
        d[i] = v.X.l[i];
        d[9 + i] = v.Y.l[i];
        d[18 + i] = v.ZZ.l[i];
        d[27 + i] = v.ZZZ.l[i];
    }
}

__device__ __forceinline__ x29 x29_load(const uint32_t* d) {
    x29 v;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        v.X.l[i] = d[i];
        v.Y.l[i] = d[9 + i];
        v.ZZ.l[i] = d[18 + i];
        v.ZZZ.l[i] = d[27 + i];
    }
    return v;
}

// 2^256-Montgomery XYZZ <-> reduced-radix XYZZ
__device__ __forceinline__ x29 x29_from_xyzz(const g1_xyzz& p) {
    x29 r;
    r.X = q_from_fq(p.X);
    r.Y = q_from_fq(p.Y);
    r.ZZ = q_from_fq(p.ZZ);
    r.ZZZ = q_from_fq(p.ZZZ);
    return r;
}

// Affine coordinates of p in the 29-bit form, from one inversion of ZZ ZZZ = z^5
// (x = X ZZZ / z^5, y = Y ZZ / z^5); *inf reports infinity (the coordinates are then 0).
__device__ __forceinline__ void x29_to_aff(const x29& p, fq29& x, fq29& y, bool* inf) {
    *inf = x29_is_inf(p);
    const fq29 I = q_inv(q_mul(p.ZZ, p.ZZZ));
    x = q_mul(p.X, q_mul(I, p.ZZZ));
    y = q_mul(p.Y, q_mul(I, p.ZZ));
}

// canonical output point: x || y as 8 little-endian u32 each, all zeros for infinity
__device__ __forceinline__ void x29_store_canon(uint32_t* out, const x29& p) {
    bool inf;
    fq29 x, y;
    x29_to_aff(p, x, y, &inf);
    fq29 one = q_zero();
    one.l[0] = 1;
    uint32_t wx[8], wy[8];
    q_pack(q_canon(q_mul(x, one)), wx);  // x 2^261 / 2^261, reduced to [0, p)
    q_pack(q_canon(q_mul(y, one)), wy);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        out[i] = inf ? 0u : wx[i];
        out[8 + i] = inf ? 0u : wy[i];
    }
}

// k * p for k < 2^bits, uniform per lane
__device__ __forceinline__ x29 x29_mul_small_uniform(const x29& p, uint32_t k, int bits) {
    x29 r = x29_inf();
    for (int b = bits - 1; b >= 0; b--) {
        r = x29_dbl(r);
        x29 s = x29_add(r, p);
        r = x29_select((k >> b) & 1, s, r);
    }
    return r;
}

}  // namespace zgr
