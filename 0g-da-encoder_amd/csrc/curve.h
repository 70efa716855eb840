// BN254 G1 (y^2 = x^3 + 3 over Fq) on gfx950.
//
// Accumulators use XYZZ coordinates (X, Y, ZZ, ZZZ) <-> affine (X/ZZ, Y/ZZZ), with
// ZZ == 0 meaning the point at infinity.  Formulas (EFD, "xyzz" short Weierstrass,
// a = 0):  madd-2008-s (8M + 2S), add-2008-s (12M + 2S), dbl-2008-s-1 (6M + 3S),
// mdbl-2008-s-1.  The exceptional cases (P == Q, P == -Q, infinity operands) are
// handled explicitly so every sum is exact.
#pragma once
#include "field.h"

namespace zgr {

struct g1_aff {  // Montgomery coordinates; never infinity inside param buffers
    fq x, y;
};

struct g1_xyzz {
    fq X, Y, ZZ, ZZZ;
};

__device__ __forceinline__ g1_xyzz xyzz_inf() {
    g1_xyzz r;
    r.X = f_zero<FqP>();
    r.Y = f_zero<FqP>();
    r.ZZ = f_zero<FqP>();
    r.ZZZ = f_zero<FqP>();
    return r;
}

__device__ __forceinline__ bool xyzz_is_inf(const g1_xyzz& p) { return f_is_zero(p.ZZ); }

__device__ __forceinline__ g1_xyzz xyzz_from_aff(const g1_aff& a) {
    g1_xyzz r;
    r.X = a.x;
    r.Y = a.y;
    r.ZZ = f_one<FqP>();
    r.ZZZ = f_one<FqP>();
    return r;
}

__device__ __forceinline__ g1_xyzz xyzz_select(bool c, const g1_xyzz& a, const g1_xyzz& b) {
    g1_xyzz r;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        r.X.v[i] = c ? a.X.v[i] : b.X.v[i];
        r.Y.v[i] = c ? a.Y.v[i] : b.Y.v[i];
        r.ZZ.v[i] = c ? a.ZZ.v[i] : b.ZZ.v[i];
        r.ZZZ.v[i] = c ? a.ZZZ.v[i] : b.ZZZ.v[i];
    }
    return r;
}

__device__ __forceinline__ g1_aff aff_neg(const g1_aff& a) {
    g1_aff r;
    r.x = a.x;
    r.y = f_neg(a.y);
    return r;
}

// dbl-2008-s-1, a = 0
__device__ __forceinline__ g1_xyzz xyzz_dbl(const g1_xyzz& p) {
    fq U = f_dbl(p.Y);
    fq V = f_sqr(U);
    fq W = f_mul(U, V);
    fq S = f_mul(p.X, V);
    fq X2 = f_sqr(p.X);
    fq M = f_add(f_dbl(X2), X2);
    fq X3 = f_sub(f_sqr(M), f_dbl(S));
    fq Y3 = f_sub(f_mul(M, f_sub(S, X3)), f_mul(W, p.Y));
    g1_xyzz r;
    r.X = X3;
    r.Y = Y3;
    r.ZZ = f_mul(V, p.ZZ);
    r.ZZZ = f_mul(W, p.ZZZ);
    // infinity doubles to infinity: ZZ stays 0
    return r;
}

// add-2008-s: p + q
__device__ __forceinline__ g1_xyzz xyzz_add(const g1_xyzz& p, const g1_xyzz& q) {
    fq U1 = f_mul(p.X, q.ZZ);
    fq U2 = f_mul(q.X, p.ZZ);
    fq S1 = f_mul(p.Y, q.ZZZ);
    fq S2 = f_mul(q.Y, p.ZZZ);
    fq P = f_sub(U2, U1);
    fq R = f_sub(S2, S1);
    bool pinf = xyzz_is_inf(p), qinf = xyzz_is_inf(q);
    if (!pinf && !qinf && f_is_zero(P) && f_is_zero(R)) return xyzz_dbl(p);
    fq PP = f_sqr(P);
    fq PPP = f_mul(P, PP);
    fq Q = f_mul(U1, PP);
    fq X3 = f_sub(f_sub(f_sqr(R), PPP), f_dbl(Q));
    fq Y3 = f_sub(f_mul(R, f_sub(Q, X3)), f_mul(S1, PPP));
    g1_xyzz r;
    r.X = X3;
    r.Y = Y3;
    r.ZZ = f_mul(f_mul(p.ZZ, q.ZZ), PP);     // 0 when P == 0 (p == -q): infinity
    r.ZZZ = f_mul(f_mul(p.ZZZ, q.ZZZ), PPP);
    r = xyzz_select(pinf, q, r);
    r = xyzz_select(qinf, p, r);
    return r;
}

// mdbl-2008-s-1: double an affine point
__device__ __forceinline__ g1_xyzz aff_dbl(const g1_aff& a) {
    fq U = f_dbl(a.y);
    fq V = f_sqr(U);
    fq W = f_mul(U, V);
    fq S = f_mul(a.x, V);
    fq X2 = f_sqr(a.x);
    fq M = f_add(f_dbl(X2), X2);
    g1_xyzz r;
    r.X = f_sub(f_sqr(M), f_dbl(S));
    r.Y = f_sub(f_mul(M, f_sub(S, r.X)), f_mul(W, a.y));
    r.ZZ = V;
    r.ZZZ = W;
    return r;
}

// madd-2008-s: p + affine q
__device__ __forceinline__ g1_xyzz xyzz_madd(const g1_xyzz& p, const g1_aff& q) {
    fq U2 = f_mul(q.x, p.ZZ);
    fq S2 = f_mul(q.y, p.ZZZ);
    fq P = f_sub(U2, p.X);
    fq R = f_sub(S2, p.Y);
    bool pinf = xyzz_is_inf(p);
    if (!pinf && f_is_zero(P) && f_is_zero(R)) return aff_dbl(q);
    fq PP = f_sqr(P);
    fq PPP = f_mul(P, PP);
    fq Q = f_mul(p.X, PP);
    g1_xyzz r;
    r.X = f_sub(f_sub(f_sqr(R), PPP), f_dbl(Q));
    r.Y = f_sub(f_mul(R, f_sub(Q, r.X)), f_mul(p.Y, PPP));
    r.ZZ = f_mul(p.ZZ, PP);
    r.ZZZ = f_mul(p.ZZZ, PPP);
    return xyzz_select(pinf, xyzz_from_aff(q), r);
}

// k * p for a small unsigned k (MSB-first double-and-add)
__device__ __forceinline__ g1_xyzz xyzz_mul_small(const g1_xyzz& p, uint32_t k, int bits) {
    g1_xyzz r = xyzz_inf();
    for (int b = bits - 1; b >= 0; b--) {
        r = xyzz_dbl(r);
        if ((k >> b) & 1) r = xyzz_add(r, p);
    }
    return r;
}

// affine (Montgomery) of an XYZZ point; *inf set when p is infinity
__device__ __forceinline__ g1_aff xyzz_to_aff(const g1_xyzz& p, bool* inf) {
    *inf = xyzz_is_inf(p);
    fq I = f_inv(f_mul(p.ZZ, p.ZZZ));  // 1 / z^5 where ZZ = z^2, ZZZ = z^3
    g1_aff a;
    a.x = f_mul(p.X, f_mul(I, p.ZZZ));  // X / ZZ
    a.y = f_mul(p.Y, f_mul(I, p.ZZ));   // Y / ZZZ
    return a;
}

__device__ __forceinline__ g1_aff aff_load(const uint32_t* p) {
    g1_aff a;
    a.x = f_load<FqP>(p);
    a.y = f_load<FqP>(p + 8);
    return a;
}

__device__ __forceinline__ void aff_store(uint32_t* p, const g1_aff& a) {
    f_store(p, a.x);
    f_store(p + 8, a.y);
}

__device__ __forceinline__ g1_xyzz xyzz_load(const uint32_t* p) {
    g1_xyzz r;
    r.X = f_load<FqP>(p);
    r.Y = f_load<FqP>(p + 8);
    r.ZZ = f_load<FqP>(p + 16);
    r.ZZZ = f_load<FqP>(p + 24);
    return r;
}

__device__ __forceinline__ void xyzz_store(uint32_t* p, const g1_xyzz& a) {
    f_store(p, a.X);
    f_store(p + 8, a.Y);
    f_store(p + 16, a.ZZ);
    f_store(p + 24, a.ZZZ);
}

// canonical output point: x || y little-endian, all zeros for infinity
__device__ __forceinline__ void out_store_canon(uint32_t* p, const g1_xyzz& q) {
    bool inf;
    g1_aff a = xyzz_to_aff(q, &inf);
    fq x = f_from_mont(a.x), y = f_from_mont(a.y);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        p[i] = inf ? 0u : x.v[i];
        p[8 + i] = inf ? 0u : y.v[i];
    }
}

}  // namespace zgr
