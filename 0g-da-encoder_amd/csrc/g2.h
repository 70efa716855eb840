// BN254 G2 (the twist y^2 = x^3 + 3 / (9 + u) over Fq2 = Fq[u] / (u^2 + 1)) for the
// verifier side of the public parameters: the powers [tau^i] G2 of PowerTau
// (crates/amt/src/power_tau.rs:54-80) and the vanishes of AMTParams (gen_vanishes,
// crates/amt/src/prove_params/generate.rs:234-263).  This is parameter generation, not
// the encode loop, so it uses the plain 8-limb Montgomery Fq of field.h; points are
// XYZZ with the same a = 0 formulas as curve.h (EFD madd-2008-s, add-2008-s,
// dbl-2008-s-1), over Fq2.
#pragma once
#include "field.h"

namespace zgr {

struct fq2 {
    fq c0, c1;  // c0 + c1 u, Montgomery
};

__device__ __forceinline__ fq2 q2_zero() { return {f_zero<FqP>(), f_zero<FqP>()}; }
__device__ __forceinline__ fq2 q2_one() { return {f_one<FqP>(), f_zero<FqP>()}; }
__device__ __forceinline__ bool q2_is_zero(const fq2& a) { return f_is_zero(a.c0) && f_is_zero(a.c1); }
__device__ __forceinline__ fq2 q2_add(const fq2& a, const fq2& b) { return {f_add(a.c0, b.c0), f_add(a.c1, b.c1)}; }
__device__ __forceinline__ fq2 q2_sub(const fq2& a, const fq2& b) { return {f_sub(a.c0, b.c0), f_sub(a.c1, b.c1)}; }
__device__ __forceinline__ fq2 q2_dbl(const fq2& a) { return q2_add(a, a); }
__device__ __forceinline__ fq2 q2_neg(const fq2& a) { return {f_neg(a.c0), f_neg(a.c1)}; }

// (a0 + a1 u)(b0 + b1 u) = a0 b0 - a1 b1 + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) u
__device__ __forceinline__ fq2 q2_mul(const fq2& a, const fq2& b) {
    const fq v0 = f_mul(a.c0, b.c0), v1 = f_mul(a.c1, b.c1);
    const fq s = f_mul(f_add(a.c0, a.c1), f_add(b.c0, b.c1));
    return {f_sub(v0, v1), f_sub(f_sub(s, v0), v1)};
}

// (a0 + a1 u)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u
__device__ __forceinline__ fq2 q2_sqr(const fq2& a) {
    const fq m = f_mul(a.c0, a.c1);
    return {f_mul(f_add(a.c0, a.c1), f_sub(a.c0, a.c1)), f_dbl(m)};
}

// 1 / (a0 + a1 u) = (a0 - a1 u) / (a0^2 + a1^2); 0 maps to 0
__device__ __forceinline__ fq2 q2_inv(const fq2& a) {
    const fq t = f_inv(f_add(f_sqr(a.c0), f_sqr(a.c1)));
    return {f_mul(a.c0, t), f_neg(f_mul(a.c1, t))};
}

struct g2_aff {
    fq2 x, y;
};

struct g2_xyzz {
    fq2 X, Y, ZZ, ZZZ;  // affine (X / ZZ, Y / ZZZ); ZZ == 0 is the point at infinity
};

__device__ __forceinline__ g2_xyzz g2_inf() { return {q2_zero(), q2_zero(), q2_zero(), q2_zero()}; }
__device__ __forceinline__ bool g2_is_inf(const g2_xyzz& p) { return q2_is_zero(p.ZZ); }
__device__ __forceinline__ g2_xyzz g2_from_aff(const g2_aff& a) { return {a.x, a.y, q2_one(), q2_one()}; }

__device__ __forceinline__ g2_xyzz g2_select(bool c, const g2_xyzz& a, const g2_xyzz& b) { return c ? a : b; }

// dbl-2008-s-1 (a = 0); infinity doubles to infinity
__device__ __forceinline__ g2_xyzz g2_dbl(const g2_xyzz& p) {
    const fq2 U = q2_dbl(p.Y), V = q2_sqr(U), W = q2_mul(U, V), S = q2_mul(p.X, V), X2 = q2_sqr(p.X);
    const fq2 M = q2_add(q2_dbl(X2), X2);
    const fq2 X3 = q2_sub(q2_sqr(M), q2_dbl(S));
    return {X3, q2_sub(q2_mul(M, q2_sub(S, X3)), q2_mul(W, p.Y)), q2_mul(V, p.ZZ), q2_mul(W, p.ZZZ)};
}

// mdbl-2008-s-1: double an affine point
__device__ __forceinline__ g2_xyzz g2_aff_dbl(const g2_aff& a) {
    const fq2 U = q2_dbl(a.y), V = q2_sqr(U), W = q2_mul(U, V), S = q2_mul(a.x, V), X2 = q2_sqr(a.x);
    const fq2 M = q2_add(q2_dbl(X2), X2);
    const fq2 X3 = q2_sub(q2_sqr(M), q2_dbl(S));
    return {X3, q2_sub(q2_mul(M, q2_sub(S, X3)), q2_mul(W, a.y)), V, W};
}

// add-2008-s, exact for every pair (equal points double, opposite points give infinity)
__device__ __forceinline__ g2_xyzz g2_add(const g2_xyzz& p, const g2_xyzz& q) {
    if (g2_is_inf(p)) return q;
    if (g2_is_inf(q)) return p;
    const fq2 U1 = q2_mul(p.X, q.ZZ), U2 = q2_mul(q.X, p.ZZ), S1 = q2_mul(p.Y, q.ZZZ), S2 = q2_mul(q.Y, p.ZZZ);
    const fq2 P = q2_sub(U2, U1), R = q2_sub(S2, S1);
    if (q2_is_zero(P) && q2_is_zero(R)) return g2_dbl(p);
    const fq2 PP = q2_sqr(P), PPP = q2_mul(P, PP), Q = q2_mul(U1, PP);
    const fq2 X3 = q2_sub(q2_sub(q2_sqr(R), PPP), q2_dbl(Q));
    return {X3, q2_sub(q2_mul(R, q2_sub(Q, X3)), q2_mul(S1, PPP)), q2_mul(q2_mul(p.ZZ, q.ZZ), PP),
            q2_mul(q2_mul(p.ZZZ, q.ZZZ), PPP)};
}

// madd-2008-s: p + affine q, exact
__device__ __forceinline__ g2_xyzz g2_madd(const g2_xyzz& p, const g2_aff& q) {
    if (g2_is_inf(p)) return g2_from_aff(q);
    const fq2 U2 = q2_mul(q.x, p.ZZ), S2 = q2_mul(q.y, p.ZZZ);
    const fq2 P = q2_sub(U2, p.X), R = q2_sub(S2, p.Y);
    if (q2_is_zero(P) && q2_is_zero(R)) return g2_aff_dbl(q);
    const fq2 PP = q2_sqr(P), PPP = q2_mul(P, PP), Q = q2_mul(p.X, PP);
    const fq2 X3 = q2_sub(q2_sub(q2_sqr(R), PPP), q2_dbl(Q));
    return {X3, q2_sub(q2_mul(R, q2_sub(Q, X3)), q2_mul(p.Y, PPP)), q2_mul(p.ZZ, PP), q2_mul(p.ZZZ, PPP)};
}

// k p for a canonical scalar k (MSB-first double-and-add)
__device__ __forceinline__ g2_xyzz g2_mul(const g2_xyzz& p, const fr& k) {
    g2_xyzz r = g2_inf();
    for (int w = 7; w >= 0; w--)
        for (int b = 31; b >= 0; b--) {
            r = g2_dbl(r);
            if ((k.v[w] >> b) & 1) r = g2_add(r, p);
        }
    return r;
}

// affine (Montgomery); *inf set for the point at infinity (then x = y = 0)
__device__ __forceinline__ g2_aff g2_to_aff(const g2_xyzz& p, bool* inf) {
    *inf = g2_is_inf(p);
    const fq2 I = q2_inv(q2_mul(p.ZZ, p.ZZZ));  // 1 / z^5
    g2_aff a;
    a.x = q2_mul(p.X, q2_mul(I, p.ZZZ));
    a.y = q2_mul(p.Y, q2_mul(I, p.ZZ));
    return a;
}

// memory form: x.c0 || x.c1 || y.c0 || y.c1, 8 u32 each (the bamt G2 layout when Montgomery,
// fast_serde_bn254.rs:150-160)
__device__ __forceinline__ g2_aff g2_load(const uint32_t* p) {
    return {{f_load<FqP>(p), f_load<FqP>(p + 8)}, {f_load<FqP>(p + 16), f_load<FqP>(p + 24)}};
}

__device__ __forceinline__ void g2_store(uint32_t* p, const g2_aff& a) {
    f_store(p, a.x.c0);
    f_store(p + 8, a.x.c1);
    f_store(p + 16, a.y.c0);
    f_store(p + 24, a.y.c1);
}

// the generator of ark-bn254 (G2Affine::generator(); the EIP-197 P2), canonical limbs
struct G2Gen {
    static constexpr uint32_t X0[8] = {0xd992f6edu, 0x46debd5cu, 0xf75edaddu, 0x674322d4u,
                                       0x5e5c4479u, 0x426a0066u, 0x121f1e76u, 0x1800deefu};
    static constexpr uint32_t X1[8] = {0xaef312c2u, 0x97e485b7u, 0x35a9e712u, 0xf1aa4933u,
                                       0x31fb5d25u, 0x7260bfb7u, 0x920d483au, 0x198e9393u};
    static constexpr uint32_t Y0[8] = {0x66fa7daau, 0x4ce6cc01u, 0x0c43d37bu, 0xe3d1e769u,
                                       0x8dcb408fu, 0x4aab7180u, 0xdb8c6debu, 0x12c85ea5u};
    static constexpr uint32_t Y1[8] = {0xd122975bu, 0x55acdadcu, 0x70b38ef3u, 0xbc4b3133u,
                                       0x690c3395u, 0xec9e99adu, 0x585ff075u, 0x090689d0u};
};

__device__ __forceinline__ fq f_from_words(const uint32_t (&w)[8]) {
    fq r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = w[i];
    return r;
}

__device__ __forceinline__ g2_aff g2_generator() {
    return {{f_to_mont(f_from_words(G2Gen::X0)), f_to_mont(f_from_words(G2Gen::X1))},
            {f_to_mont(f_from_words(G2Gen::Y0)), f_to_mont(f_from_words(G2Gen::Y1))}};
}

}  // namespace zgr
