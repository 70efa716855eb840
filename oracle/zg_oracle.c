/*
 * zg_oracle.c -- CPU restatement of the 0G DA encoding path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP path, never the
 * product.  It is loaded by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py, nowhere else.
 *
 * Arithmetic: BN254 (arkworks ark-bn254 0.4.0 semantics), 4 x 64-bit limbs,
 * Montgomery form with R = 2^256, CIOS multiplication via unsigned __int128.
 * The reference crate names below are paths inside 0glabs/0g-da-encoder.
 */
#include "zg_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

typedef struct { uint64_t v[4]; } fe;

typedef struct {
    uint64_t m[4];
    uint64_t inv;   /* -m^{-1} mod 2^64 */
    uint64_t one[4]; /* R mod m */
    uint64_t r2[4];  /* R^2 mod m */
} fparams;

/* Fr = BN254 scalar field r, Fq = base field p. */
static const fparams FR = {
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0xc2e1f593efffffffULL,
    {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL},
    {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL}};
static const fparams FQ = {
    {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0x87d20782e4866389ULL,
    {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL},
    {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL}};

/* TWO_ADIC_ROOT_OF_UNITY of Fr = 5^((r-1)/2^28) (ark-bn254 FrConfig), canonical. */
static const uint64_t G28[4] = {0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL,
                                0x00e0a7eb8ef62abcULL, 0x2a3c09f0a58a7e85ULL};
#define TWO_ADICITY 28

static int g_threads = 1;

int zgo_set_threads(int n) {
    int old = g_threads;
    g_threads = n < 1 ? 1 : n;
    return old;
}

/* ------------------------------------------------------------------ */
/* field arithmetic                                                    */
/* ------------------------------------------------------------------ */

static int geq4(const uint64_t *a, const uint64_t *b) {
    for (int i = 3; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i];
    return 1;
}

static void sub4(uint64_t *o, const uint64_t *a, const uint64_t *b) {
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) {
        u128 d = (u128)a[i] - b[i] - br;
        o[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) & 1;
    }
}

static void add4(uint64_t *o, const uint64_t *a, const uint64_t *b) {
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
        u128 s = (u128)a[i] + b[i] + c;
        o[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
}

static void fmul(fe *o, const fe *a, const fe *b, const fparams *P) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c += (u128)a->v[j] * b->v[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        uint64_t m = t[0] * P->inv;
        c = (u128)m * P->m[0] + t[0];
        c >>= 64;
        for (int j = 1; j < 4; j++) {
            c += (u128)m * P->m[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    uint64_t r[4] = {t[0], t[1], t[2], t[3]};
    if (t[4] || geq4(r, P->m)) sub4(r, r, P->m);
    memcpy(o->v, r, 32);
}

static void fadd(fe *o, const fe *a, const fe *b, const fparams *P) {
    uint64_t s[4];
    add4(s, a->v, b->v); /* moduli < 2^254: no overflow */
    if (geq4(s, P->m)) sub4(s, s, P->m);
    memcpy(o->v, s, 32);
}

static void fsub(fe *o, const fe *a, const fe *b, const fparams *P) {
    uint64_t s[4];
    if (geq4(a->v, b->v)) {
        sub4(s, a->v, b->v);
    } else {
        add4(s, a->v, P->m);
        sub4(s, s, b->v);
    }
    memcpy(o->v, s, 32);
}

static int fis_zero(const fe *a) { return !(a->v[0] | a->v[1] | a->v[2] | a->v[3]); }
static void fset_one(fe *o, const fparams *P) { memcpy(o->v, P->one, 32); }

static void fto_mont(fe *o, const uint64_t *canon, const fparams *P) {
    fe a, r2;
    memcpy(a.v, canon, 32);
    memcpy(r2.v, P->r2, 32);
    fmul(o, &a, &r2, P);
}

static void ffrom_mont(uint64_t *canon, const fe *a, const fparams *P) {
    fe one = {{1, 0, 0, 0}}, r;
    fmul(&r, a, &one, P);
    memcpy(canon, r.v, 32);
}

/* a^e, e canonical little-endian limbs (n64 limbs) */
static void fpow(fe *o, const fe *a, const uint64_t *e, int n64, const fparams *P) {
    fe r, b = *a;
    fset_one(&r, P);
    for (int i = 0; i < n64; i++) {
        uint64_t w = e[i];
        for (int j = 0; j < 64; j++) {
            if (w & 1) fmul(&r, &r, &b, P);
            fmul(&b, &b, &b, P);
            w >>= 1;
        }
    }
    *o = r;
}

static void fpow_u64(fe *o, const fe *a, uint64_t e, const fparams *P) { fpow(o, a, &e, 1, P); }

static void finv(fe *o, const fe *a, const fparams *P) {
    uint64_t e[4];
    uint64_t two[4] = {2, 0, 0, 0};
    sub4(e, P->m, two);
    fpow(o, a, e, 4, P);
}

/* ---- exported field helpers ---- */
static const fparams *pick(int which) { return which ? &FQ : &FR; }

void zgo_fmul(int which, const uint64_t a[4], const uint64_t b[4], uint64_t out[4]) {
    fe x, y, z;
    memcpy(x.v, a, 32);
    memcpy(y.v, b, 32);
    fmul(&z, &x, &y, pick(which));
    memcpy(out, z.v, 32);
}
void zgo_to_mont(int which, const uint64_t a[4], uint64_t out[4]) {
    fe z;
    fto_mont(&z, a, pick(which));
    memcpy(out, z.v, 32);
}
void zgo_from_mont(int which, const uint64_t a[4], uint64_t out[4]) {
    fe x;
    memcpy(x.v, a, 32);
    ffrom_mont(out, &x, pick(which));
}
void zgo_from_mont_n(int which, const uint64_t *a, size_t n, uint64_t *out) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (n >= 4096)
    for (size_t i = 0; i < n; i++) {
        fe x;
        memcpy(x.v, a + 4 * i, 32);
        ffrom_mont(out + 4 * i, &x, pick(which));
    }
}

void zgo_finv(int which, const uint64_t a[4], uint64_t out[4]) {
    fe x, z;
    memcpy(x.v, a, 32);
    finv(&z, &x, pick(which));
    memcpy(out, z.v, 32);
}

/* ------------------------------------------------------------------ */
/* G1: y^2 = x^3 + 3 over Fq, Jacobian coordinates (Z = 0 is infinity) */
/* ------------------------------------------------------------------ */

typedef struct { fe x, y; int inf; } g1a; /* mont coordinates */
typedef struct { fe X, Y, Z; } g1j;

static void j_set_inf(g1j *p) { memset(p, 0, sizeof(*p)); }
static int j_is_inf(const g1j *p) { return fis_zero(&p->Z); }

static void j_from_a(g1j *o, const g1a *a) {
    if (a->inf) { j_set_inf(o); return; }
    o->X = a->x;
    o->Y = a->y;
    fset_one(&o->Z, &FQ);
}

/* dbl-2009-l (a = 0) */
static void j_dbl(g1j *o, const g1j *p) {
    if (j_is_inf(p)) { *o = *p; return; }
    const fparams *P = &FQ;
    fe A, B, C, D, E, F, t, X3, Y3, Z3;
    fmul(&A, &p->X, &p->X, P);
    fmul(&B, &p->Y, &p->Y, P);
    fmul(&C, &B, &B, P);
    fadd(&t, &p->X, &B, P);
    fmul(&t, &t, &t, P);
    fsub(&t, &t, &A, P);
    fsub(&t, &t, &C, P);
    fadd(&D, &t, &t, P);
    fadd(&E, &A, &A, P);
    fadd(&E, &E, &A, P);
    fmul(&F, &E, &E, P);
    fsub(&X3, &F, &D, P);
    fsub(&X3, &X3, &D, P);
    fsub(&t, &D, &X3, P);
    fmul(&Y3, &E, &t, P);
    fadd(&C, &C, &C, P);
    fadd(&C, &C, &C, P);
    fadd(&C, &C, &C, P);
    fsub(&Y3, &Y3, &C, P);
    fmul(&Z3, &p->Y, &p->Z, P);
    fadd(&Z3, &Z3, &Z3, P);
    o->X = X3;
    o->Y = Y3;
    o->Z = Z3;
}

/* add-2007-bl with the exceptional cases handled */
static void j_add(g1j *o, const g1j *p, const g1j *q) {
    if (j_is_inf(p)) { *o = *q; return; }
    if (j_is_inf(q)) { *o = *p; return; }
    const fparams *P = &FQ;
    fe Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, r, V, t, X3, Y3, Z3;
    fmul(&Z1Z1, &p->Z, &p->Z, P);
    fmul(&Z2Z2, &q->Z, &q->Z, P);
    fmul(&U1, &p->X, &Z2Z2, P);
    fmul(&U2, &q->X, &Z1Z1, P);
    fmul(&S1, &p->Y, &q->Z, P);
    fmul(&S1, &S1, &Z2Z2, P);
    fmul(&S2, &q->Y, &p->Z, P);
    fmul(&S2, &S2, &Z1Z1, P);
    fsub(&H, &U2, &U1, P);
    fsub(&r, &S2, &S1, P);
    if (fis_zero(&H)) {
        if (fis_zero(&r)) { j_dbl(o, p); return; }
        j_set_inf(o);
        return;
    }
    fadd(&I, &H, &H, P);
    fmul(&I, &I, &I, P);
    fmul(&J, &H, &I, P);
    fadd(&r, &r, &r, P);
    fmul(&V, &U1, &I, P);
    fmul(&X3, &r, &r, P);
    fsub(&X3, &X3, &J, P);
    fsub(&X3, &X3, &V, P);
    fsub(&X3, &X3, &V, P);
    fsub(&t, &V, &X3, P);
    fmul(&Y3, &r, &t, P);
    fmul(&t, &S1, &J, P);
    fadd(&t, &t, &t, P);
    fsub(&Y3, &Y3, &t, P);
    fadd(&Z3, &p->Z, &q->Z, P);
    fmul(&Z3, &Z3, &Z3, P);
    fsub(&Z3, &Z3, &Z1Z1, P);
    fsub(&Z3, &Z3, &Z2Z2, P);
    fmul(&Z3, &Z3, &H, P);
    o->X = X3;
    o->Y = Y3;
    o->Z = Z3;
}

/* madd-2007-bl: Jacobian + affine */
static void j_madd(g1j *o, const g1j *p, const g1a *q) {
    if (q->inf) { *o = *p; return; }
    if (j_is_inf(p)) { j_from_a(o, q); return; }
    const fparams *P = &FQ;
    fe Z1Z1, U2, S2, H, HH, I, J, r, V, t, X3, Y3, Z3;
    fmul(&Z1Z1, &p->Z, &p->Z, P);
    fmul(&U2, &q->x, &Z1Z1, P);
    fmul(&S2, &q->y, &p->Z, P);
    fmul(&S2, &S2, &Z1Z1, P);
    fsub(&H, &U2, &p->X, P);
    fsub(&r, &S2, &p->Y, P);
    if (fis_zero(&H)) {
        if (fis_zero(&r)) { j_dbl(o, p); return; }
        j_set_inf(o);
        return;
    }
    fmul(&HH, &H, &H, P);
    fadd(&I, &HH, &HH, P);
    fadd(&I, &I, &I, P);
    fmul(&J, &H, &I, P);
    fadd(&r, &r, &r, P);
    fmul(&V, &p->X, &I, P);
    fmul(&X3, &r, &r, P);
    fsub(&X3, &X3, &J, P);
    fsub(&X3, &X3, &V, P);
    fsub(&X3, &X3, &V, P);
    fsub(&t, &V, &X3, P);
    fmul(&Y3, &r, &t, P);
    fmul(&t, &p->Y, &J, P);
    fadd(&t, &t, &t, P);
    fsub(&Y3, &Y3, &t, P);
    fadd(&Z3, &p->Z, &H, P);
    fmul(&Z3, &Z3, &Z3, P);
    fsub(&Z3, &Z3, &Z1Z1, P);
    fsub(&Z3, &Z3, &HH, P);
    o->X = X3;
    o->Y = Y3;
    o->Z = Z3;
}

static void j_to_a(g1a *o, const g1j *p) {
    if (j_is_inf(p)) { memset(o, 0, sizeof(*o)); o->inf = 1; return; }
    fe zi, zi2, zi3;
    finv(&zi, &p->Z, &FQ);
    fmul(&zi2, &zi, &zi, &FQ);
    fmul(&zi3, &zi2, &zi, &FQ);
    fmul(&o->x, &p->X, &zi2, &FQ);
    fmul(&o->y, &p->Y, &zi3, &FQ);
    o->inf = 0;
}

/* scalar multiplication, scalar canonical (LSB-first double-and-add) */
static void j_mul(g1j *o, const g1j *p, const uint64_t *k) {
    g1j r, b = *p;
    j_set_inf(&r);
    for (int i = 0; i < 4; i++) {
        uint64_t w = k[i];
        for (int j = 0; j < 64; j++) {
            if (w & 1) j_add(&r, &r, &b);
            j_dbl(&b, &b);
            w >>= 1;
        }
    }
    *o = r;
}

/* output encoding: canonical x || y, zeros for infinity */
static void a_store_canon(uint8_t *out, const g1a *a) {
    if (a->inf) { memset(out, 0, 64); return; }
    uint64_t x[4], y[4];
    ffrom_mont(x, &a->x, &FQ);
    ffrom_mont(y, &a->y, &FQ);
    memcpy(out, x, 32);
    memcpy(out + 32, y, 32);
}

static void a_load_canon(g1a *a, const uint8_t *in) {
    uint64_t x[4], y[4];
    memcpy(x, in, 32);
    memcpy(y, in + 32, 32);
    if (!(x[0] | x[1] | x[2] | x[3] | y[0] | y[1] | y[2] | y[3])) {
        memset(a, 0, sizeof(*a));
        a->inf = 1;
        return;
    }
    fto_mont(&a->x, x, &FQ);
    fto_mont(&a->y, y, &FQ);
    a->inf = 0;
}

/* param-buffer encoding: x_mont || y_mont (bamt, fast_serde_bn254.rs:139-147) */
static void a_load_mont(g1a *a, const uint64_t *in) {
    memcpy(a->x.v, in, 32);
    memcpy(a->y.v, in + 4, 32);
    a->inf = 0;
}

static void a_store_mont(uint64_t *out, const g1a *a) {
    memcpy(out, a->x.v, 32);
    memcpy(out + 4, a->y.v, 32);
}

static void j_store_canon(uint8_t *out, const g1j *p) {
    g1a a;
    j_to_a(&a, p);
    a_store_canon(out, &a);
}

static g1a gen_affine(void) {
    g1a g;
    uint64_t one[4] = {1, 0, 0, 0}, two[4] = {2, 0, 0, 0};
    fto_mont(&g.x, one, &FQ);
    fto_mont(&g.y, two, &FQ);
    g.inf = 0;
    return g;
}

void zgo_g1_mul_gen(const uint64_t k[4], uint8_t out[64]) {
    g1a g = gen_affine();
    g1j G, R;
    j_from_a(&G, &g);
    j_mul(&R, &G, k);
    j_store_canon(out, &R);
}

void zgo_g1_mul(const uint8_t p[64], const uint64_t k[4], uint8_t out[64]) {
    g1a a;
    g1j P, R;
    a_load_canon(&a, p);
    j_from_a(&P, &a);
    j_mul(&R, &P, k);
    j_store_canon(out, &R);
}

void zgo_g1_add(const uint8_t p[64], const uint8_t q[64], uint8_t out[64]) {
    g1a a, b;
    g1j P, R;
    a_load_canon(&a, p);
    a_load_canon(&b, q);
    j_from_a(&P, &a);
    j_madd(&R, &P, &b);
    j_store_canon(out, &R);
}

/* ------------------------------------------------------------------ */
/* MSM: Pippenger bucket method (the group element is algorithm-independent;
 * the reference calls ark-ec VariableBaseMSM::msm, prove.rs:17,31,55,77)    */
/* ------------------------------------------------------------------ */

static unsigned get_bits(const uint64_t *k, int s, int c) {
    int limb = s >> 6, off = s & 63;
    if (limb >= 4) return 0;
    u128 v = k[limb] >> off;
    if (off + c > 64 && limb < 3) v |= (u128)k[limb + 1] << (64 - off);
    return (unsigned)(v & ((1u << c) - 1));
}

static void msm_jac(g1j *out, const uint64_t *bases_mont, const uint64_t *scalars, size_t n) {
    j_set_inf(out);
    if (n == 0) return;
    int lg = 0;
    while ((1ull << (lg + 1)) <= n) lg++;
    /* ark-ec window heuristic: 3 below 32 points, else ln(n) + 2 */
    int c = n < 32 ? 3 : (lg * 69 / 100) + 2;
    int nw = (254 + c - 1) / c;
    size_t nb = ((size_t)1 << c) - 1;
    g1j *bk = (g1j *)malloc(nb * sizeof(g1j));
    for (int w = nw - 1; w >= 0; w--) {
        for (int d = 0; d < c; d++) j_dbl(out, out);
        for (size_t b = 0; b < nb; b++) j_set_inf(&bk[b]);
        for (size_t i = 0; i < n; i++) {
            unsigned dgt = get_bits(scalars + 4 * i, w * c, c);
            if (!dgt) continue;
            g1a a;
            a_load_mont(&a, bases_mont + 8 * i);
            j_madd(&bk[dgt - 1], &bk[dgt - 1], &a);
        }
        g1j run, sum;
        j_set_inf(&run);
        j_set_inf(&sum);
        for (size_t b = nb; b-- > 0;) {
            j_add(&run, &run, &bk[b]);
            j_add(&sum, &sum, &run);
        }
        j_add(out, out, &sum);
    }
    free(bk);
}

void zgo_msm(const uint64_t *bases_mont, const uint64_t *scalars, size_t n, uint8_t out[64]) {
    g1j r;
    msm_jac(&r, bases_mont, scalars, n);
    j_store_canon(out, &r);
}

/* ------------------------------------------------------------------ */
/* radix-2 domain, ark-poly Radix2EvaluationDomain semantics           */
/* ------------------------------------------------------------------ */

static size_t bitrev(size_t i, int bits) {
    size_t r = 0;
    for (int b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

/* crates/amt/src/utils.rs:63-68 */
static size_t swap_bits(size_t n, int lo, int hi) {
    size_t lowest = n & (((size_t)1 << lo) - 1);
    size_t next = (n >> lo) & (((size_t)1 << hi) - 1);
    return (lowest << hi) | next;
}

/* get_root_of_unity(2^log_n): square the 2^28-th root (28 - log_n) times (mont) */
static void domain_root_m(fe *o, int log_n) {
    fto_mont(o, G28, &FR);
    for (int i = log_n; i < TWO_ADICITY; i++) fmul(o, o, o, &FR);
}

void zgo_domain_root(int log_n, uint64_t out[4]) {
    fe w;
    domain_root_m(&w, log_n);
    ffrom_mont(out, &w, &FR);
}

/* AMTParams::coset_factor, crates/amt/src/prove_params/generate.rs:128-137 */
static void coset_factor_m(fe *o, int log_n, int idx) {
    uint64_t pw = bitrev((size_t)idx, TWO_ADICITY - log_n);
    fe g;
    fto_mont(&g, G28, &FR);
    fpow_u64(o, &g, pw, &FR);
}

void zgo_coset_factor(int log_n, int idx, uint64_t out[4]) {
    fe c;
    coset_factor_m(&c, log_n, idx);
    ffrom_mont(out, &c, &FR);
}

/* DFT out[t] = sum_i a[i] w^(i t): iterative radix-2 DIT after a bit-reversal */
static void fft_core(fe *a, int log_n, const fe *omega) {
    size_t n = (size_t)1 << log_n;
    for (size_t i = 0; i < n; i++) {
        size_t r = bitrev(i, log_n);
        if (i < r) { fe t = a[i]; a[i] = a[r]; a[r] = t; }
    }
    fe *tw = (fe *)malloc((n / 2 + 1) * sizeof(fe));
    for (int s = 1; s <= log_n; s++) {
        size_t len = (size_t)1 << s, half = len >> 1;
        fe wl;
        fpow_u64(&wl, omega, (uint64_t)(n >> s), &FR);
        fset_one(&tw[0], &FR);
        for (size_t j = 1; j < half; j++) fmul(&tw[j], &tw[j - 1], &wl, &FR);
#pragma omp parallel for num_threads(g_threads) schedule(static) if (n >= 4096)
        for (size_t i = 0; i < n; i += len) {
            for (size_t j = 0; j < half; j++) {
                fe u = a[i + j], v;
                fmul(&v, &a[i + j + half], &tw[j], &FR);
                fadd(&a[i + j], &u, &v, &FR);
                fsub(&a[i + j + half], &u, &v, &FR);
            }
        }
    }
    free(tw);
}

void zgo_fft(uint64_t *a_mont, int log_n) {
    fe w;
    domain_root_m(&w, log_n);
    fft_core((fe *)a_mont, log_n, &w);
}

void zgo_ifft(uint64_t *a_mont, int log_n) {
    fe w, wi, ninv, nn;
    domain_root_m(&w, log_n);
    finv(&wi, &w, &FR);
    fft_core((fe *)a_mont, log_n, &wi);
    uint64_t nc[4] = {(uint64_t)1 << log_n, 0, 0, 0};
    fto_mont(&nn, nc, &FR);
    finv(&ninv, &nn, &FR);
    fe *a = (fe *)a_mont;
    size_t n = (size_t)1 << log_n;
    for (size_t i = 0; i < n; i++) fmul(&a[i], &a[i], &ninv, &FR);
}

/* to_coset_blob, crates/amt/src/blob/encode.rs:135-158 */
void zgo_to_coset_blob(const uint64_t *data_mont, int log_n, int coset, uint64_t *out_mont) {
    size_t n = (size_t)1 << log_n;
    memcpy(out_mont, data_mont, n * 32);
    if (coset == 0) return;
    zgo_ifft(out_mont, log_n);
    fe cf, base;
    coset_factor_m(&cf, log_n, coset);
    fe *a = (fe *)out_mont;
    fset_one(&base, &FR);
    for (size_t i = 0; i < n; i++) {
        fmul(&a[i], &a[i], &base, &FR);
        fmul(&base, &base, &cf, &FR);
    }
    zgo_fft(out_mont, log_n);
}

/* ------------------------------------------------------------------ */
/* G1 EC-FFT: the reference's Pairing::fast_fft / fast_ifft            */
/* (crates/amt/src/ec_algebra.rs:34-46, ark-poly fft over G1Projective) */
/* ------------------------------------------------------------------ */

static void ecfft_core(g1j *a, int log_n, const fe *omega) {
    size_t n = (size_t)1 << log_n;
    for (size_t i = 0; i < n; i++) {
        size_t r = bitrev(i, log_n);
        if (i < r) { g1j t = a[i]; a[i] = a[r]; a[r] = t; }
    }
    for (int s = 1; s <= log_n; s++) {
        size_t len = (size_t)1 << s, half = len >> 1;
        fe wl;
        fpow_u64(&wl, omega, (uint64_t)(n >> s), &FR);
        for (size_t i = 0; i < n; i += len) {
            fe w;
            fset_one(&w, &FR);
            for (size_t j = 0; j < half; j++) {
                uint64_t k[4];
                ffrom_mont(k, &w, &FR);
                g1j u = a[i + j], v, nv;
                j_mul(&v, &a[i + j + half], k);
                j_add(&a[i + j], &u, &v);
                nv = v;
                if (!j_is_inf(&nv)) fsub(&nv.Y, &(fe){{0, 0, 0, 0}}, &nv.Y, &FQ);
                j_add(&a[i + j + half], &u, &nv);
                fmul(&w, &w, &wl, &FR);
            }
        }
    }
}

static void ecfft_jac(g1j *a, int log_n, int inverse) {
    fe w;
    domain_root_m(&w, log_n);
    if (!inverse) {
        ecfft_core(a, log_n, &w);
        return;
    }
    fe wi, nn, ninv;
    finv(&wi, &w, &FR);
    ecfft_core(a, log_n, &wi);
    uint64_t nc[4] = {(uint64_t)1 << log_n, 0, 0, 0}, k[4];
    fto_mont(&nn, nc, &FR);
    finv(&ninv, &nn, &FR);
    ffrom_mont(k, &ninv, &FR);
    size_t n = (size_t)1 << log_n;
    for (size_t i = 0; i < n; i++) j_mul(&a[i], &a[i], k);
}

void zgo_ecfft(const uint8_t *in, int log_n, int inverse, uint8_t *out) {
    size_t n = (size_t)1 << log_n;
    g1j *a = (g1j *)malloc(n * sizeof(g1j));
    for (size_t i = 0; i < n; i++) {
        g1a t;
        a_load_canon(&t, in + 64 * i);
        j_from_a(&a[i], &t);
    }
    ecfft_jac(a, log_n, inverse);
    for (size_t i = 0; i < n; i++) j_store_canon(out + 64 * i, &a[i]);
    free(a);
}

/* ------------------------------------------------------------------ */
/* params from tau                                                     */
/* ------------------------------------------------------------------ */

/* store line element t at bit-reversed position (enact, generate.rs:122-126) */
static void store_enacted(uint64_t *line, int log_n, size_t t, const g1j *p) {
    g1a a;
    j_to_a(&a, p);
    a_store_mont(line + 8 * bitrev(t, log_n), &a);
}

/* Closed form of generate.rs:139-232 with tau' = tau / coset_factor:
 *   basis_t     = [ A_t (tau'^N - 1) ] G1
 *   quot^(d)_t  = [ A_t (tau'^m - w^(t m)) ] G1,  m = N / 2^d
 *   high_t      = [ tau^(3N) A_t (tau'^N - 1) ] G1
 * with A_t = w^t / (N (tau' - w^t)); each line is then bit-reversed. */
static int params_closed(const uint64_t tau[4], int log_n, int h, int coset, uint64_t *out) {
    size_t n = (size_t)1 << log_n;
    fe t_m, cf, cfi, tp, w, nn, ninv, tpN, one, tauH;
    fto_mont(&t_m, tau, &FR);
    coset_factor_m(&cf, log_n, coset);
    finv(&cfi, &cf, &FR);
    fmul(&tp, &t_m, &cfi, &FR);
    domain_root_m(&w, log_n);
    uint64_t nc[4] = {n, 0, 0, 0};
    fto_mont(&nn, nc, &FR);
    finv(&ninv, &nn, &FR);
    fpow_u64(&tpN, &tp, n, &FR);
    fset_one(&one, &FR);
    /* high_start = 2^(log_n + 2) - 2^log_n (power_tau.rs:66-71) */
    fpow_u64(&tauH, &t_m, 3 * (uint64_t)n, &FR);
    g1a g = gen_affine();
    g1j G;
    j_from_a(&G, &g);
    int bad = 0;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16) reduction(| : bad)
    for (size_t t = 0; t < n; t++) {
        fe wt, den, A, s, tmp, tpm, wtm;
        uint64_t k[4];
        g1j P;
        fpow_u64(&wt, &w, t, &FR);
        fsub(&den, &tp, &wt, &FR);
        if (fis_zero(&den)) { bad = 1; continue; }
        finv(&A, &den, &FR);
        fmul(&A, &A, &wt, &FR);
        fmul(&A, &A, &ninv, &FR);
        /* basis */
        fsub(&tmp, &tpN, &one, &FR);
        fmul(&s, &A, &tmp, &FR);
        ffrom_mont(k, &s, &FR);
        j_mul(&P, &G, k);
        store_enacted(out + (size_t)h * n * 8, log_n, t, &P);
        /* high basis */
        fmul(&s, &s, &tauH, &FR);
        ffrom_mont(k, &s, &FR);
        j_mul(&P, &G, k);
        store_enacted(out + (size_t)(h + 1) * n * 8, log_n, t, &P);
        /* quotients */
        for (int d = 1; d <= h; d++) {
            uint64_t m = n >> d;
            fpow_u64(&tpm, &tp, m, &FR);
            fpow_u64(&wtm, &wt, m, &FR);
            fsub(&tmp, &tpm, &wtm, &FR);
            fmul(&s, &A, &tmp, &FR);
            ffrom_mont(k, &s, &FR);
            j_mul(&P, &G, k);
            store_enacted(out + (size_t)(d - 1) * n * 8, log_n, t, &P);
        }
    }
    return bad ? -1 : 0;
}

/* Step-for-step restatement: PowerTau::setup_inner (power_tau.rs:54-80), the coset
 * re-scaling of from_pp (generate.rs:152-170), gen_basis (:197-203), gen_quotients
 * (:205-232) through the G1 EC-FFT, and enact (:122-126). */
static int params_ecfft(const uint64_t tau[4], int log_n, int h, int coset, uint64_t *out) {
    size_t n = (size_t)1 << log_n;
    g1a g = gen_affine();
    g1j G;
    j_from_a(&G, &g);
    g1j *g1pp = (g1j *)malloc(n * sizeof(g1j));
    g1j *high = (g1j *)malloc(n * sizeof(g1j));
    g1j *buf = (g1j *)malloc(n * sizeof(g1j));
    fe t_m, tp, cf, wsc, one;
    fto_mont(&t_m, tau, &FR);
    fset_one(&one, &FR);
    coset_factor_m(&cf, log_n, coset);
    finv(&wsc, &cf, &FR); /* w = 1 / coset_factor */
    fe tauH;
    fpow_u64(&tauH, &t_m, 3 * (uint64_t)n, &FR);
    for (size_t i = 0; i < n; i++) {
        uint64_t k[4];
        fe s, wi;
        fpow_u64(&s, &t_m, i, &FR);
        ffrom_mont(k, &s, &FR);
        j_mul(&g1pp[i], &G, k);
        fmul(&s, &s, &tauH, &FR);
        ffrom_mont(k, &s, &FR);
        j_mul(&high[i], &G, k);
        if (coset > 0) {
            fpow_u64(&wi, &wsc, i, &FR);
            ffrom_mont(k, &wi, &FR);
            j_mul(&g1pp[i], &g1pp[i], k);
            j_mul(&high[i], &high[i], k);
        }
    }
    (void)tp;
    /* basis = ifft(g1pp) */
    memcpy(buf, g1pp, n * sizeof(g1j));
    ecfft_jac(buf, log_n, 1);
    for (size_t t = 0; t < n; t++) store_enacted(out + (size_t)h * n * 8, log_n, t, &buf[t]);
    /* high_basis = ifft(high_g1pp) */
    memcpy(buf, high, n * sizeof(g1j));
    ecfft_jac(buf, log_n, 1);
    for (size_t t = 0; t < n; t++) store_enacted(out + (size_t)(h + 1) * n * 8, log_n, t, &buf[t]);
    /* quotients */
    fe nn, ninv;
    uint64_t nc[4] = {n, 0, 0, 0}, kinv[4];
    fto_mont(&nn, nc, &FR);
    finv(&ninv, &nn, &FR);
    ffrom_mont(kinv, &ninv, &FR);
    for (int d = 1; d <= h; d++) {
        size_t m = n >> d;
        for (size_t i = 0; i < n; i++) j_set_inf(&buf[i]);
        for (size_t i = 1; i <= m; i++) buf[i] = g1pp[m - i];
        ecfft_jac(buf, log_n, 0);
        for (size_t t = 0; t < n; t++) {
            j_mul(&buf[t], &buf[t], kinv);
            store_enacted(out + (size_t)(d - 1) * n * 8, log_n, t, &buf[t]);
        }
    }
    free(g1pp);
    free(high);
    free(buf);
    return 0;
}

/* One point of the closed form above: line L (quotients 0..h-1, basis h, high h+1)
 * at stored position pos, written as x_mont || y_mont. */
int zgo_params_point(const uint64_t tau[4], int log_n, int h, int coset, int line, uint64_t pos,
                     uint64_t out[8]) {
    uint64_t n = (uint64_t)1 << log_n;
    uint64_t t = bitrev(pos, log_n);
    fe t_m, cf, cfi, tp, w, wt, nn, ninv, den, A, s, tmp, one;
    fto_mont(&t_m, tau, &FR);
    coset_factor_m(&cf, log_n, coset);
    finv(&cfi, &cf, &FR);
    fmul(&tp, &t_m, &cfi, &FR);
    domain_root_m(&w, log_n);
    fpow_u64(&wt, &w, t, &FR);
    uint64_t nc[4] = {n, 0, 0, 0};
    fto_mont(&nn, nc, &FR);
    finv(&ninv, &nn, &FR);
    fset_one(&one, &FR);
    fsub(&den, &tp, &wt, &FR);
    if (fis_zero(&den)) return -1;
    finv(&A, &den, &FR);
    fmul(&A, &A, &wt, &FR);
    fmul(&A, &A, &ninv, &FR);
    if (line >= h) {
        fpow_u64(&tmp, &tp, n, &FR);
        fsub(&tmp, &tmp, &one, &FR);
        fmul(&s, &A, &tmp, &FR);
        if (line == h + 1) {
            fpow_u64(&tmp, &t_m, 3 * n, &FR);
            fmul(&s, &s, &tmp, &FR);
        }
    } else {
        uint64_t m = n >> (line + 1);
        fe tpm, wtm;
        fpow_u64(&tpm, &tp, m, &FR);
        fpow_u64(&wtm, &wt, m, &FR);
        fsub(&tmp, &tpm, &wtm, &FR);
        fmul(&s, &A, &tmp, &FR);
    }
    uint64_t k[4];
    ffrom_mont(k, &s, &FR);
    g1a g = gen_affine(), a;
    g1j G, Pt;
    j_from_a(&G, &g);
    j_mul(&Pt, &G, k);
    j_to_a(&a, &Pt);
    a_store_mont(out, &a);
    return 0;
}

/* f(x) for the degree < N interpolant of evaluations y over the 2^log_n domain:
 * f(x) = sum_j y_j L_j(x), L_j(x) = (w^j / N)(x^N - 1)/(x - w^j)  (x off the domain) */
int zgo_eval_interp(const uint64_t *y_mont, int log_n, const uint64_t x[4], uint64_t out[4]) {
    size_t n = (size_t)1 << log_n;
    fe xm, w, wj, one, xN, nn, ninv, acc;
    fto_mont(&xm, x, &FR);
    domain_root_m(&w, log_n);
    fset_one(&one, &FR);
    fpow_u64(&xN, &xm, n, &FR);
    uint64_t nc[4] = {n, 0, 0, 0};
    fto_mont(&nn, nc, &FR);
    finv(&ninv, &nn, &FR);
    fe *den = (fe *)malloc(n * sizeof(fe)), *pre = (fe *)malloc((n + 1) * sizeof(fe));
    fe *wp = (fe *)malloc(n * sizeof(fe));
    wj = one;
    for (size_t j = 0; j < n; j++) {
        wp[j] = wj;
        fsub(&den[j], &xm, &wj, &FR);
        if (fis_zero(&den[j])) { free(den); free(pre); free(wp); return -1; }
        fmul(&wj, &wj, &w, &FR);
    }
    /* batch inversion (Montgomery's trick) */
    pre[0] = one;
    for (size_t j = 0; j < n; j++) fmul(&pre[j + 1], &pre[j], &den[j], &FR);
    fe inv_all;
    finv(&inv_all, &pre[n], &FR);
    memset(&acc, 0, sizeof(acc));
    for (size_t j = n; j-- > 0;) {
        fe dinv, term, yj;
        fmul(&dinv, &inv_all, &pre[j], &FR);
        fmul(&inv_all, &inv_all, &den[j], &FR);
        memcpy(yj.v, y_mont + 4 * j, 32);
        fmul(&term, &yj, &wp[j], &FR);
        fmul(&term, &term, &dinv, &FR);
        fadd(&acc, &acc, &term, &FR);
    }
    fe t;
    fsub(&t, &xN, &one, &FR);
    fmul(&acc, &acc, &t, &FR);
    fmul(&acc, &acc, &ninv, &FR);
    ffrom_mont(out, &acc, &FR);
    free(den);
    free(pre);
    free(wp);
    return 0;
}

int zgo_params_from_tau(const uint64_t tau[4], int log_n, int h, int coset, int method,
                        uint64_t *out_bases) {
    if (log_n < 1 || log_n > TWO_ADICITY || h < 1 || h > log_n) return -2;
    if (method == 0) return params_closed(tau, log_n, h, coset, out_bases);
    return params_ecfft(tau, log_n, h, coset, out_bases);
}

/* ------------------------------------------------------------------ */
/* AMT prover (crates/amt/src/prove_params/prove.rs:20-103)            */
/* ------------------------------------------------------------------ */

static void to_canon_scalars(uint64_t *out, const uint64_t *mont, size_t n) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t i = 0; i < n; i++) {
        fe a;
        memcpy(a.v, mont + 4 * i, 32);
        ffrom_mont(out + 4 * i, &a, &FR);
    }
}

/* chunked MSMs: out[b] = MSM(bases[b*cs ..], s[b*cs ..]) for b < cnt */
static void chunk_msms(g1j *out, const uint64_t *bases, const uint64_t *s, size_t cs, size_t cnt) {
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (size_t b = 0; b < cnt; b++) msm_jac(&out[b], bases + 8 * b * cs, s + 4 * b * cs, cs);
}

/* one large MSM split over threads (used for the high commitment) */
static void par_msm(g1j *out, const uint64_t *bases, const uint64_t *s, size_t n) {
    int parts = g_threads;
    while (parts > 1 && (n % (size_t)parts)) parts--;
    g1j *tmp = (g1j *)malloc(parts * sizeof(g1j));
    chunk_msms(tmp, bases, s, n / parts, parts);
    j_set_inf(out);
    for (int i = 0; i < parts; i++) j_add(out, out, &tmp[i]);
    free(tmp);
}

int zgo_amt_prove(const uint64_t *bases_mont, int log_n, int h, const uint64_t *ri_mont,
                  uint8_t *out_commitment, uint8_t *out_comm_tree, uint8_t *out_proof_tree,
                  uint8_t *out_high) {
    size_t n = (size_t)1 << log_n;
    size_t nb = (size_t)1 << h, bs = n / nb;
    uint64_t *s = (uint64_t *)malloc(n * 32);
    to_canon_scalars(s, ri_mont, n);
    g1j *tmp = (g1j *)malloc(nb * sizeof(g1j));
    /* gen_prove_tree (prove.rs:58-83): depth d has 2^d MSMs of n / 2^d points */
    size_t off = 0;
    for (int d = 1; d <= h; d++) {
        size_t items = (size_t)1 << d, cs = n / items;
        chunk_msms(tmp, bases_mont + (size_t)(d - 1) * n * 8, s, cs, items);
        for (size_t k = 0; k < items; k++) j_store_canon(out_proof_tree + 64 * (off + k), &tmp[k]);
        off += items;
    }
    /* gen_commitment_tree + build_commitment_tree (prove.rs:20-52) */
    chunk_msms(tmp, bases_mont + (size_t)h * n * 8, s, bs, nb);
    size_t len = nb;
    while (len > 1) {
        /* this layer (len points) is commitments[log2(len) - 1], which starts at
         * point offset sum_{d < log2(len) - 1} 2^(d+1) = len - 2 */
        size_t loff = len - 2;
        for (size_t k = 0; k < len; k++) j_store_canon(out_comm_tree + 64 * (loff + k), &tmp[k]);
        for (size_t k = 0; k < len / 2; k++) j_add(&tmp[k], &tmp[2 * k], &tmp[2 * k + 1]);
        len /= 2;
    }
    j_store_canon(out_commitment, &tmp[0]);
    /* gen_high_commitment (prove.rs:54-56) */
    g1j hc;
    par_msm(&hc, bases_mont + (size_t)(h + 1) * n * 8, s, n);
    j_store_canon(out_high, &hc);
    free(tmp);
    free(s);
    return 0;
}

/* ag-cuda-ec multiple_multiexp_mt semantics as called by prove_gpu.rs:77-84 */
int zgo_multiple_multiexp(const uint64_t *bases_mont, int log_n, int h, const uint64_t *ri_mont,
                          uint8_t *out_points) {
    size_t n = (size_t)1 << log_n;
    size_t nb = (size_t)1 << h, bs = n / nb;
    uint64_t *s = (uint64_t *)malloc(n * 32);
    to_canon_scalars(s, ri_mont, n);
    g1j *tmp = (g1j *)malloc(nb * sizeof(g1j));
    for (int L = 0; L < h + 2; L++) {
        chunk_msms(tmp, bases_mont + (size_t)L * n * 8, s, bs, nb);
        for (size_t b = 0; b < nb; b++) j_store_canon(out_points + 64 * (L * nb + b), &tmp[b]);
    }
    free(tmp);
    free(s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Keccak-256 (tiny-keccak 2.0.2 Keccak::v256: rate 136, pad 0x01)     */
/* ------------------------------------------------------------------ */

static const uint64_t KRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

static uint64_t rotl(uint64_t x, int s) { return s ? (x << s) | (x >> (64 - s)) : x; }

/* Keccak-f[1600] written from the FIPS 202 step definitions (theta, rho, pi, chi, iota) */
static void keccak_f(uint64_t A[25]) {
    for (int rnd = 0; rnd < 24; rnd++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        /* rho + pi: offsets from the (x, y) -> (y, 2x + 3y) walk */
        int x = 1, y = 0;
        B[0] = A[0];
        uint64_t cur = A[1];
        for (int t = 0; t < 24; t++) {
            int nx = y, ny = (2 * x + 3 * y) % 5;
            int off = ((t + 1) * (t + 2) / 2) % 64;
            uint64_t nxt = A[nx + 5 * ny];
            B[nx + 5 * ny] = rotl(cur, off);
            cur = nxt;
            x = nx;
            y = ny;
        }
        for (int yy = 0; yy < 5; yy++)
            for (int xx = 0; xx < 5; xx++)
                A[xx + 5 * yy] = B[xx + 5 * yy] ^ (~B[(xx + 1) % 5 + 5 * yy] & B[(xx + 2) % 5 + 5 * yy]);
        A[0] ^= KRC[rnd];
    }
}

void zgo_keccak256(const uint8_t *msg, size_t len, uint8_t pad, uint8_t out[32]) {
    uint64_t A[25];
    memset(A, 0, sizeof(A));
    const size_t rate = 136;
    while (len >= rate) {
        for (size_t i = 0; i < rate / 8; i++) {
            uint64_t w;
            memcpy(&w, msg + 8 * i, 8);
            A[i] ^= w;
        }
        keccak_f(A);
        msg += rate;
        len -= rate;
    }
    uint8_t blk[136];
    memset(blk, 0, rate);
    memcpy(blk, msg, len);
    blk[len] ^= pad;
    blk[rate - 1] ^= 0x80;
    for (size_t i = 0; i < rate / 8; i++) {
        uint64_t w;
        memcpy(&w, blk + 8 * i, 8);
        A[i] ^= w;
    }
    keccak_f(A);
    memcpy(out, A, 32);
}

/* EncodedBlobMerkle::build (crates/encoder/src/merkle/blob.rs:15-34),
 * keccak_chunked (crates/encoder/src/utils.rs:44-57), compute_file_root
 * (crates/encoder/src/encoder/blob.rs:64-72) */
int zgo_merkle(const uint8_t *data, size_t n_elems, int coset_n, int levels,
               uint8_t *out_layers, uint8_t out_file_root[32]) {
    if (n_elems % 8) return -1;
    size_t cnt = n_elems / 8;
    uint8_t *cur = (uint8_t *)malloc(cnt * 32);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t i = 0; i < cnt; i++) zgo_keccak256(data + 256 * i, 256, 0x01, cur + 32 * i);
    /* collect all layers bottom-up */
    int nl = 1;
    size_t cap = 64;
    uint8_t **layers = (uint8_t **)malloc(cap * sizeof(uint8_t *));
    size_t *lens = (size_t *)malloc(cap * sizeof(size_t));
    layers[0] = cur;
    lens[0] = cnt;
    while (cnt > (size_t)coset_n) {
        if (cnt % 2) break;
        size_t nc = cnt / 2;
        uint8_t *nx = (uint8_t *)malloc(nc * 32);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (size_t i = 0; i < nc; i++) zgo_keccak256(cur + 64 * i, 64, 0x01, nx + 32 * i);
        layers[nl] = nx;
        lens[nl] = nc;
        nl++;
        cur = nx;
        cnt = nc;
    }
    /* tree[k] = layers[nl - 1 - k] */
    int ok = levels < nl;
    if (ok && out_layers) {
        size_t off = 0;
        for (int k = 0; k <= levels; k++) {
            int li = nl - 1 - k;
            memcpy(out_layers + off, layers[li], lens[li] * 32);
            off += lens[li] * 32;
        }
    }
    const uint8_t *roots = layers[nl - 1];
    if (lens[nl - 1] == 1) {
        memcpy(out_file_root, roots, 32);
    } else {
        uint8_t buf[64], t[32];
        memcpy(buf, roots, 64);
        zgo_keccak256(buf, 64, 0x01, t);
        if (lens[nl - 1] == 3) {
            memcpy(buf, t, 32);
            memcpy(buf + 32, roots + 64, 32);
            zgo_keccak256(buf, 64, 0x01, t);
        }
        memcpy(out_file_root, t, 32);
    }
    for (int i = 0; i < nl; i++) free(layers[i]);
    free(layers);
    free(lens);
    return ok ? nl : -1;
}

/* ------------------------------------------------------------------ */
/* packing: RawData::try_from (raw_data.rs:55-73), RawBlob::from       */
/* (raw_blob.rs:33-40), raw_unit_to_scalar (utils.rs:10-15)            */
/* ------------------------------------------------------------------ */

int zgo_raw_to_blob(const uint8_t *raw, size_t len, int log_n, uint64_t *out_mont) {
    size_t n = (size_t)1 << log_n;
    size_t max_blob = 31 * n, max_raw = max_blob - 4;
    if (len > max_raw) return 1;
    uint8_t *pad = (uint8_t *)calloc(max_blob, 1);
    if (len) memcpy(pad, raw, len);
    uint32_t l32 = (uint32_t)len;
    memcpy(pad + max_raw, &l32, 4);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t i = 0; i < n; i++) {
        uint64_t v[4] = {0, 0, 0, 0};
        memcpy(v, pad + 31 * i, 31);
        fe m;
        fto_mont(&m, v, &FR);
        memcpy(out_mont + 4 * i, m.v, 32);
    }
    free(pad);
    return 0;
}

/* ------------------------------------------------------------------ */
/* full encode: EncodedBlob::build (crates/encoder/src/encoder/blob.rs:32-42) */
/* ------------------------------------------------------------------ */

int zgo_encode(const uint64_t *const *bases, int log_col, int log_row, int coset_n,
               const uint8_t *raw, size_t len,
               uint8_t *out_data, uint8_t *out_commit, uint8_t *out_comm_tree,
               uint8_t *out_proof_tree, uint8_t *out_high, uint8_t *out_merkle,
               uint8_t *out_root) {
    int log_n = log_col + log_row, h = log_row;
    size_t n = (size_t)1 << log_n;
    size_t tree_pts = ((size_t)1 << (h + 1)) - 2;
    uint64_t *blob = (uint64_t *)malloc(n * 32);
    if (zgo_raw_to_blob(raw, len, log_n, blob)) { free(blob); return 1; }
    /* process_blob (crates/amt/src/blob/encode.rs:91-108):
     * change_matrix_direction(LOG_COL, LOG_ROW) -> column-major */
    uint64_t *pts = (uint64_t *)malloc(n * 32);
    for (size_t i = 0; i < n; i++) memcpy(pts + 4 * swap_bits(i, log_col, log_row), blob + 4 * i, 32);
    uint64_t *y = (uint64_t *)malloc(n * 32);
    uint64_t *ri = (uint64_t *)malloc(n * 32);
    int rc = 0;
    for (int c = 0; c < coset_n; c++) {
        zgo_to_coset_blob(pts, log_n, c, y);
        /* HalfBlob::generate (encode.rs:171-183): index_reverse, prove, restore,
         * change_matrix_direction(LOG_ROW, LOG_COL) -> row-major */
        for (size_t i = 0; i < n; i++) memcpy(ri + 4 * bitrev(i, log_n), y + 4 * i, 32);
        uint8_t com[64];
        zgo_amt_prove(bases[c], log_n, h, ri, com, out_comm_tree + 64 * tree_pts * c,
                      out_proof_tree + 64 * tree_pts * c, out_high + 64 * c);
        if (c == 0) memcpy(out_commit, com, 64);
        else if (memcmp(out_commit, com, 64)) rc = 2; /* assert_commitment_consistent */
        uint8_t *dst = out_data + (size_t)c * n * 32;
        for (size_t i = 0; i < n; i++) {
            fe a;
            memcpy(a.v, y + 4 * i, 32);
            /* scalar_to_h256 (crates/encoder/src/utils.rs:17-20) */
            ffrom_mont((uint64_t *)(dst + 32 * swap_bits(i, log_row, log_col)), &a, &FR);
        }
    }
    zgo_merkle(out_data, (size_t)coset_n * n, coset_n, log_row, out_merkle, out_root);
    free(blob);
    free(pts);
    free(y);
    free(ri);
    return rc;
}

/* ------------------------------------------------------------------ */
/* RS recovery: recover_from_da_slice (crates/recovery/src/lib.rs:15-43) */
/* ------------------------------------------------------------------ */

/* A dense polynomial: coefficients c[0..len), len >= 1. */
typedef struct { fe *c; size_t len; } rpoly;

static int ilog2_sz(size_t n) {
    int l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}

/* Poly::dense_from_vec / to_vec (crates/recovery/src/poly.rs:236-246, 196-216):
 * drop trailing zeros; the zero polynomial keeps one zero coefficient. */
static void rp_trim(rpoly *p) {
    while (p->len > 1 && fis_zero(&p->c[p->len - 1])) p->len--;
}

/* fx_to_fkx (crates/recovery/src/utils.rs:67-79): c_i -> c_i k^i */
static void rp_scale_powers(fe *c, size_t len, const fe *k) {
    fe base;
    fset_one(&base, &FR);
    for (size_t i = 0; i < len; i++) {
        fmul(&c[i], &c[i], &base, &FR);
        fmul(&base, &base, k, &FR);
    }
}

/* coeffs_to_evals_coset (utils.rs:81-105): evaluations of f on coset idx of the
 * N-domain (f(cw * w_N^s), s < N), then change_matrix_direction(ROW_LOG, COL_LOG)
 * to row-major.  The FFT runs over max(next_pow2(len), N) points and keeps every
 * (that / N)-th value. */
static void rp_evals_coset(const rpoly *f, int log_col, int log_row, int idx, fe *out) {
    int log_n = log_col + log_row;
    size_t n = (size_t)1 << log_n;
    int log_f = ilog2_sz(f->len);
    if (log_f < log_n) log_f = log_n;
    size_t F = (size_t)1 << log_f;
    fe *a = (fe *)calloc(F, sizeof(fe));
    memcpy(a, f->c, f->len * sizeof(fe));
    if (idx) {
        fe cw;
        coset_factor_m(&cw, log_n, idx);
        rp_scale_powers(a, f->len, &cw);
    }
    zgo_fft((uint64_t *)a, log_f);
    size_t step = F / n;
    for (size_t s = 0; s < n; s++) out[swap_bits(s, log_row, log_col)] = a[s * step];
    free(a);
}

/* coeffs_to_evals_larger (utils.rs:107-111): the CM cosets one after the other */
static void rp_evals_larger(const rpoly *f, int log_col, int log_row, int cm, fe *out) {
    size_t n = (size_t)1 << (log_col + log_row);
    for (int c = 0; c < cm; c++) rp_evals_coset(f, log_col, log_row, c, out + (size_t)c * n);
}

/* evals_to_poly (utils.rs:117-143): per chunk change_matrix_direction(COL_LOG, ROW_LOG),
 * chunks in bit-reversed coset order, change_matrix_direction(log N, log CM), then
 * the inverse DFT over the CM * N domain. */
static rpoly rp_from_evals(const fe *evals, int log_col, int log_row, int cm) {
    int log_n = log_col + log_row, log_cm = ilog2_sz((size_t)cm);
    size_t n = (size_t)1 << log_n, big = n * (size_t)cm;
    fe *t = (fe *)malloc(big * sizeof(fe));
    for (int x = 0; x < cm; x++) {
        const fe *chunk = evals + bitrev((size_t)x, log_cm) * n;
        for (size_t i = 0; i < n; i++) {
            size_t s = swap_bits(i, log_col, log_row); /* row-major -> natural order */
            t[swap_bits((size_t)x * n + s, log_n, log_cm)] = chunk[i];
        }
    }
    zgo_ifft((uint64_t *)t, log_n + log_cm);
    rpoly p = {t, big};
    rp_trim(&p);
    return p;
}

/* Status codes of zgo_recover (the reference's outcome each stands for):
 *   0 Ok(raw data)
 *   1 Err("Not enough valid lines")                      lib.rs:34-36
 *   2 Err("Cannot recover data: RowIdOverflow")          data_poly.rs:39-42, lib.rs:38-39
 *   3 Err("Incorrect scalar")                            raw_blob.rs:47-49
 *   4 Err("Incorrect length field")                      raw_data.rs:27-29
 *   5 Err("Non zero in padding range")                   raw_data.rs:31-33
 *   6 panic: an element >= r (MontConfig::from_bigint(..).unwrap(), utils.rs:26-34)
 *   7 panic: recovered degree above N (assert, data_poly.rs:112)
 *   8 Err("Cannot recover data: ExtaustiveK")            data_poly.rs:117
 * rows: n_rows rows of 2^log_col canonical 32 B elements (merkle_row bytes), ids
 * strictly increasing (a BTreeMap's keys).  out_raw: 31 * N bytes (the padded
 * RawData); *out_len = its actual length. */
int zgo_recover(const uint32_t *ids, const uint8_t *rows, size_t n_rows, int log_col, int log_row,
                int coset_n, uint8_t *out_raw, size_t *out_len) {
    const int log_n = log_col + log_row;
    const size_t ROW = (size_t)1 << log_row, COL = (size_t)1 << log_col, n = (size_t)1 << log_n;
    int cm = 1;
    while (cm < coset_n) cm <<= 1; /* COSET_MORE = COSET_N.next_power_of_two() (zpoly.rs:13) */
    const size_t big = n * (size_t)cm, cm_rows = ROW * (size_t)cm;
    /* raw_slice_to_line (utils.rs:26-34) converts every row before the count is
     * looked at (lib.rs:21-36) */
    for (size_t e = 0; e < n_rows * COL; e++) {
        uint64_t v[4];
        memcpy(v, rows + e * 32, 32);
        if (geq4(v, FR.m)) return 6;
    }
    if (n_rows < ROW) return 1;
    if (ids[n_rows - 1] >= ROW * (size_t)coset_n) return 2; /* check_input (data_poly.rs:29-51) */
    /* convert_input (data_poly.rs:53-70) */
    fe *data = (fe *)calloc(big, sizeof(fe));
    unsigned char *present = (unsigned char *)calloc(cm_rows, 1);
    for (size_t k = 0; k < n_rows; k++) {
        present[ids[k]] = 1;
        for (size_t j = 0; j < COL; j++) {
            uint64_t v[4];
            memcpy(v, rows + (k * COL + j) * 32, 32);
            fto_mont(&data[(size_t)ids[k] * COL + j], v, &FR);
        }
    }
    /* zpoly (zpoly.rs:63-82): prod over erased rows of (x^COL + f0(row)), f0 from ZBlob
     * (zpoly.rs:21-61): f0(c, l) = m0 * w_power^l, w_power = w_N^COL, m0 = -1 for
     * coset 0 and -coset_factor(N, c)^COL otherwise.  A product of polynomials in
     * x^COL: multiplied here in y = x^COL and spread out. */
    fe wp, one, mone;
    domain_root_m(&wp, log_n);
    fpow_u64(&wp, &wp, (uint64_t)COL, &FR);
    fset_one(&one, &FR);
    fe zero_fe;
    memset(&zero_fe, 0, sizeof zero_fe);
    fsub(&mone, &zero_fe, &one, &FR);
    size_t ne = 0;
    for (size_t r = 0; r < cm_rows; r++) ne += !present[r];
    fe *P = (fe *)calloc(ne + 1, sizeof(fe)); /* P(y) = prod (y + f0) */
    P[0] = one;
    size_t deg = 0;
    for (size_t r = 0; r < cm_rows; r++) {
        if (present[r]) continue;
        size_t c = r / ROW, l = r % ROW;
        fe m0, f0;
        if (c == 0) {
            m0 = mone;
        } else {
            fe cf;
            coset_factor_m(&cf, log_n, (int)c);
            fpow_u64(&cf, &cf, (uint64_t)COL, &FR);
            fsub(&m0, &zero_fe, &cf, &FR);
        }
        fpow_u64(&f0, &wp, (uint64_t)l, &FR);
        fmul(&f0, &f0, &m0, &FR);
        /* P <- P * (y + f0) */
        P[deg + 1] = P[deg];
        for (size_t i = deg; i > 0; i--) {
            fe t;
            fmul(&t, &P[i], &f0, &FR);
            fadd(&P[i], &P[i - 1], &t, &FR);
        }
        fmul(&P[0], &P[0], &f0, &FR);
        deg++;
    }
    rpoly Z = {(fe *)calloc(deg * COL + 1, sizeof(fe)), deg * COL + 1};
    for (size_t i = 0; i <= deg; i++) Z.c[i * COL] = P[i];
    free(P);
    free(present);
    /* data_times_zpoly (data_times_zpoly.rs:14-32) */
    fe *ev = (fe *)malloc(big * sizeof(fe));
    rp_evals_larger(&Z, log_col, log_row, cm, ev);
    for (size_t i = 0; i < big; i++) fmul(&ev[i], &ev[i], &data[i], &FR);
    rpoly DZ = rp_from_evals(ev, log_col, log_row, cm);
    free(data);
    /* the retry loop over k (data_poly.rs:86-116); k = 5, 6, ... here in place of
     * random draws: any k that is not a CM*N-th root of unity gives Z(kx) != 0 */
    int rc = 8;
    fe *zk = (fe *)malloc(big * sizeof(fe));
    for (uint64_t kk = 5; kk < 105 && rc == 8; kk++) {
        uint64_t kc[4] = {kk, 0, 0, 0};
        fe k, kinv;
        fto_mont(&k, kc, &FR);
        finv(&kinv, &k, &FR);
        rpoly Zk = {(fe *)malloc(Z.len * sizeof(fe)), Z.len};
        memcpy(Zk.c, Z.c, Z.len * sizeof(fe));
        rp_scale_powers(Zk.c, Zk.len, &k);
        rp_evals_larger(&Zk, log_col, log_row, cm, zk);
        free(Zk.c);
        int has_zero = 0;
        for (size_t i = 0; i < big && !has_zero; i++) has_zero = fis_zero(&zk[i]);
        if (has_zero) continue;
        { /* batch_inversion (Montgomery's trick) */
            fe *pre = (fe *)malloc(big * sizeof(fe)), acc, inv;
            fset_one(&acc, &FR);
            for (size_t i = 0; i < big; i++) {
                pre[i] = acc;
                fmul(&acc, &acc, &zk[i], &FR);
            }
            finv(&inv, &acc, &FR);
            for (size_t i = big; i-- > 0;) {
                fe t;
                fmul(&t, &inv, &pre[i], &FR);
                fmul(&inv, &inv, &zk[i], &FR);
                zk[i] = t;
            }
            free(pre);
        }
        rpoly DZk = {(fe *)malloc(DZ.len * sizeof(fe)), DZ.len};
        memcpy(DZk.c, DZ.c, DZ.len * sizeof(fe));
        rp_scale_powers(DZk.c, DZk.len, &k);
        rp_evals_larger(&DZk, log_col, log_row, cm, ev);
        free(DZk.c);
        for (size_t i = 0; i < big; i++) fmul(&ev[i], &ev[i], &zk[i], &FR);
        rpoly D = rp_from_evals(ev, log_col, log_row, cm);
        rp_scale_powers(D.c, D.len, &kinv);
        if (D.len > n + 1) { free(D.c); rc = 7; break; }
        /* RawBlob::new(coeffs_to_evals(data_coeffs)) then RawData::try_from(RawBlob)
         * (raw_blob.rs:43-57, raw_data.rs:14-38) */
        fe *blob = (fe *)malloc(n * sizeof(fe));
        rp_evals_coset(&D, log_col, log_row, 0, blob);
        free(D.c);
        rc = 0;
        for (size_t i = 0; i < n && !rc; i++) {
            uint64_t v[4];
            ffrom_mont(v, &blob[i], &FR);
            uint8_t b32[32];
            memcpy(b32, v, 32);
            if (b32[31]) rc = 3;
            memcpy(out_raw + 31 * i, b32, 31);
        }
        free(blob);
        if (!rc) {
            size_t max_raw = 31 * n - 4;
            uint32_t l32;
            memcpy(&l32, out_raw + max_raw, 4);
            if (l32 > max_raw) rc = 4;
            for (size_t i = l32; i < max_raw && !rc; i++)
                if (out_raw[i]) rc = 5;
            *out_len = l32;
        }
    }
    free(zk);
    free(ev);
    free(DZ.c);
    free(Z.c);
    return rc;
}

/* ------------------------------------------------------------------ */
/* deferred verification MSM: DeferredVerifier::record_msm + fast_check_msm */
/* (crates/amt/src/deferred_verification.rs:71-92, 133-150) over rows verified by */
/* BlobRow::verify (crates/amt/src/blob/encode.rs:224-239) -> verify_amt_proof   */
/* (crates/amt/src/verify_params.rs:176-186)                                   */
/* ------------------------------------------------------------------ */

/* out = sum_k alpha_k * sum_i basis_c[bitrev(l) * COL + i] * index_reverse(row_k)[i],
 * row k = (c, l) = divmod(ids[k], ROW); basis_c = line h of bases[c] (param layout).
 * rows: canonical 32 B elements, COL per row; alphas canonical.  Returns 0, or -1 for
 * an id past the encoded blob. */
int zgo_verify_msm(const uint64_t *const *bases, int log_col, int log_row, int coset_n, const uint32_t *ids,
                   const uint8_t *rows, const uint8_t *alphas, size_t n_rows, uint8_t out[64]) {
    const size_t ROW = (size_t)1 << log_row, COL = (size_t)1 << log_col, n = ROW * COL;
    for (size_t k = 0; k < n_rows; k++)
        if (ids[k] >= ROW * (size_t)coset_n) return -1;
    uint64_t *pts = (uint64_t *)malloc((n_rows * COL + 1) * 64);
    uint64_t *sc = (uint64_t *)malloc((n_rows * COL + 1) * 32);
    for (size_t k = 0; k < n_rows; k++) {
        const size_t c = ids[k] / ROW, l = ids[k] % ROW, b = bitrev(l, log_row);
        fe a;
        uint64_t av[4];
        memcpy(av, alphas + 32 * k, 32);
        fto_mont(&a, av, &FR);
        for (size_t i = 0; i < COL; i++) {
            /* data = index_reverse(row): data[i] = row[bitrev(i)] */
            uint64_t v[4];
            memcpy(v, rows + 32 * (k * COL + bitrev(i, log_col)), 32);
            fe m;
            fto_mont(&m, v, &FR);
            fmul(&m, &m, &a, &FR);
            ffrom_mont(sc + 4 * (k * COL + i), &m, &FR);
            memcpy(pts + 8 * (k * COL + i), bases[c] + 8 * ((size_t)log_row * n + b * COL + i), 64);
        }
    }
    g1j r;
    par_msm(&r, pts, sc, n_rows * COL);
    j_store_canon(out, &r);
    free(pts);
    free(sc);
    return 0;
}

/* ------------------------------------------------------------------ */
/* power_tau (crates/amt/src/power_tau.rs:28-42): out[i] = [tau^(start + i)] G1,      */
/* i < 2^log_n, canonical affine (64 B)                                             */
/* ------------------------------------------------------------------ */
void zgo_power_tau(const uint64_t tau[4], uint64_t start, int log_n, uint8_t *out) {
    const size_t n = (size_t)1 << log_n;
    fe t;
    fto_mont(&t, tau, &FR);
#pragma omp parallel num_threads(g_threads)
    {
        const int nt = omp_get_num_threads(), id = omp_get_thread_num();
        const size_t lo = n * (size_t)id / (size_t)nt, hi = n * (size_t)(id + 1) / (size_t)nt;
        fe k;
        fpow_u64(&k, &t, start + lo, &FR);
        for (size_t i = lo; i < hi; i++) {
            uint64_t kc[4];
            ffrom_mont(kc, &k, &FR);
            zgo_g1_mul_gen(kc, out + 64 * i);
            fmul(&k, &k, &t, &FR);
        }
    }
}
