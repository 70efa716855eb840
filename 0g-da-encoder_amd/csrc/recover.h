// Device kernels of RS recovery (recover_from_da_slice, crates/recovery/src/lib.rs:15-43).
// Included once, by zgr.hip.
//
// The reference decodes through a vanishing polynomial over the 4N domain
// (crates/recovery/src/data_poly.rs:72-118).  This implementation decodes the same
// code row by row instead.  Row l of coset c holds D(s zeta^j), j < COL, where D is
// the data polynomial (degree < N), s = cf_c w_N^l and zeta = w_N^ROW = w_COL.
// Writing D(x) = sum_{i<COL} x^i D_i(x^COL) (each D_i of degree < ROW):
//   iDFT_COL of the row = s^i D_i(u), u = s^COL,
// so each row gives one value of every D_i at its own point u.  Any ROW rows fix the
// D_i by interpolation; the erased coset-0 rows (points x_l = w_ROW^l) come out of one
// Lagrange matrix product, and a forward DFT_COL per row turns them back into data.
// Extra rows are checked against the interpolant through a random combination of
// their D_i values (Schwartz-Zippel) and through their D_0 alone: rows on a polynomial
// of degree exactly N, which the reference's degree assert admits (data_poly.rs:112),
// differ from the interpolant only in D_0, by one constant times A(u), and that constant
// is folded onto the coset-0 values as coeffs_to_evals does; anything else is the case
// where the assert fires.
#pragma once
#include "field.h"
#include "field29.h"
#include "fr29.h"

namespace zgr {

// w_N^e for any e, from the half table tw[k] = w_N^k (k < N/2): w_N^(N/2) = -1
__device__ __forceinline__ fr rec_wpow(const uint32_t* tw_half, uint64_t e, int log_n) {
    const uint64_t n = 1ull << log_n, h = n >> 1;
    e &= n - 1;
    fr v = f_load<FrP>(tw_half + 8 * (size_t)(e & (h - 1)));
    return e >= h ? f_neg(v) : v;
}

// canonical 32 B LE < r ?
__device__ __forceinline__ bool rec_canonical(const fr& a) {
    for (int j = 7; j >= 0; j--) {
        if (a.v[j] != FrP::M[j]) return a.v[j] < FrP::M[j];
    }
    return false;
}

// Per-call constants, in two steps:
//   k_rec_inverses (one thread each): inv[c] = cf_c^-1 (c < coset_n), inv[coset_n] = 1 / COL,
//                                     inv[coset_n + 1] = gamma (gamma arrives canonical)
//   k_rec_consts: cpow[c * COL + i] = cf_c^-i / COL (row iDFT scaling), gpow[i] = gamma^i
//                 (random combination of a row's D_i); one thread per entry.
// No Fermat inversion: cf_c is a power of the 2^28-th root, so cf_c^-1 = cf_c^(2^28 - 1),
// and 1 / 2^k = r - (r - 1) / 2^k because 2^28 divides r - 1.
__global__ void k_rec_inverses(const uint32_t* consts, const uint32_t* gamma_canon, int coset_n, int log_col,
                               uint32_t* inv) {
    const int t = threadIdx.x;
    if (blockIdx.x) return;
    if (t < coset_n) {
        f_store(inv + 8 * t, f_pow_u64(f_load<FrP>(consts + 8 * (K_CF + t)), (1ull << 28) - 1));
    } else if (t == coset_n) {
        uint32_t q[8], d[8];
        for (int j = 0; j < 8; j++) {  // (r - 1) >> log_col
            const uint32_t lo = j == 0 ? FrP::M[0] - 1 : FrP::M[j];
            const uint32_t hi = j + 1 < 8 ? FrP::M[j + 1] : 0;
            q[j] = log_col ? (lo >> log_col) | (hi << (32 - log_col)) : lo;
        }
        uint32_t br = 0;
        for (int j = 0; j < 8; j++) {
            const uint64_t x = (uint64_t)FrP::M[j] - q[j] - br;
            d[j] = (uint32_t)x;
            br = (uint32_t)(x >> 63);
        }
        fr c;
        for (int j = 0; j < 8; j++) c.v[j] = d[j];
        f_store(inv + 8 * coset_n, f_to_mont(c));
    } else if (t == coset_n + 1) {
        f_store(inv + 8 * (coset_n + 1), f_to_mont(f_load<FrP>(gamma_canon)));
    }
}

__global__ void k_rec_consts(const uint32_t* inv, int coset_n, int log_col, uint32_t* cpow, uint32_t* gpow) {
    const uint32_t col = 1u << log_col;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (uint32_t)coset_n * col) {
        const uint32_t c = t >> log_col, i = t & (col - 1);
        f_store(cpow + 8 * (size_t)t, f_mul(f_pow_u64(f_load<FrP>(inv + 8 * c), i), f_load<FrP>(inv + 8 * coset_n)));
    } else if (t < (uint32_t)(coset_n + 1) * col) {
        const uint32_t i = t - (uint32_t)coset_n * col;
        f_store(gpow + 8 * (size_t)i, f_pow_u64(f_load<FrP>(inv + 8 * (coset_n + 1)), i));
    }
}

__global__ void k_rec_clear(int* flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

// raw_slice_to_line (crates/recovery/src/utils.rs:26-34) accepts only canonical
// elements: flag bit 0 when any of the n 32 B elements is >= r.  One thread each.
__global__ void k_rec_canon(const uint8_t* rows, uint64_t n, int* flag) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (!rec_canonical(f_load<FrP>((const uint32_t*)(rows + 32 * e)))) atomicOr(flag, 1);
}

// Interpolation point u = (cf_c w_N^l)^COL of each received row (row id = c * ROW + l)
__global__ void k_rec_points(const uint32_t* ids, uint32_t nrows, const uint32_t* consts, const uint32_t* tw,
                             int log_col, int log_row, uint32_t* u) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int log_n = log_col + log_row;
    const uint32_t id = ids[k], c = id >> log_row, l = id & ((1u << log_row) - 1);
    fr cf = f_load<FrP>(consts + 8 * (K_CF + c));
    for (int b = 0; b < log_col; b++) cf = f_sqr(cf);
    f_store(u + 8 * (size_t)k, f_mul(cf, rec_wpow(tw, (uint64_t)l << log_col, log_n)));
}

// x_l = w_N^(l COL) for the erased coset-0 rows l = erased[t]
__global__ void k_rec_targets(const uint32_t* erased, uint32_t ne, const uint32_t* tw, int log_col, int log_row,
                              uint32_t* x) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne) return;
    f_store(x + 8 * (size_t)t, rec_wpow(tw, (uint64_t)erased[t] << log_col, log_col + log_row));
}

// LDS radix-2 DIT over COL elements held in bit-reversed order (natural out),
// twiddles tw[j << (log_n - s)] = w_COL^(j 2^(log_col - s)) from the N-domain table
__device__ __forceinline__ void rec_lds_dit(fr* a, const uint32_t* tw, int log_col, int log_n) {
    const uint32_t half_col = 1u << (log_col - 1);
    for (int s = 1; s <= log_col; s++) {
        const uint32_t half = 1u << (s - 1);
        for (uint32_t t = threadIdx.x; t < half_col; t += blockDim.x) {
            const uint32_t j = t & (half - 1);
            const uint32_t i0 = ((t >> (s - 1)) << s) | j, i1 = i0 + half;
            fr v = f_mul(a[i1], f_load<FrP>(tw + 8 * (size_t)(j << (log_n - s))));
            fr u = a[i0];
            a[i0] = f_add(u, v);
            a[i1] = f_sub(u, v);
        }
        __syncthreads();
    }
}

// One workgroup per received row k: canonical check, iDFT_COL, scaling to the row's
// D_i values V[k][i] = cf_c^-i w_N^(-l i) iDFT(row)_i, and the row's random
// combination yv[k] = sum_i gamma^i V[k][i].  The elements are canonical (k_rec_canon).
__global__ __launch_bounds__(256) void k_rec_rows(const uint8_t* rows, const uint32_t* ids, const uint32_t* cpow,
                                                  const uint32_t* gpow, const uint32_t* tw_inv, int log_col,
                                                  int log_row, uint32_t* V, uint32_t* yv) {
    extern __shared__ fr rec_lds[];  // COL elements, then blockDim partial sums
    fr* lds = rec_lds;
    const uint32_t k = blockIdx.x, col = 1u << log_col;
    const int log_n = log_col + log_row;
    const uint32_t id = ids[k], c = id >> log_row, l = id & ((1u << log_row) - 1);
    const uint32_t* src = (const uint32_t*)(rows + (size_t)k * col * 32);
    for (uint32_t j = threadIdx.x; j < col; j += blockDim.x)
        lds[bitrev_bits(j, log_col)] = f_to_mont(f_load<FrP>(src + 8 * (size_t)j));
    __syncthreads();
    rec_lds_dit(lds, tw_inv, log_col, log_n);
    fr part = f_zero<FrP>();
    for (uint32_t i = threadIdx.x; i < col; i += blockDim.x) {
        fr v = f_mul(lds[i], f_load<FrP>(cpow + 8 * ((size_t)c * col + i)));
        v = f_mul(v, rec_wpow(tw_inv, (uint64_t)l * i, log_n));
        f_store(V + 8 * ((size_t)k * col + i), v);
        part = f_add(part, f_mul(v, f_load<FrP>(gpow + 8 * (size_t)i)));
    }
    fr* red = lds + col;
    red[threadIdx.x] = part;
    __syncthreads();
    for (uint32_t o = blockDim.x >> 1; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = f_add(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) f_store(yv + 8 * (size_t)k, red[0]);
}

// One workgroup (256 threads) per point a = a_pts[blockIdx.x]:
//   out = prod_{q < m, q != skip} (a - u_q), inverted when invert != 0,
// with skip = blockIdx.x when self_skip (B_r = prod_{q != r} (u_r - u_q) over the
// interpolation set, the inverse barycentric weights), else no skip (A(x) = prod_q (x - u_q)).
// Strided partial products, then a tree in LDS.
__global__ __launch_bounds__(256) void k_rec_prod(const uint32_t* a_pts, const uint32_t* u, uint32_t m, int self_skip,
                                                  int invert, uint32_t* out) {
    __shared__ fr part[256];
    const uint32_t b = blockIdx.x;
    const fr a = f_load<FrP>(a_pts + 8 * (size_t)b);
    fr p = f_one<FrP>();
    for (uint32_t q = threadIdx.x; q < m; q += blockDim.x)
        if (!(self_skip && q == b)) p = f_mul(p, f_sub(a, f_load<FrP>(u + 8 * (size_t)q)));
    part[threadIdx.x] = p;
    __syncthreads();
    for (uint32_t o = blockDim.x >> 1; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] = f_mul(part[threadIdx.x], part[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) f_store(out + 8 * (size_t)b, invert ? f_inv(part[0]) : part[0]);
}

constexpr uint32_t REC_CHUNK = 64;  // entries of one batched inversion in k_rec_lagrange

// out[i] (9 u32, 2^261 form) from in[i] (Fr, 2^256 Montgomery), i < n
__global__ void k_rec_to29(const uint32_t* in, uint32_t n, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fq29 v = r29_mul(fr29_unpack(in + 8 * (size_t)i), q_const(R29M::C_TO));
#pragma unroll
    for (int k = 0; k < 9; k++) out[9 * (size_t)i + k] = v.l[k];
}

__device__ __forceinline__ fq29 rec_load29(const uint32_t* p) {
    fq29 v;
#pragma unroll
    for (int k = 0; k < 9; k++) v.l[k] = p[k];
    return v;
}

// L[t][r] = 1 / ((x_t - u_r) B_r), B_r = prod_{q != r} (u_r - u_q): the Lagrange matrix
// without its row factor A(x_t), b_r = 1 / B_r folded into the same inversion.  One
// thread per (target t, chunk of REC_CHUNK sources), Montgomery's trick over the chunk
// in the 29-bit form (x, u, B come converted by k_rec_to29), the prefix products parked
// in L itself; L is written in the 2^256 Montgomery form, canonical.
__global__ void k_rec_lagrange(const uint32_t* t29, uint32_t nt, const uint32_t* u29, const uint32_t* B29,
                               uint32_t m, uint32_t* L) {
    const uint32_t chunks = (m + REC_CHUNK - 1) / REC_CHUNK;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nt * chunks) return;
    const uint32_t t = g / chunks, r0 = (g % chunks) * REC_CHUNK;
    const uint32_t r1 = min(r0 + REC_CHUNK, m);
    const fq29 x = rec_load29(t29 + 9 * (size_t)t);
    uint32_t* row = L + 8 * (size_t)t * m;
    fq29 acc = q_const(R29M::ONE);
    for (uint32_t r = r0; r < r1; r++) {
        fr29_pack(acc, row + 8 * (size_t)r);  // prefix prod_{r0 <= q < r} d_q (< 2r < 2^255)
        const fq29 d = r29_mul(r29_sub_lazy(x, rec_load29(u29 + 9 * (size_t)r)), rec_load29(B29 + 9 * (size_t)r));
        acc = r29_mul(acc, d);
    }
    fq29 inv = r29_inv(acc);
    const fq29 c_from = q_const(R29M::C_FROM);
    for (uint32_t r = r1; r-- > r0;) {
        const fq29 d = r29_mul(r29_sub_lazy(x, rec_load29(u29 + 9 * (size_t)r)), rec_load29(B29 + 9 * (size_t)r));
        fq29 v = r29_mul(r29_mul(inv, fr29_unpack(row + 8 * (size_t)r)), c_from);  // 2^256 form, < 2r
        fq29 w;  // v - r, kept when it does not borrow
        int32_t br = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const int32_t y = (int32_t)v.l[k] - (int32_t)R29::P[k] - br;
            br = (y >> 31) & 1;
            w.l[k] = (uint32_t)y & R29_MASK;
        }
#pragma unroll
        for (int k = 0; k < 9; k++) v.l[k] = br ? v.l[k] : w.l[k];
        fr29_pack(v, row + 8 * (size_t)r);
        inv = r29_mul(inv, d);
    }
}

constexpr int REC_TL = 4;   // GEMM tile: targets (one per wave)
constexpr int REC_TI = 64;  // GEMM tile: D_i columns (one per lane)
constexpr int REC_TK = 14;  // GEMM tile: sources per LDS stage (two carry passes)

// Out[t][i] = sum_{r < m} L[t][r] V[r][i] for t < nt, i < col (Montgomery form in and
// out, values < r): 256 threads, a 4 x 64 output tile per workgroup, one output per
// thread.  The products accumulate unreduced in 18 64-bit columns; with L, V < r and
// m <= 2^10 the sum is below 2^518, the single Montgomery reduction by 2^261 leaves it
// below 2^257, and six conditional subtractions of r 2^k bring it below r.  That
// reduction divides by 2^261 where the Montgomery product divides by 2^256, so the
// result is doubled five times.
__global__ __launch_bounds__(256) void k_rec_gemm(const uint32_t* L, const uint32_t* V, uint32_t nt, uint32_t m,
                                                  uint32_t col, uint32_t* Out) {
    __shared__ uint32_t Ls[REC_TL][REC_TK][9];
    __shared__ uint32_t Vs[REC_TK][REC_TI][9];
    const uint32_t tx = threadIdx.x % REC_TI, ty = threadIdx.x / REC_TI;
    const uint32_t t0 = blockIdx.y * REC_TL, i0 = blockIdx.x * REC_TI;
    uint64_t T[18];
#pragma unroll
    for (int k = 0; k < 18; k++) T[k] = 0;
    for (uint32_t k0 = 0; k0 < m; k0 += REC_TK) {
        for (uint32_t e = threadIdx.x; e < REC_TL * REC_TK; e += blockDim.x) {
            const uint32_t a = e / REC_TK, b = e % REC_TK;
            const uint32_t t = t0 + a, r = k0 + b;
            const fr x = (t < nt && r < m) ? f_load<FrP>(L + 8 * ((size_t)t * m + r)) : f_zero<FrP>();
            fr29_limbs(x, Ls[a][b]);
        }
        for (uint32_t e = threadIdx.x; e < REC_TK * REC_TI; e += blockDim.x) {
            const uint32_t a = e / REC_TI, b = e % REC_TI;
            const uint32_t r = k0 + a, i = i0 + b;
            const fr x = (r < m && i < col) ? f_load<FrP>(V + 8 * ((size_t)r * col + i)) : f_zero<FrP>();
            fr29_limbs(x, Vs[a][b]);
        }
        __syncthreads();
#pragma unroll 1
        for (int half = 0; half < 2; half++) {
            for (int kk = 7 * half; kk < 7 * half + 7; kk++) {
                uint32_t l[9], v[9];
#pragma unroll
                for (int j = 0; j < 9; j++) {
                    l[j] = Ls[ty][kk][j];
                    v[j] = Vs[kk][tx][j];
                }
#pragma unroll
                for (int i = 0; i < 9; i++)
#pragma unroll
                    for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)l[i] * v[j];
            }
#pragma unroll
            for (int k = 0; k < 17; k++) {  // carry pass: same value, columns back below 2^29
                T[k + 1] += T[k] >> 29;
                T[k] &= R29_MASK;
            }
        }
        __syncthreads();
    }
    // Montgomery reduction by 2^261
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t mm = ((uint32_t)T[i] * R29::INV) & R29_MASK;
#pragma unroll
        for (int j = 0; j < 9; j++) T[i + j] += (uint64_t)mm * R29::P[j];
        T[i + 1] += T[i] >> 29;
    }
    uint32_t v[9];
    uint64_t cy = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const uint64_t x = T[9 + k] + cy;
        v[k] = (uint32_t)x & R29_MASK;
        cy = x >> 29;
    }
    v[8] |= (uint32_t)cy << 29;  // below 2^259: the top limb keeps its extra bits
#pragma unroll
    for (int q = 0; q < 6; q++) {  // v -= r 2^(5-q) where that stays non-negative
        uint32_t d[9];
        int32_t br = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const int64_t x = (int64_t)v[k] - (int64_t)R29::RK[q][k] - br;
            br = x < 0;
            d[k] = k < 8 ? (uint32_t)x & R29_MASK : (uint32_t)x;
        }
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = br ? v[k] : d[k];
    }
    fr o;
#pragma unroll
    for (int w = 0; w < 8; w++) {  // 29-bit limbs -> 32-bit words
        const int bit = 32 * w, k = bit / 29, sh = bit % 29;
        uint64_t x = (uint64_t)v[k] >> sh;
        if (k + 1 < 9) x |= (uint64_t)v[k + 1] << (29 - sh);
        if (k + 2 < 9 && 29 - sh + 29 < 32) x |= (uint64_t)v[k + 2] << (58 - sh);
        o.v[w] = (uint32_t)x;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) o = f_add(o, o);
    const uint32_t i = i0 + tx, t = t0 + ty;
    if (i < col && t < nt) f_store(Out + 8 * ((size_t)t * col + i), o);
}

// ---------------------------------------------------------------------------
// The same GEMM on the matrix cores (v_mfma_i32_32x32x32_i8).  Every element, a
// Montgomery residue below r < 2^254, is written with 32 balanced base-256 digits
// (-128..127, int8; the top digit stays below 64, so no carry leaves it), so
// Out[t][i] * 2^256 = sum_s D_s[t][i] 2^(8 s) with, for the 63 digit-weight diagonals
// s = a + b,
//     D_s[t][i] = sum_{a + b = s} sum_r Lp_a[t][r] Vp_b[r][i]    (|D_s| <= 32 * 2^10 * 2^14 = 2^29)
// a sum of int8 matrix products.  k_rec_gemm_mfma computes the D_s of a 32 x 32 output
// tile, k_rec_fold folds them: sum_s |D_s| (+-2^(8 s + 5) mod r) in 29-bit columns, one
// Montgomery reduction by 2^261 (= X 2^-256 mod r, the Montgomery form of the product
// sum), one conditional subtraction of r.
// MFMA operand layout (32x32x32, int8): lane l holds A[l % 32][16 (l / 32) .. + 16) and
// B[16 (l / 32) .. + 16)[l % 32] as 16 packed bytes; accumulator register q of lane l is
// C[8 (q / 4) + 4 (l / 32) + q % 4][l % 32].
constexpr int RM_LIMBS = 32;
constexpr int RM_DIAG = 2 * RM_LIMBS - 1;
constexpr int RM_WAVES = 16;                              // k_rec_gemm_mfma: 1024 threads
constexpr int RM_ACC = (RM_DIAG + RM_WAVES - 1) / RM_WAVES;  // diagonals per wave (s = w + 16 q)
constexpr uint32_t RM_MAX_M = 2048;                       // keeps D_s below 2^31
static_assert(2 * RM_LIMBS == 4 * RM_WAVES, "rm_stage: four slabs per wave cover both operands' digits");

typedef int rm_v16i __attribute__((ext_vector_type(16)));
typedef int rm_v4i __attribute__((ext_vector_type(4)));

// {2^(8 s + 5) mod r, r - that}, 29-bit limbs
__constant__ uint32_t RM_KS[RM_DIAG][2][9] = {
    {{0x00000020u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x0fffffe1u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00002000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x0fffe001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00200000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x0fe00001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000001u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9eu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000100u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fab9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00010000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0eac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x01000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1e0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000008u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2448u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000800u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c1c50u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00080000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e542450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x08000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x065c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000040u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090b3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00004000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d050f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00400000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x079090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000002u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d281u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000200u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d083u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00020000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1583d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x02000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1385d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000010u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40b0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00001000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db30c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00100000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02cb40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x10000000u, 0x00000000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x12db40c0u, 0x00a6e140u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000080u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e0c1u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00008000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a66141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00800000u, 0x00000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x0026e141u, 0x0e5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000004u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2630u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000400u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2234u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00040000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e582634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x04000000u, 0x00000000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0a5c2634u, 0x0030644eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000020u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030642eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00002000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030444eu}},
    {{0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00200000u}, {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0010644eu}},
    {{0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu, 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u}, {0x000000aau, 0x0068a1ebu, 0x09301dc5u, 0x108041aau, 0x0addcb27u, 0x0598fff2u, 0x0ed19539u, 0x09315e8bu, 0x00229c18u}},
    {{0x1fff56b8u, 0x0aa227a3u, 0x021c552fu, 0x1da684b1u, 0x1a6426d0u, 0x14969824u, 0x065cad9bu, 0x00dce9edu, 0x002c0019u}, {0x1000a949u, 0x146d84fbu, 0x0c3fcf21u, 0x0a2a0c42u, 0x1b21abb2u, 0x0e44a89bu, 0x1a4a33a5u, 0x0d7f3c46u, 0x00046435u}},
    {{0x1f56b718u, 0x1bf33373u, 0x18d445f3u, 0x11815470u, 0x02e009fcu, 0x1fe57637u, 0x057178a7u, 0x19674e0eu, 0x002531eeu}, {0x10a948e9u, 0x031c792bu, 0x1587de5du, 0x164f3c82u, 0x12a5c886u, 0x02f5ca89u, 0x1b356899u, 0x14f4d825u, 0x000b325fu}},
    {{0x16b7173cu, 0x0b3349dcu, 0x15ba25e1u, 0x05a57662u, 0x0594d010u, 0x1598a393u, 0x11b43229u, 0x08c0ce57u, 0x002522bbu}, {0x1948e8c5u, 0x13dc62c2u, 0x18a1fe6fu, 0x022b1a90u, 0x0ff10273u, 0x0d429d2du, 0x0ef2af17u, 0x059b57dcu, 0x000b4193u}},
    {{0x17173b3cu, 0x0b49b297u, 0x1b9a135bu, 0x09c76849u, 0x1a5ae3b1u, 0x08c5ffa8u, 0x146db3d7u, 0x024117b9u, 0x0015ef36u}, {0x18e8c4c5u, 0x13c5fa07u, 0x12c210f5u, 0x1e0928a9u, 0x1b2aeed1u, 0x1a154117u, 0x0c392d69u, 0x0c1b0e7au, 0x001a7518u}},
    {{0x173b3b8cu, 0x16985f72u, 0x1852e6a9u, 0x1ce69b8cu, 0x1a404dd5u, 0x1aa65184u, 0x0215c5c7u, 0x1f566a11u, 0x0001c285u}, {0x18c4c475u, 0x08774d2cu, 0x16093da7u, 0x0ae9f566u, 0x1b4584adu, 0x0834ef3bu, 0x1e911b79u, 0x0f05bc22u, 0x002ea1c8u}},
    {{0x0b3b8bf7u, 0x00d2611eu, 0x11a962dcu, 0x00467433u, 0x1e996f4au, 0x0c9c3e0bu, 0x0fe7dc8bu, 0x152cb93cu, 0x000eff38u}, {0x04c4740au, 0x1e3d4b81u, 0x1cb2c174u, 0x078a1cbfu, 0x16ec6339u, 0x163f02b4u, 0x10bf04b5u, 0x192f6cf7u, 0x00216515u}},
    {{0x0b8bf6b1u, 0x1c8ad921u, 0x1af3a709u, 0x1d17786cu, 0x15235381u, 0x1a95107fu, 0x145d084eu, 0x1e497271u, 0x00104473u}, {0x04740950u, 0x0284d37eu, 0x13687d47u, 0x0ab91886u, 0x00627f01u, 0x08463041u, 0x0c49d8f2u, 0x1012b3c2u, 0x00201fdau}},
    {{0x0bf6b0aau, 0x1b9523c7u, 0x00b2d6b0u, 0x1767bb0fu, 0x085ec9d1u, 0x1f68beefu, 0x04f8a2f6u, 0x167d9c29u, 0x0002c197u}, {0x04094f57u, 0x037a88d8u, 0x0da94da0u, 0x1068d5e4u, 0x0d2708b1u, 0x037281d1u, 0x1bae3e4au, 0x17de8a0au, 0x002da2b6u}},
    {{0x16b0a9f2u, 0x024856a6u, 0x09ccb46fu, 0x1a5321b5u, 0x11784e8du, 0x00c164b9u, 0x0f82a56cu, 0x1492124fu, 0x001c1b69u}, {0x194f560fu, 0x1cc755f8u, 0x048f6fe1u, 0x0d7d6f3eu, 0x040d83f5u, 0x0219dc07u, 0x11243bd5u, 0x19ca13e4u, 0x001448e4u}},
    {{0x10a9f16cu, 0x1346da7fu, 0x1f6f7042u, 0x0e8de88fu, 0x06f0d9f2u, 0x1aa34a28u, 0x022b3264u, 0x04cc3969u, 0x00216c4au}, {0x1f560e95u, 0x0bc8d21fu, 0x0eecb40eu, 0x1942a863u, 0x0e94f890u, 0x0837f698u, 0x1e7baedcu, 0x098feccau, 0x000ef804u}},
    {{0x09f16b50u, 0x0c13d1ddu, 0x10174aefu, 0x0e84e89cu, 0x04d93839u, 0x0c8da3c1u, 0x18778815u, 0x0cdf254du, 0x00275437u}, {0x060e94b1u, 0x12fbdac2u, 0x1e44d961u, 0x194ba856u, 0x10ac9a49u, 0x164d9cffu, 0x082f592bu, 0x017d00e6u, 0x00091017u}},
    {{0x116b4f30u, 0x17159bb7u, 0x0c6d6d96u, 0x0b72d6b3u, 0x1c7d2ed1u, 0x1b7f249au, 0x0ff11081u, 0x1446437fu, 0x0002b7a9u}, {0x1e94b0d1u, 0x07fa10e7u, 0x01eeb6bau, 0x1c5dba40u, 0x1908a3b1u, 0x075c1c25u, 0x10b5d0bfu, 0x1a15e2b4u, 0x002daca4u}},
    {{0x0b4f2ff2u, 0x02c046d2u, 0x04639a4bu, 0x056ec613u, 0x0fdd4e2eu, 0x1727105au, 0x07f0304cu, 0x1d3968a7u, 0x00122d57u}, {0x04b0d00fu, 0x1c4f65cdu, 0x09f88a05u, 0x0261cae0u, 0x05a88455u, 0x0bb43066u, 0x18b6b0f4u, 0x1122bd8cu, 0x001e36f6u}},
    {{0x0f2ff1a0u, 0x1a66168au, 0x010cacb8u, 0x008fb7d8u, 0x0b1f3cf4u, 0x14d8123eu, 0x119bd450u, 0x16da53bdu, 0x0007ba7eu}, {0x00d00e61u, 0x04a99615u, 0x0d4f7798u, 0x0740d91bu, 0x0a66958fu, 0x0e032e82u, 0x0f0b0cf0u, 0x1781d276u, 0x0028a9cfu}},
    {{0x0ff19fd8u, 0x0ba3918du, 0x0e470c2cu, 0x172131feu, 0x02540f82u, 0x05d0203eu, 0x01c11e7bu, 0x1bedc56cu, 0x002ad274u}, {0x000e6029u, 0x136c1b12u, 0x00151824u, 0x10af5ef5u, 0x1331c300u, 0x1d0b2082u, 0x1ee5c2c5u, 0x126e60c7u, 0x000591d9u}},
    {{0x119fd71eu, 0x17bb28b0u, 0x19b41ce1u, 0x1b120786u, 0x13ebaadbu, 0x0a9113fau, 0x0dcb9fb8u, 0x006bb221u, 0x0019e79eu}, {0x1e6028e3u, 0x075483eeu, 0x14a8076fu, 0x0cbe896cu, 0x019a27a7u, 0x184a2cc6u, 0x12db4188u, 0x0df07412u, 0x00167cb0u}},
    {{0x0fd71d77u, 0x1bc54f31u, 0x04cd7268u, 0x0369f485u, 0x070d339cu, 0x09be5383u, 0x12512c7fu, 0x1c61af97u, 0x0001f007u}, {0x0028e28au, 0x034a5d6eu, 0x098eb1e8u, 0x04669c6eu, 0x0e789ee7u, 0x191ced3du, 0x0e55b4c1u, 0x11fa769cu, 0x002e7446u}},
    {{0x171d76f6u, 0x0eb27343u, 0x1dd8fdb4u, 0x1bcedba3u, 0x15f962fau, 0x01c2fbb1u, 0x0aa7b2c3u, 0x1216198au, 0x000c1cd2u}, {0x18e2890bu, 0x105d395bu, 0x1083269cu, 0x0c01b54fu, 0x1f8c6f88u, 0x0118450eu, 0x15ff2e7eu, 0x1c460ca9u, 0x0024477bu}},
    {{0x1d76f5c0u, 0x0e881bd8u, 0x01f4a037u, 0x1ab76712u, 0x17ee5a0eu, 0x0c2b8184u, 0x1dfa72c8u, 0x1f0ffd53u, 0x0003bef3u}, {0x12890a41u, 0x108790c6u, 0x0c678419u, 0x0d1929e1u, 0x1d977874u, 0x16afbf3bu, 0x02ac6e78u, 0x0f4c28e0u, 0x002ca55au}},
    {{0x06f5bfedu, 0x19f20915u, 0x03c98571u, 0x02ec4ffeu, 0x156b6f18u, 0x153bb672u, 0x0e10108cu, 0x1f267e13u, 0x00278225u}, {0x090a4014u, 0x051da38au, 0x0a929edfu, 0x04e440f5u, 0x001a636bu, 0x0d9f8a4eu, 0x1296d0b4u, 0x0f35a820u, 0x0008e228u}},
    {{0x05bfec2fu, 0x163d2700u, 0x104bcbb4u, 0x0b09a75du, 0x192e3af1u, 0x06b6955eu, 0x07d2a686u, 0x0d42e2f8u, 0x000041edu}, {0x0a4013d2u, 0x08d2859fu, 0x1e10589cu, 0x1cc6e995u, 0x1c579791u, 0x1c24ab61u, 0x18d43abau, 0x0119433bu, 0x00302261u}},
    {{0x0fec2effu, 0x1e17538eu, 0x1d6f9060u, 0x01d6cc8eu, 0x18b51ed5u, 0x13ba1e08u, 0x11ffa4f4u, 0x1486d20au, 0x0011891bu}, {0x0013d102u, 0x00f85911u, 0x10ec93f0u, 0x05f9c464u, 0x1cd0b3aeu, 0x0f2122b7u, 0x0ea73c4cu, 0x19d55429u, 0x001edb32u}},
    {{0x0c2efea4u, 0x0db1852du, 0x067353d7u, 0x07d8776eu, 0x19072de4u, 0x1352c387u, 0x03ac0139u, 0x1db44fdeu, 0x00250f72u}, {0x03d1015du, 0x115e2772u, 0x07e8d079u, 0x1ff81985u, 0x1c7ea49eu, 0x0f887d38u, 0x1cfae007u, 0x10a7d655u, 0x000b54dbu}},
    {{0x0efea33cu, 0x09850343u, 0x14c8096fu, 0x1cc873cfu, 0x0cb8b7c2u, 0x02e5f444u, 0x0c3cc3c5u, 0x15c29e49u, 0x0002a6ddu}, {0x01015cc5u, 0x158aa95cu, 0x19941ae1u, 0x0b081d23u, 0x08cd1ac0u, 0x1ff54c7cu, 0x146a1d7bu, 0x189987eau, 0x002dbd70u}},
    {{0x1ea33bf2u, 0x1227d2beu, 0x1eff72deu, 0x1b0be255u, 0x0b663fb8u, 0x1df6b9dcu, 0x13a37387u, 0x19943289u, 0x00016163u}, {0x115cc40fu, 0x0ce7d9e0u, 0x0f5cb172u, 0x0cc4ae9du, 0x0a1f92cau, 0x04e486e4u, 0x0d036db9u, 0x14c7f3aau, 0x002f02eau}},
    {{0x133bf1f9u, 0x0e650698u, 0x1aede05au, 0x152e5f4fu, 0x0f96f741u, 0x02bb1716u, 0x1ee35f28u, 0x0fad7e30u, 0x000ea5a7u}, {0x1cc40e08u, 0x10aaa606u, 0x136e43f6u, 0x12a231a3u, 0x05eedb41u, 0x002029aau, 0x01c38219u, 0x1eaea803u, 0x0021bea6u}},
    {{0x0bf1f8b3u, 0x0d4faca0u, 0x1c296e18u, 0x14a3b69du, 0x1db6f02fu, 0x1f249c88u, 0x112d6781u, 0x1bc6b351u, 0x00177be4u}, {0x040e074eu, 0x11bfffffu, 0x1232b638u, 0x132cda55u, 0x17cee253u, 0x03b6a437u, 0x0f7979bfu, 0x129572e2u, 0x0018e869u}},
    {{0x11f8b284u, 0x0415031du, 0x14cc8132u, 0x1ab067f5u, 0x0a1e3812u, 0x02692b9au, 0x1c926672u, 0x1210d056u, 0x000b4edeu}, {0x1e074d7du, 0x1afaa981u, 0x198fa31eu, 0x0d2028fdu, 0x0b679a70u, 0x00721526u, 0x04147acfu, 0x1c4b55ddu, 0x0025156fu}},
    {{0x08b283c5u, 0x0c6654cdu, 0x1d44d377u, 0x03568d8au, 0x08608e96u, 0x00a3ade9u, 0x0bf08813u, 0x019388e7u, 0x0027c07cu}, {0x074d7c3cu, 0x12a957d2u, 0x111750d9u, 0x047a0368u, 0x0d2543edu, 0x023792d7u, 0x14b6592eu, 0x0cc89d4cu, 0x0008a3d2u}},
    {{0x1283c42eu, 0x0b79326eu, 0x1d3dacf7u, 0x0d76a335u, 0x18c7e671u, 0x0bd2cb35u, 0x07a34ba0u, 0x0bf190b3u, 0x000e33b2u}, {0x1d7c3bd3u, 0x13967a30u, 0x111e7759u, 0x1a59edbdu, 0x1cbdec11u, 0x1708758au, 0x190395a0u, 0x026a9580u, 0x0022309cu}}};

// the 32 balanced base-256 digits of v < 2^254 (little-endian words)
__device__ __forceinline__ void rm_digits(const fr& v, int8_t* d) {
    uint32_t c = 0;
#pragma unroll
    for (int a = 0; a < RM_LIMBS; a++) {
        const uint32_t x = ((v.v[a >> 2] >> (8 * (a & 3))) & 0xffu) + c;
        c = x >= 128u;
        d[a] = (int8_t)(int32_t)(x - 256u * c);
    }
}

// Limb planes in slabs: plane a of an x-by-mp operand (x = tp or cp, mp: multiples of
// 32) stores, for row tile x / 32 and K step k / 32, the 32 x 32 bytes [row % 32][k % 32]
// as one contiguous 1 KB slab, so that staging a slab is fully coalesced.
__device__ __forceinline__ size_t rm_slab_index(int a, uint32_t x_tiles, uint32_t k_steps, uint32_t row,
                                                uint32_t k) {
    return ((((size_t)a * x_tiles + (row >> 5)) * k_steps + (k >> 5)) << 10) + ((row & 31) << 5) + (k & 31);
}

// Lp = digit planes of L[t][r] (t < nt, r < m), zero in the padding up to tp x mp
__global__ void k_rec_planes_l(const uint32_t* L, uint32_t nt, uint32_t m, uint32_t tp, uint32_t mp, int8_t* Lp) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)tp * mp) return;
    const uint32_t t = (uint32_t)(e / mp), r = (uint32_t)(e % mp);
    fr v{};
    if (t < nt && r < m) v = f_load<FrP>(L + 8 * ((size_t)t * m + r));
    int8_t d[RM_LIMBS];
    rm_digits(v, d);
#pragma unroll
    for (int a = 0; a < RM_LIMBS; a++) Lp[rm_slab_index(a, tp >> 5, mp >> 5, t, r)] = d[a];
}

// Vp = digit planes of V transposed, Vp[b][i][r] = digit b of V[r][i] (r < m, i < col),
// so that a lane's 8 bytes of B are contiguous; zero in the padding up to cp x mp
__global__ void k_rec_planes_v(const uint32_t* V, uint32_t m, uint32_t col, uint32_t mp, uint32_t cp, int8_t* Vp) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)cp * mp) return;
    const uint32_t i = (uint32_t)(e % cp), r = (uint32_t)(e / cp);  // consecutive threads: consecutive i
    fr v{};
    if (i < col && r < m) v = f_load<FrP>(V + 8 * ((size_t)r * col + i));
    int8_t d[RM_LIMBS];
    rm_digits(v, d);
#pragma unroll
    for (int b = 0; b < RM_LIMBS; b++) Vp[rm_slab_index(b, cp >> 5, mp >> 5, i, r)] = d[b];
}

// a lane's 16-byte fragment of slab x
__device__ __forceinline__ rm_v4i rm_frag(const uint64_t (*S)[32][4], int x, int row, int kh) {
    const uint64_t lo = S[x][row][2 * kh], hi = S[x][row][2 * kh + 1];
    rm_v4i f = {(int)(uint32_t)lo, (int)(uint32_t)(lo >> 32), (int)(uint32_t)hi, (int)(uint32_t)(hi >> 32)};
    return f;
}

// One K step of wave W: for every digit a, its A fragment once, then an MFMA for each of
// the wave's diagonals s = W + 16 q that pairs it with a digit b = s - a.  W, a and q are
// compile-time constants after unrolling, so the wave runs exactly its MFMAs as
// straight-line code (no branch between them for the hazard checks to guard).
template <int W>
__device__ __forceinline__ void rm_step(const uint64_t (*As)[32][4], const uint64_t (*Bs)[32][4], rm_v16i* acc,
                                        int row, int kh) {
#pragma unroll
    for (int a = 0; a < RM_LIMBS; a++) {
        const rm_v4i af = rm_frag(As, a, row, kh);
#pragma unroll
        for (int q = 0; q < RM_ACC; q++) {
            const int s = W + RM_WAVES * q, b = s - a;
            if (s < RM_DIAG && b >= 0 && b < RM_LIMBS)
                acc[q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, rm_frag(Bs, b, row, kh), acc[q], 0, 0, 0);
        }
        // a compiler barrier per digit: the fragments a wave reuses are read again from LDS
        // rather than kept in registers, where they do not fit beside the accumulators
        asm volatile("" ::: "memory");
    }
}

// rm_step<w> for the runtime (wave-uniform) wave index w
template <int W>
__device__ __forceinline__ void rm_dispatch(int w, const uint64_t (*As)[32][4], const uint64_t (*Bs)[32][4],
                                            rm_v16i* acc, int row, int kh) {
    if (w == W) {
        rm_step<W>(As, Bs, acc, row, kh);
    } else {
        if constexpr (W + 1 < RM_WAVES) rm_dispatch<W + 1>(w, As, Bs, acc, row, kh);
    }
}

// Copy the K-step k0 slabs of the tile into one LDS buffer pair: 64 slabs of 1 KB, four
// per wave, each one global_load_lds_dwordx4 (lane l's 16 bytes land at base + 16 l).  The
// copies complete asynchronously (vmcnt), so a buffer fills while the other is read.
__device__ __forceinline__ void rm_stage(const int8_t* Lp, const int8_t* Vp, uint32_t tp, uint32_t cp,
                                         uint32_t k_steps, uint32_t t0, uint32_t i0, uint32_t k0, int w, int l,
                                         uint64_t (*As)[32][4], uint64_t (*Bs)[32][4]) {
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int j = w + RM_WAVES * jj;  // slab: A digit j (j < 32) or B digit j - 32
        if (j < RM_LIMBS)
            __builtin_amdgcn_global_load_lds((const void*)(Lp + rm_slab_index(j, tp >> 5, k_steps, t0, k0) + 16 * l),
                                             (void __attribute__((address_space(3)))*)&As[j][0][0], 16, 0, 0);
        else
            __builtin_amdgcn_global_load_lds(
                (const void*)(Vp + rm_slab_index(j - RM_LIMBS, cp >> 5, k_steps, i0, k0) + 16 * l),
                (void __attribute__((address_space(3)))*)&Bs[j - RM_LIMBS][0][0], 16, 0, 0);
    }
}

// D[s][t][i] for one 32 x 32 tile (t0 = 32 blockIdx.y, i0 = 32 blockIdx.x); tp, cp and
// mp multiples of 32.  Wave w accumulates the diagonals s = w + 16 q; a wave keeps its A
// fragment of digit a while it runs the b = s - a its diagonals need.  The 32 digit slabs
// of both operands of a K step (64 KB) sit in one of two LDS buffer pairs (128 KB): the
// next step's slabs are copied into the other pair while this step's MFMAs run.
__global__ __launch_bounds__(RM_WAVES * 64) void k_rec_gemm_mfma(const int8_t* Lp, const int8_t* Vp, uint32_t tp,
                                                                 uint32_t mp, uint32_t cp, int* D) {
    __shared__ uint64_t As0[RM_LIMBS][32][4], Bs0[RM_LIMBS][32][4], As1[RM_LIMBS][32][4], Bs1[RM_LIMBS][32][4];
    const uint32_t t0 = blockIdx.y * 32, i0 = blockIdx.x * 32;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int row = l & 31, kh = l >> 5;
    rm_v16i acc[RM_ACC];
#pragma unroll
    for (int q = 0; q < RM_ACC; q++)
#pragma unroll
        for (int k = 0; k < 16; k++) acc[q][k] = 0;
    const uint32_t k_steps = mp >> 5;
    rm_stage(Lp, Vp, tp, cp, k_steps, t0, i0, 0, w, l, As0, Bs0);
    for (uint32_t ks = 0; ks < k_steps; ks += 2) {  // block-uniform; k_steps >= 1
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copies into pair 0
        __syncthreads();                // everyone's, and pair 1 is no longer read
        if (ks + 1 < k_steps) rm_stage(Lp, Vp, tp, cp, k_steps, t0, i0, (ks + 1) * 32, w, l, As1, Bs1);
        rm_dispatch<0>(w, As0, Bs0, acc, row, kh);  // wave-uniform
        if (ks + 1 >= k_steps) break;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (ks + 2 < k_steps) rm_stage(Lp, Vp, tp, cp, k_steps, t0, i0, (ks + 2) * 32, w, l, As0, Bs0);
        rm_dispatch<0>(w, As1, Bs1, acc, row, kh);
    }
#pragma unroll
    for (int q = 0; q < RM_ACC; q++) {
        const int s = w + RM_WAVES * q;
        if (s >= RM_DIAG) continue;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t t = t0 + 8 * (k >> 2) + 4 * kh + (k & 3);
            D[((size_t)s * tp + t) * cp + i0 + row] = acc[q][k];
        }
    }
}

// Out[t][i] (Montgomery, < r) from the diagonals: sum_s |D_s| K_s, with K_s = 2^(8 s + 5)
// mod r for D_s >= 0 and r - that otherwise, in 29-bit columns (a carry pass every 24
// terms keeps them below 2^63), a Montgomery reduction by 2^261 (the sum is below
// 2^289 < r 2^261, so the result is below 2r), one conditional subtraction.
__global__ void k_rec_fold(const int* D, uint32_t tp, uint32_t cp, uint32_t nt, uint32_t col, uint32_t* Out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
    if (i >= col || t >= nt) return;
    uint64_t c[19];
#pragma unroll
    for (int k = 0; k < 19; k++) c[k] = 0;
    for (int s0 = 0; s0 < RM_DIAG; s0 += 24) {
        for (int s = s0; s < s0 + 24 && s < RM_DIAG; s++) {
            const int32_t ds = D[((size_t)s * tp + t) * cp + i];
            const uint64_t d = (uint32_t)(ds < 0 ? -ds : ds);
            const bool neg = ds < 0;
#pragma unroll
            for (int k = 0; k < 9; k++) c[k] += d * (neg ? RM_KS[s][1][k] : RM_KS[s][0][k]);  // uniform loads
        }
#pragma unroll
        for (int k = 0; k < 10; k++) {  // carry pass: columns 0..9 back below 2^29 (+ carry)
            c[k + 1] += c[k] >> 29;
            c[k] &= R29_MASK;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {  // REDC by 2^261
        const uint32_t mk = ((uint32_t)c[k] * R29::INV) & R29_MASK;
#pragma unroll
        for (int j = 0; j < 9; j++) c[k + j] += (uint64_t)mk * R29::P[j];
        c[k + 1] += c[k] >> 29;
    }
    uint32_t v[9];
    uint64_t carry = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const uint64_t x = c[9 + k] + carry;
        v[k] = (uint32_t)x & R29_MASK;
        carry = x >> 29;
    }
    uint32_t d[9];  // v - r, kept when it does not borrow
    int32_t br = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int32_t x = (int32_t)v[k] - (int32_t)R29::P[k] - br;
        br = (x >> 31) & 1;
        d[k] = (uint32_t)x & R29_MASK;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = br ? v[k] : d[k];
    fr o;
#pragma unroll
    for (int ww = 0; ww < 8; ww++) {  // 29-bit limbs -> 32-bit words
        const int bit = 32 * ww, k = bit / 29, sh = bit % 29;
        uint64_t x = (uint64_t)v[k] >> sh;
        if (k + 1 < 9) x |= (uint64_t)v[k + 1] << (29 - sh);
        if (k + 2 < 9 && 29 - sh + 29 < 32) x |= (uint64_t)v[k + 2] << (58 - sh);
        o.v[ww] = (uint32_t)x;
    }
    f_store(Out + 8 * ((size_t)t * col + i), o);
}

// One workgroup (256 threads) per extra row x (beyond the ROW interpolation rows): the
// residuals of its random combination and of its D_0 against the interpolant I:
//   e  = yv_x   - A(u_x) sum_r L[x][r] yv_r
//   e0 = V_x[0] - A(u_x) sum_r L[x][r] V_r[0]
// Rows on one polynomial of degree < N give e = e0 = 0.  The reference also accepts
// degree exactly N (data_coeffs.len() <= RAW_BLOB_SIZE + 1, data_poly.rs:112): its x^N
// term reaches only D_0, whose degree-ROW part over the interpolation points is
// c A(u) for one constant c, so e = e0 = c A(u_x) on every extra row.  Flag bit 1 when
// e != e0 (another D_i has a higher degree); cx[x] = e0 / A(u_x) for k_rec_fold_const.
__global__ __launch_bounds__(256) void k_rec_check(const uint32_t* L, const uint32_t* A, const uint32_t* yv,
                                                   const uint32_t* V, uint32_t m, uint32_t col, uint32_t* cx,
                                                   int* flag) {
    __shared__ fr part[256], part0[256];
    const uint32_t x = blockIdx.x;
    fr s = f_zero<FrP>(), s0 = f_zero<FrP>();
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) {
        const fr l = f_load<FrP>(L + 8 * ((size_t)x * m + r));
        s = f_add(s, f_mul(l, f_load<FrP>(yv + 8 * (size_t)r)));
        s0 = f_add(s0, f_mul(l, f_load<FrP>(V + 8 * ((size_t)r * col))));
    }
    part[threadIdx.x] = s;
    part0[threadIdx.x] = s0;
    __syncthreads();
    for (uint32_t o = blockDim.x >> 1; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            part[threadIdx.x] = f_add(part[threadIdx.x], part[threadIdx.x + o]);
            part0[threadIdx.x] = f_add(part0[threadIdx.x], part0[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const fr a = f_load<FrP>(A + 8 * (size_t)x);
        const fr e = f_sub(f_load<FrP>(yv + 8 * (size_t)(m + x)), f_mul(part[0], a));
        const fr e0 = f_sub(f_load<FrP>(V + 8 * ((size_t)(m + x) * col)), f_mul(part0[0], a));
        if (!f_eq(e, e0)) atomicOr(flag, 2);
        f_store(cx + 8 * (size_t)x, f_mul(e0, f_inv(a)));  // A(u_x) != 0: the points are distinct
    }
}

// The degree-N coefficient c = cx[0] when every extra row gives the same one (else flag
// bit 1: degree above N); c = 0 for rows on a polynomial of degree < N.  One workgroup.
__global__ __launch_bounds__(256) void k_rec_fold_const(const uint32_t* cx, uint32_t nx, uint32_t* c, int* flag) {
    const fr c0 = f_load<FrP>(cx);
    for (uint32_t x = threadIdx.x + 1; x < nx; x += blockDim.x)
        if (!f_eq(f_load<FrP>(cx + 8 * (size_t)x), c0)) atomicOr(flag, 2);
    if (threadIdx.x == 0) f_store(c, c0);
}

// One workgroup per erased coset-0 row l = erased[t]: w_i = A(x_l) Out[t][i] w_N^(l i),
// forward DFT_COL, canonical 32 B into blob row l.  cfold (null: none) is the degree-N
// coefficient c of k_rec_fold_const: D_0(x_l) = A(x_l) (Out[t][0] + c), which at the
// coset-0 points (x^N = 1) is coeffs_to_evals' fold of the top coefficient.
__global__ __launch_bounds__(256) void k_rec_out(const uint32_t* Out, const uint32_t* A, const uint32_t* erased,
                                                 const uint32_t* tw, int log_col, int log_row, const uint32_t* cfold,
                                                 uint32_t* blob) {
    extern __shared__ fr rec_lds[];
    fr* lds = rec_lds;
    const uint32_t t = blockIdx.x, col = 1u << log_col, l = erased[t];
    const int log_n = log_col + log_row;
    const fr a = f_load<FrP>(A + 8 * (size_t)t);
    for (uint32_t i = threadIdx.x; i < col; i += blockDim.x) {
        fr o = f_load<FrP>(Out + 8 * ((size_t)t * col + i));
        if (cfold && i == 0) o = f_add(o, f_load<FrP>(cfold));
        fr w = f_mul(o, a);
        lds[bitrev_bits(i, log_col)] = f_mul(w, rec_wpow(tw, (uint64_t)l * i, log_n));
    }
    __syncthreads();
    rec_lds_dit(lds, tw, log_col, log_n);
    for (uint32_t j = threadIdx.x; j < col; j += blockDim.x)
        f_store(blob + 8 * ((size_t)l * col + j), f_from_mont(lds[j]));
}

// Received coset-0 rows straight into the blob (row k of the received list is row
// ids[k] < ROW), one thread per element
__global__ void k_rec_pass(const uint8_t* rows, const uint32_t* ids, uint32_t n0, int log_col, uint32_t* blob) {
    const uint32_t col = 1u << log_col;
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n0 * col) return;
    const uint32_t k = (uint32_t)(e >> log_col), j = (uint32_t)(e & (col - 1));
    const uint4* s = (const uint4*)(rows + ((size_t)k * col + j) * 32);
    uint4* d = (uint4*)(blob + 8 * ((size_t)ids[k] * col + j));
    d[0] = s[0];
    d[1] = s[1];
}

// RawBlob -> padded RawData bytes (raw_blob.rs:43-57): byte q of the output is byte
// q % 31 of element q / 31 (row-major blob); flag bit 2 when an element's top byte is
// not zero.  One thread per output u32 word (31 N is a multiple of 4 when N is).
__global__ void k_rec_pack(const uint32_t* blob, uint32_t n, uint32_t* raw_words, int* flag) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)31 * n / 4;
    if (w < n && (blob[8 * w + 7] >> 24)) atomicOr(flag, 4);
    if (w >= total) return;
    const uint8_t* b = (const uint8_t*)blob;
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t p = 4 * w + q, e = p / 31, o = p % 31;
        v |= (uint32_t)b[32 * e + o] << (8 * q);
    }
    raw_words[w] = v;
}

// RawData::get_actual_length (raw_data.rs:22-38): flag bit 3 when the length field
// exceeds max_raw, bit 4 when a byte in [length, max_raw) is not zero.  One thread
// per byte of [0, max_raw).
__global__ void k_rec_padding(const uint8_t* raw, uint64_t max_raw, int* flag) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t len = (uint32_t)raw[max_raw] | ((uint32_t)raw[max_raw + 1] << 8) |
                         ((uint32_t)raw[max_raw + 2] << 16) | ((uint32_t)raw[max_raw + 3] << 24);
    if (q == 0 && len > max_raw) atomicOr(flag, 8);
    if (q >= max_raw || len > max_raw) return;
    if (q >= len && raw[q]) atomicOr(flag, 16);
}

}  // namespace zgr
