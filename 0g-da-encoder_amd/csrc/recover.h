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
// their D_i values (Schwartz-Zippel), the case where the reference's degree assert
// fires (data_poly.rs:112).
#pragma once
#include "field.h"

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

// L[t][r] = 1 / ((x_t - u_r) B_r), B_r = prod_{q != r} (u_r - u_q): the Lagrange matrix
// without its row factor A(x_t), b_r = 1 / B_r folded into the same inversion.  One
// thread per (target t, chunk of REC_CHUNK sources), Montgomery's trick over the chunk
// with the prefix products parked in L itself.
__global__ void k_rec_lagrange(const uint32_t* t_pts, uint32_t nt, const uint32_t* u, const uint32_t* B, uint32_t m,
                               uint32_t* L) {
    const uint32_t chunks = (m + REC_CHUNK - 1) / REC_CHUNK;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nt * chunks) return;
    const uint32_t t = g / chunks, r0 = (g % chunks) * REC_CHUNK;
    const uint32_t r1 = min(r0 + REC_CHUNK, m);
    const fr x = f_load<FrP>(t_pts + 8 * (size_t)t);
    uint32_t* row = L + 8 * (size_t)t * m;
    fr acc = f_one<FrP>();
    for (uint32_t r = r0; r < r1; r++) {
        f_store(row + 8 * (size_t)r, acc);  // prefix prod_{r0 <= q < r} d_q
        acc = f_mul(acc, f_mul(f_sub(x, f_load<FrP>(u + 8 * (size_t)r)), f_load<FrP>(B + 8 * (size_t)r)));
    }
    fr inv = f_inv(acc);
    for (uint32_t r = r1; r-- > r0;) {
        const fr d = f_mul(f_sub(x, f_load<FrP>(u + 8 * (size_t)r)), f_load<FrP>(B + 8 * (size_t)r));
        f_store(row + 8 * (size_t)r, f_mul(inv, f_load<FrP>(row + 8 * (size_t)r)));
        inv = f_mul(inv, d);
    }
}

// Fr in 29-bit limbs for the GEMM: products of two limbs stay below 2^58, so a 64-bit
// column takes 63 of them (7 terms of 9 products each) between carry passes, and the
// Montgomery reduction runs once per output instead of once per product.
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
__device__ __forceinline__ void rec_to29(const fr& a, uint32_t* o) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int bit = 29 * k, w = bit >> 5, sh = bit & 31;
        uint32_t v = a.v[w] >> sh;
        if (sh > 3 && w + 1 < 8) v |= a.v[w + 1] << (32 - sh);
        o[k] = v & R29_MASK;
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
            rec_to29(x, Ls[a][b]);
        }
        for (uint32_t e = threadIdx.x; e < REC_TK * REC_TI; e += blockDim.x) {
            const uint32_t a = e / REC_TI, b = e % REC_TI;
            const uint32_t r = k0 + a, i = i0 + b;
            const fr x = (r < m && i < col) ? f_load<FrP>(V + 8 * ((size_t)r * col + i)) : f_zero<FrP>();
            rec_to29(x, Vs[a][b]);
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

// Consistency of the extra rows (indices m..nrows-1 of the received list): their
// combination yv must equal the interpolant A(u_x) sum_r L[x][r] yv[r].  One workgroup
// (256 threads) per extra row x.  flag bit 1.
__global__ __launch_bounds__(256) void k_rec_check(const uint32_t* L, const uint32_t* A, const uint32_t* yv,
                                                   uint32_t m, int* flag) {
    __shared__ fr part[256];
    const uint32_t x = blockIdx.x;
    fr s = f_zero<FrP>();
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x)
        s = f_add(s, f_mul(f_load<FrP>(L + 8 * ((size_t)x * m + r)), f_load<FrP>(yv + 8 * (size_t)r)));
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = blockDim.x >> 1; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] = f_add(part[threadIdx.x], part[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const fr v = f_mul(part[0], f_load<FrP>(A + 8 * (size_t)x));
        if (!f_eq(v, f_load<FrP>(yv + 8 * (size_t)(m + x)))) atomicOr(flag, 2);
    }
}

// One workgroup per erased coset-0 row l = erased[t]: w_i = A(x_l) Out[t][i] w_N^(l i),
// forward DFT_COL, canonical 32 B into blob row l.
__global__ __launch_bounds__(256) void k_rec_out(const uint32_t* Out, const uint32_t* A, const uint32_t* erased,
                                                 const uint32_t* tw, int log_col, int log_row, uint32_t* blob) {
    extern __shared__ fr rec_lds[];
    fr* lds = rec_lds;
    const uint32_t t = blockIdx.x, col = 1u << log_col, l = erased[t];
    const int log_n = log_col + log_row;
    const fr a = f_load<FrP>(A + 8 * (size_t)t);
    for (uint32_t i = threadIdx.x; i < col; i += blockDim.x) {
        fr w = f_mul(f_load<FrP>(Out + 8 * ((size_t)t * col + i)), a);
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
