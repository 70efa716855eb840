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

// Per-call constants:
//   cpow[c * COL + i] = cf_c^-i / COL      (row iDFT scaling, c < coset_n)
//   gpow[i]           = gamma^i             (random combination of a row's D_i)
// gamma arrives canonical.  One thread per (c, i) plus COL threads for gpow.
__global__ void k_rec_consts(const uint32_t* consts, const uint32_t* gamma_canon, int coset_n, int log_col,
                             uint32_t* cpow, uint32_t* gpow) {
    const uint32_t col = 1u << log_col;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (uint32_t)coset_n * col) {
        const uint32_t c = t >> log_col, i = t & (col - 1);
        fr cfi = f_inv(f_load<FrP>(consts + 8 * (K_CF + c)));
        fr colinv = f_inv(fr_from_u64(col));
        f_store(cpow + 8 * (size_t)t, f_mul(f_pow_u64(cfi, i), colinv));
    } else if (t < (uint32_t)(coset_n + 1) * col) {
        const uint32_t i = t - (uint32_t)coset_n * col;
        fr g = f_to_mont(f_load<FrP>(gamma_canon));
        f_store(gpow + 8 * (size_t)i, f_pow_u64(g, i));
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

// Barycentric weights of the interpolation set (the first m = ROW points of u):
// b_r = 1 / prod_{r' != r} (u_r - u_r'), one thread per r
__global__ void k_rec_bary(const uint32_t* u, uint32_t m, uint32_t* b) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const fr ur = f_load<FrP>(u + 8 * (size_t)r);
    fr p = f_one<FrP>();
    for (uint32_t q = 0; q < m; q++)
        if (q != r) p = f_mul(p, f_sub(ur, f_load<FrP>(u + 8 * (size_t)q)));
    f_store(b + 8 * (size_t)r, f_inv(p));
}

// A(t) = prod_r (t - u_r) over the interpolation set, one thread per target
__global__ void k_rec_node(const uint32_t* t_pts, uint32_t nt, const uint32_t* u, uint32_t m, uint32_t* A) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const fr x = f_load<FrP>(t_pts + 8 * (size_t)t);
    fr p = f_one<FrP>();
    for (uint32_t q = 0; q < m; q++) p = f_mul(p, f_sub(x, f_load<FrP>(u + 8 * (size_t)q)));
    f_store(A + 8 * (size_t)t, p);
}

constexpr uint32_t REC_CHUNK = 64;  // entries of one batched inversion in k_rec_lagrange

// L[t][r] = b_r / (x_t - u_r) (the Lagrange matrix without its row factor A(x_t)):
// one thread per (target t, chunk of REC_CHUNK sources), Montgomery's trick over the
// chunk with the prefix products parked in L itself.  m is a multiple of REC_CHUNK or
// smaller than it.
__global__ void k_rec_lagrange(const uint32_t* t_pts, uint32_t nt, const uint32_t* u, const uint32_t* b, uint32_t m,
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
        acc = f_mul(acc, f_sub(x, f_load<FrP>(u + 8 * (size_t)r)));
    }
    fr inv = f_inv(acc);
    for (uint32_t r = r1; r-- > r0;) {
        const fr d = f_sub(x, f_load<FrP>(u + 8 * (size_t)r));
        const fr di = f_mul(inv, f_load<FrP>(row + 8 * (size_t)r));
        inv = f_mul(inv, d);
        f_store(row + 8 * (size_t)r, f_mul(di, f_load<FrP>(b + 8 * (size_t)r)));
    }
}

constexpr int REC_TL = 16;  // GEMM tile: targets
constexpr int REC_TI = 64;  // GEMM tile: D_i columns
constexpr int REC_TK = 16;  // GEMM tile: sources per LDS stage

// Out[t][i] = sum_{r < m} L[t][r] V[r][i] for t < nt, i < col: 256 threads, a
// 16 x 64 output tile per workgroup, each thread 4 targets of one column.
__global__ __launch_bounds__(256) void k_rec_gemm(const uint32_t* L, const uint32_t* V, uint32_t nt, uint32_t m,
                                                  uint32_t col, uint32_t* Out) {
    __shared__ fr Ls[REC_TL][REC_TK];
    __shared__ fr Vs[REC_TK][REC_TI];
    const uint32_t tx = threadIdx.x % REC_TI, ty = threadIdx.x / REC_TI;  // ty < 4
    const uint32_t t0 = blockIdx.y * REC_TL, i0 = blockIdx.x * REC_TI;
    fr acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = f_zero<FrP>();
    for (uint32_t k0 = 0; k0 < m; k0 += REC_TK) {
        {  // L tile: 16 x 16 = 256 entries, one per thread
            const uint32_t a = threadIdx.x / REC_TK, bb = threadIdx.x % REC_TK;
            const uint32_t t = t0 + a, r = k0 + bb;
            Ls[a][bb] = (t < nt && r < m) ? f_load<FrP>(L + 8 * ((size_t)t * m + r)) : f_zero<FrP>();
        }
        for (uint32_t e = threadIdx.x; e < REC_TK * REC_TI; e += blockDim.x) {  // V tile: 16 x 64
            const uint32_t a = e / REC_TI, bb = e % REC_TI;
            const uint32_t r = k0 + a, i = i0 + bb;
            Vs[a][bb] = (r < m && i < col) ? f_load<FrP>(V + 8 * ((size_t)r * col + i)) : f_zero<FrP>();
        }
        __syncthreads();
#pragma unroll 2
        for (int kk = 0; kk < REC_TK; kk++) {
            const fr v = Vs[kk][tx];
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = f_add(acc[q], f_mul(Ls[ty * 4 + q][kk], v));
        }
        __syncthreads();
    }
    const uint32_t i = i0 + tx;
    if (i >= col) return;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t t = t0 + ty * 4 + q;
        if (t < nt) f_store(Out + 8 * ((size_t)t * col + i), acc[q]);
    }
}

// Consistency of the extra rows (indices m..nrows-1 of the received list): their
// combination yv must equal the interpolant A(u_x) sum_r L[x][r] yv[r].  flag bit 1.
__global__ void k_rec_check(const uint32_t* L, const uint32_t* A, const uint32_t* yv, uint32_t m, uint32_t nx,
                            int* flag) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    fr s = f_zero<FrP>();
    for (uint32_t r = 0; r < m; r++)
        s = f_add(s, f_mul(f_load<FrP>(L + 8 * ((size_t)x * m + r)), f_load<FrP>(yv + 8 * (size_t)r)));
    s = f_mul(s, f_load<FrP>(A + 8 * (size_t)x));
    if (!f_eq(s, f_load<FrP>(yv + 8 * (size_t)(m + x)))) atomicOr(flag, 2);
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
