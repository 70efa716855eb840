// Device kernels of the deferred AMT verification MSM (DeferredVerifier::record_msm /
// fast_check_msm_gpu, crates/amt/src/deferred_verification.rs:71-92, 203-222).
// Included once, by zgr.hip.
//
// Verifying row l of coset c (BlobRow::verify, crates/amt/src/blob/encode.rs:224-239;
// verify_amt_proof, crates/amt/src/verify_params.rs:133-211) checks
//   sum_i basis_c[b * COL + i] * data[i] == commitment - sum_d proof[d].commitment,
// with b = bitrev_ROW_LOG(l) and data = index_reverse(row).  The deferred verifier
// scales each row's scalars by a random alpha and runs ONE MSM over every recorded row;
// here that is one MSM per coset over the resident basis line, whose scalar vector
// holds alpha_k * data_k at chunk b_k and zeros elsewhere.
#pragma once
#include "curve.h"

namespace zgr {

// s[b * COL + i] = alpha_k * row_k[bitrev(i)] (Montgomery) for the rows ks[0..nk) of
// one coset; rows and alphas canonical.  flag bit 0: an element or alpha >= r.
__global__ void k_ver_scalars(const uint8_t* rows, const uint8_t* alphas, const uint32_t* ks, uint32_t nk,
                              const uint32_t* ids, int log_col, int log_row, uint32_t* s, int* flag) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ((uint64_t)nk << log_col)) return;
    const uint32_t col = 1u << log_col;
    const uint32_t kk = (uint32_t)(e >> log_col), i = (uint32_t)(e & (col - 1)), k = ks[kk];
    const uint32_t l = ids[k] & ((1u << log_row) - 1), b = bitrev_bits(l, log_row);
    const fr a = f_load<FrP>((const uint32_t*)(alphas + 32 * (size_t)k));
    const fr v = f_load<FrP>((const uint32_t*)(rows + 32 * ((size_t)k * col + bitrev_bits(i, log_col))));
    bool bad = false;
    for (int j = 7; j >= 0; j--) {
        if (a.v[j] != FrP::M[j]) {
            bad |= a.v[j] > FrP::M[j];
            break;
        }
        if (j == 0) bad = true;
    }
    for (int j = 7; j >= 0; j--) {
        if (v.v[j] != FrP::M[j]) {
            bad |= v.v[j] > FrP::M[j];
            break;
        }
        if (j == 0) bad = true;
    }
    if (bad) atomicOr(flag, 1);
    f_store(s + 8 * ((size_t)b * col + i), f_mul(f_to_mont(a), f_to_mont(v)));
}

// out = sum of count canonical affine points (64 zero bytes = infinity), canonical
__global__ void k_ver_sum(const uint32_t* pts, int count, uint32_t* out) {
    if (threadIdx.x || blockIdx.x) return;
    x29 acc = x29_inf();  // in the 29-bit form: the one inversion of the output is its cost
    for (int q = 0; q < count; q++) acc = x29_add(acc, ec29_from_canon(pts + 16 * q));
    x29_store_canon(out, acc);
}

}  // namespace zgr
