/*
 * zg_oracle.h -- CPU restatement of the 0G DA encoding path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation in 0g-da-encoder_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path never links or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the 0glabs/0g-da-encoder repository root).
 *
 * Conventions shared with the product C-ABI (include/zgr.h):
 *   - Field elements are 4 x u64 little-endian limbs.  "mont" = Montgomery form
 *     with R = 2^256 (arkworks' internal form, crates/amt/src/prove_params/
 *     fast_serde_bn254.rs:182-187); "canonical" = the plain integer.
 *   - A G1 affine point in a *param buffer* is x_mont || y_mont (64 bytes), the
 *     bamt layout of fast_serde_bn254.rs:139-147.
 *   - A G1 affine point in an *output* is x || y canonical LE (64 bytes); the
 *     point at infinity is 64 zero bytes ((0,0) is not on y^2 = x^3 + 3).
 */
#ifndef ZG_ORACLE_H
#define ZG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field primitives (exported for pinning tests) ---- */
/* which: 0 = Fr (scalar field), 1 = Fq (base field) */
void zgo_fmul(int which, const uint64_t a[4], const uint64_t b[4], uint64_t out[4]);
void zgo_to_mont(int which, const uint64_t a[4], uint64_t out[4]);
void zgo_from_mont(int which, const uint64_t a[4], uint64_t out[4]);
void zgo_finv(int which, const uint64_t a[4], uint64_t out[4]);
/* n elements out of Montgomery form */
void zgo_from_mont_n(int which, const uint64_t *a, size_t n, uint64_t *out);

/* ---- G1 primitives ---- */
/* scalar (canonical) * generator -> canonical affine (64B, zeros = infinity) */
void zgo_g1_mul_gen(const uint64_t k[4], uint8_t out[64]);
/* scalar (canonical) * P (canonical affine) */
void zgo_g1_mul(const uint8_t p[64], const uint64_t k[4], uint8_t out[64]);
void zgo_g1_add(const uint8_t p[64], const uint8_t q[64], uint8_t out[64]);
/* MSM: sum k_i * P_i.  bases in param layout (mont), scalars canonical */
void zgo_msm(const uint64_t *bases_mont, const uint64_t *scalars, size_t n, uint8_t out[64]);

/* ---- NTT / RS extension ---- */
/* omega for a 2^log_n domain (canonical) */
void zgo_domain_root(int log_n, uint64_t out[4]);
/* coset_factor(2^log_n, idx) (canonical) */
void zgo_coset_factor(int log_n, int idx, uint64_t out[4]);
/* in-place DFT / inverse DFT of 2^log_n Fr elements in mont form */
void zgo_fft(uint64_t *a_mont, int log_n);
void zgo_ifft(uint64_t *a_mont, int log_n);
/* to_coset_blob: data (mont, column-major) -> evaluations on coset idx */
void zgo_to_coset_blob(const uint64_t *data_mont, int log_n, int coset, uint64_t *out_mont);

/* ---- EC-FFT (G1) over the 2^log_n domain, reference gen_basis/gen_quotients ---- */
/* points: canonical affine in/out (64B each, zeros = infinity) */
void zgo_ecfft(const uint8_t *in, int log_n, int inverse, uint8_t *out);

/* ---- params ---- */
/* Prove bases for one coset from tau (canonical), line-major:
 * quotients[0..h) || basis || high_basis, each 2^log_n points, mont x||y.
 * method 0 = closed form, 1 = EC-FFT restatement of generate.rs:139-232. */
int zgo_params_from_tau(const uint64_t tau[4], int log_n, int h, int coset, int method,
                        uint64_t *out_bases);

/* One closed-form prove base: line (quotients 0..h-1, basis h, high h+1) at stored
 * (bit-reversed) position pos; out = x_mont || y_mont.  Returns 0 or -1. */
int zgo_params_point(const uint64_t tau[4], int log_n, int h, int coset, int line, uint64_t pos,
                     uint64_t out[8]);
/* f(x) of the degree < 2^log_n interpolant of y (mont evaluations over the domain),
 * x and out canonical.  Returns -1 if x lies on the domain. */
int zgo_eval_interp(const uint64_t *y_mont, int log_n, const uint64_t x[4], uint64_t out[4]);

/* ---- AMT prover ---- */
/* gen_all_proofs (CPU reference path).  ri_mont: 2^log_n Fr (mont), bit-reversed.
 * out_commitment: 64B; out_comm_tree / out_proof_tree: (2^(h+1)-2) points, level d
 * (d = 0..h-1, 2^(d+1) points) stored consecutively; out_high: 64B. */
int zgo_amt_prove(const uint64_t *bases_mont, int log_n, int h, const uint64_t *ri_mont,
                  uint8_t *out_commitment, uint8_t *out_comm_tree, uint8_t *out_proof_tree,
                  uint8_t *out_high);
/* multiple_multiexp semantics: out[L*nb + b] = sum_i bases[L*N + b*bs + i]*s[b*bs+i]
 * for L < h+2 lines; canonical affine output. */
int zgo_multiple_multiexp(const uint64_t *bases_mont, int log_n, int h, const uint64_t *ri_mont,
                          uint8_t *out_points);

/* ---- Keccak / Merkle ---- */
/* pad = 0x01 for Keccak-256 (tiny-keccak v256), 0x06 for FIPS SHA3-256 */
void zgo_keccak256(const uint8_t *msg, size_t len, uint8_t pad, uint8_t out[32]);
/* EncodedBlobMerkle::build over n_elems 32B elements; writes the top (levels+1)
 * layers tree[0..=levels] consecutively (tree[0] = coset roots) and the file root.
 * Returns the number of layers in the full tree, or -1. */
int zgo_merkle(const uint8_t *data, size_t n_elems, int coset_n, int levels,
               uint8_t *out_layers, uint8_t out_file_root[32]);

/* ---- packing ---- */
/* RawData::try_from + RawBlob::from: raw bytes -> 2^log_n Fr (mont), row-major.
 * Returns 0, or 1 for TooLargeBlob. */
int zgo_raw_to_blob(const uint8_t *raw, size_t len, int log_n, uint64_t *out_mont);

/* ---- full encode: RawData -> EncodedBlob ---- */
/* bases: coset_n param buffers (each (h+2)*2^log_n points).  Outputs:
 *   out_data:   coset_n * 2^log_n * 32 B (scalar_to_h256 of every encoded element)
 *   out_commit: 64 B
 *   out_comm_tree, out_proof_tree: coset_n * (2^(h+1)-2) points
 *   out_high:   coset_n points
 *   out_merkle: tree[0..=log_row] layers (3*(2^(log_row+1)-1) nodes * 32 B)
 *   out_root:   32 B file root
 * Returns 0, 1 = TooLargeBlob, 2 = inconsistent coset commitments. */
int zgo_encode(const uint64_t *const *bases, int log_col, int log_row, int coset_n,
               const uint8_t *raw, size_t len,
               uint8_t *out_data, uint8_t *out_commit, uint8_t *out_comm_tree,
               uint8_t *out_proof_tree, uint8_t *out_high, uint8_t *out_merkle,
               uint8_t *out_root);

/* ---- RS recovery: recover_from_da_slice (crates/recovery/src/lib.rs:15-43) ---- */
/* ids: n_rows strictly increasing row ids; rows: n_rows * 2^log_col canonical 32 B
 * elements.  out_raw: 31 * 2^(log_col+log_row) bytes (padded RawData), *out_len its
 * length.  Returns 0 ok, 1 not enough lines, 2 RowIdOverflow, 3 "Incorrect scalar",
 * 4 "Incorrect length field", 5 "Non zero in padding range", 6 element >= r (the
 * reference panics), 7 recovered degree above N (the reference asserts), 8 ExtaustiveK. */
int zgo_recover(const uint32_t *ids, const uint8_t *rows, size_t n_rows, int log_col, int log_row,
                int coset_n, uint8_t *out_raw, size_t *out_len);

/* ---- deferred verification MSM (crates/amt/src/deferred_verification.rs:71-92, 133-150) ---- */
/* out = sum_k alpha_k * sum_i basis_c[bitrev(l) * COL + i] * row_k[bitrev(i)] for rows
 * ids[k] = c * ROW + l (BlobRow::verify, crates/amt/src/blob/encode.rs:224-239); rows and
 * alphas canonical; basis_c = line log_row of bases[c].  Returns 0, -1 for a bad id. */
int zgo_verify_msm(const uint64_t *const *bases, int log_col, int log_row, int coset_n, const uint32_t *ids,
                   const uint8_t *rows, const uint8_t *alphas, size_t n_rows, uint8_t out[64]);

/* power_tau (crates/amt/src/power_tau.rs:28-42): out[i] = [tau^(start + i)] G1 for
 * i < 2^log_n, canonical affine (64 B each); tau canonical. */
void zgo_power_tau(const uint64_t tau[4], uint64_t start, int log_n, uint8_t *out);

/* Thread count used by the parallel loops (OpenMP); returns the previous value. */
int zgo_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
