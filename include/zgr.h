/*
 * zgr.h -- C ABI of the MI355X (gfx950) 0G DA encoder.
 *
 * One shared library, lib/libzgr.so inside 0g-da-encoder_amd/, implements the
 * encoding path RawBlob -> EncodedBlob of 0glabs/0g-da-encoder on the GPU.  Every
 * entry point below names the reference interface it replaces (paths relative to
 * the reference repository root).
 *
 * Conventions
 *   - Plain pointers and sizes only.  Inputs are borrowed for the duration of the
 *     call; outputs are caller-allocated.  Nothing is freed across the boundary.
 *   - Fr / Fq elements are 32 bytes little-endian.  "mont" = Montgomery form with
 *     R = 2^256 (arkworks' internal representation).
 *   - Prove-base buffers (params) hold affine G1 points as x_mont || y_mont
 *     (64 B each, the bamt layout of crates/amt/src/prove_params/
 *     fast_serde_bn254.rs:139-147), line-major in the order read_gpu_bases uploads
 *     them (crates/amt/src/prove_params/prove_gpu.rs:42-47):
 *         quotients[0] .. quotients[h-1] || basis || high_basis,
 *     each line 2^log_n points, so one coset's buffer is (h + 2) * 2^log_n * 64 B.
 *   - Output points are canonical affine x || y (64 B).  The point at infinity is
 *     64 zero bytes ((0, 0) is not on y^2 = x^3 + 3).
 *   - Return value: ZGR_OK or a negative ZGR_ERR_*; zgr_last_error() gives a
 *     thread-local message.  Never unwinds.
 *   - A context serialises the calls made on it; use one context per host thread
 *     for concurrency, or the batch entry points.
 */
#ifndef ZGR_H
#define ZGR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ZGR_OK = 0,
    ZGR_ERR_INVALID_ARG = -1,
    ZGR_ERR_TOO_LARGE = -2,      /* EncoderError::TooLargeBlob (crates/encoder/src/encoder/error.rs:3-6) */
    ZGR_ERR_OOM = -3,            /* device allocation / admission failure */
    ZGR_ERR_DEVICE = -4,         /* HIP runtime error */
    ZGR_ERR_NOT_LOADED = -5,     /* params for a coset are not resident */
    ZGR_ERR_INCONSISTENT = -6,   /* coset commitments differ (crates/encoder/src/amt/blob.rs:42-47);
                                    recovery rows off one polynomial (crates/recovery/src/data_poly.rs:112) */
    ZGR_ERR_CANCELLED = -7       /* a served request withdrawn by zgr_server_cancel */
};

typedef struct zgr_ctx zgr_ctx;

/* Outputs of one encoded blob, caller-allocated (NULL = not wanted).
 * h = log_row, nb = 2^h, T = 2^(h+1) - 2 (points of one AllProofs tree). */
typedef struct {
    uint8_t *data;        /* coset_n * 2^log_n * 32: scalar_to_h256 of every encoded element,
                             coset-major then row-major (EncodedBlob::get_data,
                             crates/encoder/src/encoder/blob.rs:61) */
    uint8_t *commitment;  /* 64: EncodedBlob::get_affine_commitment (blob.rs:51) */
    uint8_t *comm_tree;   /* coset_n * T * 64: AllProofs.commitments, level d (2^(d+1) points)
                             after levels 0..d-1 (crates/amt/src/proofs.rs:8-14) */
    uint8_t *proof_tree;  /* coset_n * T * 64: AllProofs.proofs, same layout */
    uint8_t *high;        /* coset_n * 64: AllProofs.high_commitment */
    uint8_t *merkle;      /* coset_n * (2^(h+1) - 1) * 32: EncodedBlobMerkle.tree[0..=h]
                             (crates/encoder/src/merkle/blob.rs:15-34) */
    uint8_t *file_root;   /* 32: EncodedBlob::get_file_root (blob.rs:57-59) */
} zgr_encoded;

const char *zgr_last_error(void);

/* Context on one device for one blob geometry (log_col, log_row, coset_n).
 * Replaces the per-thread workspace of ag_cuda_ec::init_local_workspace
 * (crates/amt/src/prove_params/interfaces.rs:41-44).  coset_n <= 3, log_col >= 3,
 * log_row >= 1, log_col + log_row <= 22 (the NTT's two-pass limit; production is 20). */
int zgr_ctx_create(int device, int log_col, int log_row, int coset_n, zgr_ctx **out);
void zgr_ctx_destroy(zgr_ctx *ctx);

/* ---- resident public parameters (EncoderParams::warmup, crates/amt/src/blob/encode.rs:85-89;
 *      read_gpu_bases, prove_gpu.rs:31-53) ---- */
/* Upload one coset's prove bases from host memory (layout above). */
int zgr_params_load(zgr_ctx *ctx, int coset, const void *bases_mont, size_t nbytes);
/* Upload one coset's prove bases from the image of its bamt file
 * amt-prove-coset{c}-mont-{hash}-{h:02}-{n:02}.bin (AMTParams::from_dir_mont,
 * crates/amt/src/prove_params/generate.rs:27-65; read_amt_params,
 * crates/amt/src/prove_params/fast_serde_bn254.rs:58-88).  The G2 sections (g2,
 * high_g2, vanishes) belong to the verifier and are skipped.  ZGR_ERR_INVALID_ARG for a
 * wrong header ("Incorrect format"), degrees that differ from the context, or a length
 * the layout does not account for. */
int zgr_params_load_bamt(zgr_ctx *ctx, int coset, const uint8_t *bamt, size_t nbytes);
/* Insecure test setup: generate every coset's prove bases on the device from a known
 * tau (32 B canonical LE).  Same points as PowerTau::setup_inner with that tau
 * (crates/amt/src/power_tau.rs:54-80) followed by AMTParams::from_pp
 * (crates/amt/src/prove_params/generate.rs:139-195). */
int zgr_params_generate(zgr_ctx *ctx, const uint8_t tau[32]);
/* AMTParams::from_pp (crates/amt/src/prove_params/generate.rs:139-232) on the device:
 * the prove bases of one coset from powers of tau, without tau -- g1pp[i] = [tau^i] G1
 * and high_g1pp[i] = [tau^(H+i)] G1 for i < 2^log_n (PowerTau, crates/amt/src/
 * power_tau.rs:20-80, as a ceremony provides them), canonical affine x || y, 64 B each.
 * Coset scaling by cf_c^-i, then the EC-FFTs of gen_basis / gen_quotients (Pairing::
 * fast_fft / fast_ifft) and enact, straight into the resident layout; the shifted-copy
 * tables follow.  n_points must be 2^log_n. */
int zgr_params_from_pp(zgr_ctx *ctx, int coset, const uint8_t *g1pp, const uint8_t *high_g1pp,
                       size_t n_points);
/* Copy one coset's resident bases back to host memory (layout above). */
int zgr_params_download(zgr_ctx *ctx, int coset, void *out, size_t nbytes);

/* ---- the rest of the parameter pipeline: powers of tau, the verifier's G2 data and the
 *      reference's parameter files (bin/build_params.rs:24-53) ----
 * G2 points in memory are canonical affine x.c0 || x.c1 || y.c0 || y.c1 (128 B; 128 zero
 * bytes = infinity), G1 points canonical affine x || y (64 B).
 * zgr_params_generate also sets every coset's verifier data (closed form, from tau);
 * zgr_params_load_bamt keeps the file's; zgr_params_from_pp (G1 only) clears it. */
/* PowerTau::setup_inner with a known tau (crates/amt/src/power_tau.rs:54-80), on the device,
 * kept in the context: g1pp[i] = [tau^i] G1, g2pp[i] = [tau^i] G2, high_g1pp[i] =
 * [tau^(H+i)] G1 (i < 2^log_n, H = 3 * 2^log_n, power_tau.rs:66-67), high_g2 = [tau^H] G2. */
int zgr_power_tau_setup(zgr_ctx *ctx, const uint8_t tau[32]);
/* Powers of tau from elsewhere (a ceremony), kept in the context; n_points = 2^log_n. */
int zgr_power_tau_load(zgr_ctx *ctx, const uint8_t *g1pp, const uint8_t *g2pp, const uint8_t *high_g1pp,
                       const uint8_t *high_g2, size_t n_points);
/* Ceremony import: decode and check count points of a perpetual-powers-of-tau file, as
 * BatchedAccumulator::read_chunk with CheckForCorrectness::Yes does for
 * crates/ppot2ark/src/main.rs:45-119 (from_ppot_file_inner; the powersoftau crate of
 * kobigurk/phase2-bn254 rev dd6b966, pairing_ce bn256 encodings: big-endian coordinates,
 * G2 as c1 || c0, flag bits 0x80 / 0x40 in the first byte).  src holds the count
 * encodings back to back (64 / 32 B per G1 point uncompressed / compressed, 128 / 64 B
 * per G2 point); out receives canonical affine points (G1 64 B, G2 128 B, the layouts
 * above).  A point that a checked read rejects -- wrong compression mode, stray bits
 * with the infinity flag, the point at infinity, a coordinate >= p, off the curve, or a
 * G2 point outside the order-r subgroup -- fails the call with ZGR_ERR_INVALID_ARG,
 * the reason and *first_bad = its index (count when all are accepted).  Decoding,
 * square roots and the subgroup test run on the device. */
enum {
    ZGR_PPOT_G1_UNCOMPRESSED = 0,
    ZGR_PPOT_G1_COMPRESSED = 1,
    ZGR_PPOT_G2_UNCOMPRESSED = 2,
    ZGR_PPOT_G2_COMPRESSED = 3
};
int zgr_ppot_decode(zgr_ctx *ctx, int kind, const uint8_t *src, size_t count, uint8_t *out, size_t *first_bad);
/* Copy the context's powers out (NULL = not wanted); ZGR_ERR_NOT_LOADED without any. */
int zgr_power_tau_download(zgr_ctx *ctx, uint8_t *g1pp, uint8_t *g2pp, uint8_t *high_g1pp, uint8_t *high_g2);
/* AMTParams::from_pp (crates/amt/src/prove_params/generate.rs:139-195) for one coset from
 * the context's powers: the prove bases (as zgr_params_from_pp) and the verifier half --
 * vanishes[d-1] = enact(gen_vanishes(g2pp', d)) for d = 1..log_row (generate.rs:234-263,
 * a G2 DFT of 2^d points per level with the inverse root, unscaled), g2 = g2pp[0] and
 * high_g2. */
int zgr_params_from_power_tau(zgr_ctx *ctx, int coset);
/* One coset's verifier data: vanishes ((2^(log_row+1) - 2) G2, level d = 1..log_row
 * holding 2^d points at point offset 2^d - 2, in AMTParams.vanishes order), g2, high_g2. */
int zgr_params_g2(zgr_ctx *ctx, int coset, uint8_t *vanishes, uint8_t *g2, uint8_t *high_g2);

/* Images of the reference's parameter files, written from what the context holds:
 *   ZGR_FILE_AMT_MONT    amt-prove-coset{c}-mont-{hash}-{h:02}-{n:02}.bin: write_amt_params
 *                        (crates/amt/src/prove_params/fast_serde_bn254.rs:20-55)
 *   ZGR_FILE_AMT         amt-prove-coset{c}-{hash}-{h:02}-{n:02}.bin: serialize_uncompressed
 *                        of AMTParams (crates/amt/src/prove_params/serde.rs)
 *   ZGR_FILE_AMT_VERIFY  amt-verify-coset{c}-{hash}-{vd:02}-{n:02}.bin: serialize_uncompressed
 *                        of AMTVerifyParams with sub_depth = vd levels (crates/amt/src/
 *                        verify_params.rs:22-28, 81-98)
 *   ZGR_FILE_PTAU        power-tau-{hash}-{n:02}.bin: serialize_compressed of PowerTau
 *                        (power_tau.rs:20-26, 121-125); coset ignored
 *   ZGR_FILE_PTAU_MONT   power-tau-mont-{hash}-{n:02}.bin: write_power_tau
 *                        (fast_serde_bn254.rs:90-113); coset ignored
 * ({hash} = 5DWgDV for BN254, README.md:73-82.)  out == NULL: *len receives the size.
 * ZGR_ERR_INVALID_ARG when cap is smaller (with *len set); ZGR_ERR_NOT_LOADED when the
 * context lacks the data (bases, verifier data, or powers). */
enum {
    ZGR_FILE_AMT_MONT = 0,
    ZGR_FILE_AMT = 1,
    ZGR_FILE_AMT_VERIFY = 2,
    ZGR_FILE_PTAU = 3,
    ZGR_FILE_PTAU_MONT = 4
};
int zgr_write_params_file(zgr_ctx *ctx, int kind, int coset, int sub_depth, uint8_t *out, size_t cap,
                          size_t *len);

/* ---- AMT prover seam: AMTProofs::gen_amt_proofs
 *      (crates/amt/src/prove_params/interfaces.rs:7-45) ---- */
/* ri_mont: 2^log_n Fr (mont), already bit-reversed (HalfBlob::generate,
 * crates/amt/src/blob/encode.rs:171-183).  Outputs: commitment (64 B),
 * comm_tree / proof_tree (T points each), high (64 B). */
int zgr_amt_prove(zgr_ctx *ctx, int coset, const uint8_t *ri_mont, size_t n,
                  uint8_t *commitment, uint8_t *comm_tree, uint8_t *proof_tree, uint8_t *high);
/* ag_cuda_ec::multiple_multiexp_mt as called by gen_all_proofs_gpu
 * (prove_gpu.rs:77-84): out[L*nb + b] = sum_{i<bs} bases[L*N + b*bs + i] * s[b*bs + i]
 * for the h + 2 lines L.  out: (h + 2) * nb canonical points. */
int zgr_multiple_multiexp(zgr_ctx *ctx, int coset, const uint8_t *ri_mont, size_t n,
                          uint8_t *out_points);

/* ---- full encode: EncodedBlob::build over RawData::try_from + RawBlob::from
 *      (crates/encoder/src/encoder/blob.rs:32-42, raw_data.rs:55-73, raw_blob.rs:33-40),
 *      the body of EncoderService::process_data (grpc/src/service.rs:67-106) ---- */
int zgr_encode(zgr_ctx *ctx, const uint8_t *raw, size_t len, const zgr_encoded *out);
/* Same, with the raw bytes already in device memory (pointer from hipMalloc /
 * torch), outputs to host. */
int zgr_encode_device(zgr_ctx *ctx, const void *raw_dev, size_t len, const zgr_encoded *out);
/* Batched encode of nblobs blobs whose raw bytes are already in device memory; the
 * blobs share each pass over the resident bases.  Outputs to host. */
int zgr_encode_batch_device(zgr_ctx *ctx, int nblobs, const void *const *raw_dev,
                            const size_t *lens, const zgr_encoded *outs);

/* ---- serving: continuous batching over one context ----
 * The reference serves one blob per gRPC request, synchronously on a tokio worker
 * (Encoder::encode_blob -> process_data, grpc/src/service.rs:48-106).  A server keeps
 * the context's params resident and runs a worker thread that admits queued requests
 * while the device staging arena (hbm_budget bytes of raw input) has room and at most
 * max_inflight requests are admitted, uploads them, and issues each into the next
 * free pipeline slot as soon as one frees up, so requests that arrive while others
 * are on the GPU overlap with them.  Each request's outputs are exactly those of
 * zgr_encode on the same bytes. */
typedef struct zgr_server zgr_server;
int zgr_server_create(zgr_ctx *ctx, size_t hbm_budget, int max_inflight, zgr_server **out);
/* Completes every submitted request, then stops the worker.  While a server exists
 * it owns its context: other calls on that context wait until it is destroyed. */
void zgr_server_destroy(zgr_server *srv);
/* Non-blocking.  raw (host memory) and out must stay valid until the ticket
 * completes.  Returns ZGR_ERR_TOO_LARGE for oversize input, or ZGR_ERR_OOM when a
 * single request can never fit the staging arena. */
int zgr_server_submit(zgr_server *srv, const uint8_t *raw, size_t len, const zgr_encoded *out,
                      uint64_t *ticket);
/* Wait up to timeout_ms (< 0: forever) for a ticket.  Returns 1 while pending, else
 * the request's final status (ZGR_OK or an error), once. */
int zgr_server_poll(zgr_server *srv, uint64_t ticket, int timeout_ms);
/* Withdraw a request that has not been issued to the GPU yet (queued, or admitted and
 * waiting for a slot): its poll then returns ZGR_ERR_CANCELLED and its buffers are free.
 * No counterpart in the reference, whose requests run to completion once received
 * (grpc/src/service.rs:48-65).  ZGR_ERR_INVALID_ARG when the ticket is not pending or
 * is already on the GPU (the request then completes as usual). */
int zgr_server_cancel(zgr_server *srv, uint64_t ticket);
/* queued (not yet admitted), admitted (staged or on the GPU), completed counts
 * (completed includes cancelled requests) */
void zgr_server_stats(zgr_server *srv, uint64_t *queued, uint64_t *admitted, uint64_t *completed);

/* ---- failure injection, for tests of the error paths ----
 * No counterpart in the reference (whose GPU errors panic, prove_gpu.rs:48,84).  Arms a
 * host-side failure: the countdown-th next event at `site` fails with ZGR_ERR_DEVICE
 * and the message "injected fault: ..." before it launches anything, so every kernel
 * still completes normally.  countdown 0 disarms.  Safe to call while a server owns
 * the context.
 *   ZGR_FAULT_ISSUE  the issue of one blob into a pipeline slot (zgr_encode*, served
 *                    requests) */
enum { ZGR_FAULT_ISSUE = 0, ZGR_FAULT_SITES = 1 };
int zgr_test_inject_fault(zgr_ctx *ctx, int site, int countdown);

/* ---- primitives (reference call sites named per function) ---- */
/* Radix2EvaluationDomain::fft / ifft of ark-poly over the 2^log_n domain, natural
 * order in and out (to_coset_blob, crates/amt/src/blob/encode.rs:142-157).  inverse != 0
 * is ifft (inverse root, then size_inv).  batch rows of 2^log_n Fr (mont), in place,
 * host memory.  log_n <= the context's log_col + log_row. */
int zgr_ntt(zgr_ctx *ctx, uint8_t *data_mont, int log_n, int batch, int inverse);
/* The same on rows already in device memory (in place): up to 2^11 points per row one
 * launch for the whole batch, each row in LDS; larger rows two passes over HBM. */
int zgr_ntt_device(zgr_ctx *ctx, void *data_dev, int log_n, int batch, int inverse);
/* to_coset_blob (encode.rs:135-158) for coset 1..coset_n-1 of ctx's geometry:
 * data_mont (2^log_n, column-major) -> out_mont evaluations on that coset. */
int zgr_coset_extend(zgr_ctx *ctx, const uint8_t *data_mont, int coset, uint8_t *out_mont);
/* Plain MSM sum_i k_i * P_i: VariableBaseMSM::msm as the CPU prover calls it per line
 * (crates/amt/src/prove_params/prove.rs:17,31,55,77).  Bases in param layout (mont),
 * scalars mont; the result is canonical affine.  Any n (zero-padded to a power of two).
 * An all-zero base (never in a param file: fast_serde_bn254.rs:139-142 refuses to write
 * it) is read as the point at infinity. */
int zgr_msm(zgr_ctx *ctx, const uint8_t *bases_mont, const uint8_t *scalars_mont, size_t n,
            uint8_t *out_point);
/* G1 EC-FFT: Pairing::fast_fft / fast_ifft (crates/amt/src/ec_algebra.rs:34-95) over
 * 2^log_n points (canonical affine in and out); inverse applies size_inv. */
int zgr_ecfft(zgr_ctx *ctx, const uint8_t *points, int log_n, int inverse, uint8_t *out);
/* EncodedBlobMerkle::build over n_elems 32-byte elements (host memory): writes
 * tree[0..=levels] and the file root (merkle/blob.rs:15-34, encoder/blob.rs:64-72). */
int zgr_merkle(zgr_ctx *ctx, const uint8_t *data, size_t n_elems, int levels,
               uint8_t *out_layers, uint8_t *file_root);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * zgr_merkle with every buffer in device memory -- data (n_elems * 32 B), out_layers
 * (tree[0..=levels], as zgr_merkle lays them out) and file_root (32 B) -- and the
 * intermediate layers in a fixed scratch array of ZGR_MERKLE_SCRATCH_BYTES that each GPU
 * holds for the life of the process, instead of per-call allocations.  The call itself
 * allocates nothing.  n_elems * 8 must not exceed ZGR_MERKLE_SCRATCH_BYTES (n_elems <=
 * 4 Mi, a 1024 x 1024 blob's 3 Mi encoded elements fit); larger inputs get
 * ZGR_ERR_INVALID_ARG.  Synchronous on the context's stream.  Every call on one GPU uses
 * that GPU's one scratch array: calls that overlap on one GPU (from different contexts
 * or threads) race. */
#define ZGR_MERKLE_SCRATCH_BYTES ((size_t)32 << 20)
int zgr_merkle_noalloc(zgr_ctx *ctx, const void *data_dev, size_t n_elems, int levels,
                       void *out_layers_dev, void *file_root_dev);

/* ---- deferred AMT verification (the GPU half of DeferredVerifier) ---- */
/* DeferredVerifier::fast_check_msm_gpu (crates/amt/src/deferred_verification.rs:203-222)
 * over the rows record_msm collected (deferred_verification.rs:71-92), with the
 * verify basis = the resident prove basis line: for each row k = coset c, local l,
 *   out = sum_k alpha_k * sum_i basis_c[bitrev(l) * COL + i] * row_k[bitrev(i)]
 * (BlobRow::verify, crates/amt/src/blob/encode.rs:224-239, and verify_amt_proof,
 * crates/amt/src/verify_params.rs:176-186).  row_ids strictly increasing, < coset_n *
 * 2^log_row; rows: n_rows * 32 * 2^log_col canonical bytes (BlobRow.row); alphas: n_rows
 * canonical Fr (the random factors; fast_check_msm compares out with the sum of
 * alpha_k (commitment - sum_d proof_k[d].commitment)).  out_point: canonical affine
 * (64 zero bytes = infinity).  Needs the params of every coset named. */
int zgr_verify_msm(zgr_ctx *ctx, const uint32_t *row_ids, const uint8_t *rows, const uint8_t *alphas,
                   size_t n_rows, uint8_t *out_point);

/* ---- RS recovery ---- */
/* recover_from_da_slice (crates/recovery/src/lib.rs:15-43): the raw data of a blob
 * from any 2^log_row of its encoded rows.  row_ids: n_rows strictly increasing row
 * indices (the keys of the reference's BTreeMap); rows: n_rows * 32 * 2^log_col
 * bytes, each row the canonical 32 B elements of EncodedSlice::merkle_row
 * (raw_slice_to_line, crates/recovery/src/utils.rs:26-34).  Rows of another length
 * are the caller's to drop, as the reference does (lib.rs:21-32).  raw_out (raw_cap
 * >= MAX_RAW_DATA_SIZE = 31 * 2^log_n - 4) receives RawData::as_bytes, *raw_len its
 * length.  Errors, in the reference's order (zgr_last_error gives its text):
 *   ZGR_ERR_INVALID_ARG   an element >= r (the reference panics), "Not enough valid
 *                         lines", "Cannot recover data: RowIdOverflow", "Incorrect
 *                         scalar", "Incorrect length field", "Non zero in padding range";
 *   ZGR_ERR_INCONSISTENT  the rows do not lie on one polynomial of degree <= N (the
 *                         reference's degree assert, data_poly.rs:112).
 * Rows on a polynomial of degree exactly N decode as the reference's do: its top
 * coefficient is added to every coset-0 value (coeffs_to_evals, crates/recovery/src/
 * utils.rs:76-110).  * Needs no params; runs synchronously on the context's stream. */
int zgr_recover(zgr_ctx *ctx, const uint32_t *row_ids, const uint8_t *rows, size_t n_rows,
                uint8_t *raw_out, size_t raw_cap, size_t *raw_len);

/* ---- autotuning (no counterpart in the reference, whose GPU path hard-codes its window,
 *      prove_gpu.rs:16-17) ----
 * Times, on the device, the candidate chunk-group plans of the fixed-base bucket kernels
 * (how many chunks of an AMT line share one bucket set) for each density class of a
 * coset's scalars (coset 0 of a short blob has zero chunks; cosets 1 and 2 are dense),
 * and the split of the context's two-pass NTT; keeps the fastest of each for the encode
 * path and zgr_amt_prove.  Every choice gives the same outputs.  Needs coset 0's params.
 * report (may be NULL) receives a JSON object of the timings and choices (truncated to
 * cap - 1 bytes).  Without a call the defaults hold. */
int zgr_autotune(zgr_ctx *ctx, char *report, size_t cap);

/* ---- timing hooks for the benchmark (HIP events on the library's own stream) ---- */
/* Average device time in ms of the named kernel over the calls since the last reset
 * (name: "msm_accum", "ntt", "merkle", ...); count receives the launch count. */
double zgr_kernel_time_ms(zgr_ctx *ctx, const char *name, int *count);
void zgr_kernel_time_reset(zgr_ctx *ctx);
void zgr_set_profiling(zgr_ctx *ctx, int enabled);

#ifdef __cplusplus
}
#endif
#endif
