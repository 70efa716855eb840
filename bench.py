"""Benchmark: encoded 1024x1024 blobs/s on one MI355X (plain and batched), p50/p99
per-blob latency, the roofline of the dominant kernel (the MSM bucket kernel) and
the CPU oracle timed on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]

A step is one encode of B seeded full-size blobs (raw bytes already resident in
HBM) through the C ABI, outputs (encoded data, commitment, AMT proofs, Merkle
tree) copied back to host memory.  Also times RS recovery of the same blob from its
rows, the deferred-verification MSM over its rows and a 2^20-point EC-FFT (device,
and the oracle where it finishes in seconds).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "0g-da-encoder_amd"))

import numpy as np  # noqa: E402

import zg_rocm as Z  # noqa: E402

TAU_SEED = 7
DATA_SEED = 222
# Algorithmic work of one bucket-kernel launch (one coset of one blob): 12 lines x
# 2^20 scalar-point pairs x 32 signed 8-bit windows, each a mixed addition counted
# as 10 Fq multiplications of 136 32x32->64 multiply-accumulates (8-limb CIOS).
MAC32_PER_PAIR_WINDOW = 10 * 136
# VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3 T lane-ops/s,
# i.e. one v_mad_u64_u32 (32x32->64 MAC) per lane per clock.  (A dependent-chain
# microbenchmark of v_mad_u64_u32 alone reaches 25.0 T/s on the GPU.)
PEAK_MAC32 = 256 * 4 * 16 * 2.4e9
# dense int8 MFMA peak of MI355X: 5.0 POP/s = 2.5 x 10^15 multiply-accumulates per second
PEAK_INT8_MAC = 2.5e15
# HBM bytes of one coset's bucket stage (the sorts, k_fb_sums, k_fb_reduce, k_fw_combine)
# from two rocprofv3 counter passes over `bench.py --roofline-only`: FETCH_SIZE
# 19,562,518 KiB x 2.0 (a 1 GiB device copy reads as 524,316 KiB) + WRITE_SIZE
# 3,064,418 KiB x 1.0 (the copy writes as 1,048,576 KiB).
MSM_TRAFFIC_BYTES = (19562518 * 2.0 + 3064418) * 1024
# VALU lane-instructions of the same stage: rocprofv3 SQ_INSTS_VALU (wave instructions)
# over the same command, x 64 lanes.
VALU_LANE_OPS = 9909242770 * 64.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def roofline_scalars(geo):
    """Dense scalars for the isolated bucket-kernel launches: uniform Montgomery
    residues below 3 * 2^252 < r (seeded)."""
    scal = np.random.default_rng(5).integers(0, 2**63, (geo.n, 4), dtype=np.uint64)
    scal[:, 3] &= np.uint64(0x2FFFFFFFFFFFFFFF)
    return scal


_HALF_P = np.array([((Z.P_MOD - 1) // 2 >> (64 * k)) & (2**64 - 1) for k in range(4)], dtype=np.uint64)


def ppot_encode(pts, g2):
    """Synthetic ceremony input: canonical affine points (the library's memory form) in
    the compressed encodings of a response file -- big-endian x (G2: x.c1 then x.c0),
    0x80 in the first byte when y is the larger of y and -y (G2: by c1, then c0)."""
    def larger(le):  # (n, 32) little-endian integers > (p - 1) / 2
        w = np.ascontiguousarray(le).view("<u8")
        gt, eq = np.zeros(len(w), dtype=bool), np.ones(len(w), dtype=bool)
        for k in (3, 2, 1, 0):
            gt |= eq & (w[:, k] > _HALF_P[k])
            eq &= w[:, k] == _HALF_P[k]
        return gt
    if g2:
        out = np.concatenate([pts[:, 32:64][:, ::-1], pts[:, 0:32][:, ::-1]], axis=1)
        flag = np.where(pts[:, 96:128].any(axis=1), larger(pts[:, 96:128]), larger(pts[:, 64:96]))
    else:
        out = np.ascontiguousarray(pts[:, 0:32][:, ::-1])
        flag = larger(pts[:, 32:64])
    out[:, 0] |= np.where(flag, 0x80, 0).astype(np.uint8)
    return out.tobytes()


def cpu_baseline(ctx, geo, raw, threads):
    """The oracle -- a CPU restatement of the reference `parallel` path (to_coset_blob,
    gen_all_proofs, EncodedBlobMerkle::build) -- on bounded samples of one 1024x1024 blob.

    All host cores: coset 0 (identity extension + gen_all_proofs over 12 x 2^20 pairs)
    and coset 1 (iFFT / scale / FFT of 2^20 + gen_all_proofs) timed separately, per blob
    = coset 0 + 2 x coset 1 + the Merkle tree over 3 x 2^20 elements.
    One core: coset 1's extension, a quarter of its basis line (256 chunk MSMs of 1024
    points, 2^18 pairs) scaled to 12 lines x 2^20 pairs per coset, and the Merkle tree
    over an eighth of the elements, scaled x8."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    blob = O.raw_to_blob(raw, geo.log_n).reshape(geo.row_n, geo.col_n, 4)
    pts = np.ascontiguousarray(blob.transpose(1, 0, 2)).reshape(-1, 4)
    idx = np.arange(geo.n)
    rev = np.zeros(geo.n, dtype=np.int64)
    for b in range(geo.log_n):
        rev |= ((idx >> b) & 1) << (geo.log_n - 1 - b)
    data = np.random.default_rng(1).integers(0, 256, geo.coset_n * geo.n * 32, dtype=np.uint8)

    O.set_threads(threads)
    t_c = []
    for c in (0, 1):
        bases = ctx.params_download(c)
        t0 = time.time()
        y = pts if c == 0 else O.to_coset_blob(pts, geo.log_n, c)
        O.amt_prove(bases, geo.log_n, geo.h, np.ascontiguousarray(y.reshape(-1, 4)[rev]))
        t_c.append(time.time() - t0)
    t0 = time.time()
    O.merkle(data, geo.coset_n, geo.log_row)
    t_merkle = time.time() - t0
    per_blob = t_c[0] + (geo.coset_n - 1) * t_c[1] + t_merkle

    O.set_threads(1)
    t0 = time.time()
    y = O.to_coset_blob(pts, geo.log_n, 1)
    t_ext1 = time.time() - t0
    bases = ctx.params_download(1).reshape(geo.h + 2, geo.n, 8)[geo.h]
    canon = np.ascontiguousarray(O.fr_canonical_array(np.ascontiguousarray(y.reshape(-1, 4)[rev]))).reshape(-1, 4)
    chunks, bs = 256, geo.n >> geo.h
    t0 = time.time()
    for b in range(chunks):
        O.msm(bases[b * bs:(b + 1) * bs], canon[b * bs:(b + 1) * bs])
    t_chunks = time.time() - t0
    eighth = geo.coset_n * geo.n // 8
    t0 = time.time()
    O.merkle(data[:eighth * 32], geo.coset_n, geo.log_row - 3)
    t_m1 = time.time() - t0
    prove1 = t_chunks * (geo.n // (chunks * bs)) * (geo.h + 2)
    per_blob_1 = (geo.coset_n - 1) * t_ext1 + geo.coset_n * prove1 + 8 * t_m1
    O.set_threads(threads)
    return {
        "value": 1.0 / per_blob,
        "unit": "blobs/s",
        "cores": threads,
        "kind": "port",
        "label": "CPU restatement of the reference parallel path (the oracle, C + OpenMP)",
        "sample": "one 1024x1024 blob: coset 0 (gen_all_proofs, 12 x 2^20 pairs) %.1f s, coset 1 (extension + "
                  "gen_all_proofs) %.1f s, Merkle over 3 x 2^20 elements %.1f s; per blob = coset 0 + 2 x coset 1 "
                  "+ Merkle" % (t_c[0], t_c[1], t_merkle),
        "one_core": {"value": 1.0 / per_blob_1, "unit": "blobs/s", "cores": 1,
                     "sample": "coset 1 extension %.1f s; 256 of the basis line's 1024-point chunk MSMs %.1f s, "
                               "scaled to 12 lines x 2^20 pairs per coset (%.0f s per coset); Merkle over 1/8 of "
                               "the elements %.1f s, x8" % (t_ext1, t_chunks, prove1, t_m1)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--served-requests", type=int, default=48)
    ap.add_argument("--qps", type=float, default=0.0)
    ap.add_argument("--roofline-only", action="store_true")
    ap.add_argument("--msm-only", action="store_true",
                    help="only the isolated 2^20 MSM (zgr_msm), four calls, for a kernel trace")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the encoder is built and measured on a single "
              "MI355X (blobs shard over GPUs with no exchange between them). --gpus %d is not run."
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    geo = Z.PRODUCTION
    ctx = Z.Context(geo, device=0)
    t0 = time.time()
    ctx.params_generate(Z.tau_from_seed(TAU_SEED))
    t_params = time.time() - t0
    log("params (3 cosets x 12 x 2^20 bases) generated on device in %.2f s" % t_params)
    if args.msm_only:
        # one warm-up and three timed calls of zgr_msm over 2^20 basis-line bases and dense
        # scalars (the same inputs as the msm_2pow20 field of the full run)
        mbases = ctx.params_download(0).reshape(geo.h + 2, geo.n, 8)[geo.h].copy()
        mscal = roofline_scalars(geo)
        ctx.msm(mbases, mscal)
        mt = []
        for _ in range(3):
            s0 = time.perf_counter()
            ctx.msm(mbases, mscal)
            mt.append(time.perf_counter() - s0)
        ctx.close()
        print(json.dumps({"msm_only": True, "n": geo.n, "calls": 4, "wall_ms": round(float(np.median(mt)) * 1e3, 2)}))
        return
    if args.roofline_only:
        # workload for the counter passes: three isolated launches of the bucket kernels
        # (one coset of amt_prove each, the default plan: the tuner does not run, since
        # its timings under a profiler could pick another plan) and one 1 GiB
        # device-to-device copy whose FETCH_SIZE / WRITE_SIZE calibrate the counters
        scal = roofline_scalars(geo)
        for _ in range(3):
            ctx.amt_prove(1, scal)
        # and three of each NTT shape of the bench (k_ntt_pass, k_ntt_unrev)
        nrows = np.random.default_rng(DATA_SEED + 11).integers(0, 2**63, (1024, 1 << 10, 4), dtype=np.uint64)
        nrows[:, :, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
        ndev = Z.DeviceBuffer(nrows.tobytes())
        for log_n, batch in ((10, 1024), (20, 1)):
            for _ in range(3):
                ctx.ntt_device(ndev.address, log_n, batch)
        del ndev
        import ctypes
        pa, pb = ctypes.c_void_p(), ctypes.c_void_p()
        Z.hip().hipMalloc(ctypes.byref(pa), 1 << 30)
        Z.hip().hipMalloc(ctypes.byref(pb), 1 << 30)
        Z.hip().hipMemcpy(pb, pa, 1 << 30, 3)  # hipMemcpyDeviceToDevice
        Z.hip().hipDeviceSynchronize()
        Z.hip().hipFree(pa)
        Z.hip().hipFree(pb)
        ctx.close()
        print(json.dumps({"roofline_only": True, "msm_launches": 3, "ntt_calls": {"1024x2^10": 3, "2^20": 3},
                          "copy_bytes": 1 << 30}))
        return
    # the in-project tuner: chunk-group plan per scalar-density class, NTT split
    t0 = time.time()
    tune = ctx.autotune()
    tune["seconds"] = round(time.time() - t0, 2)
    log("autotune: %s" % tune)

    def resident(raws):
        dev = [Z.DeviceBuffer(r) for r in raws]
        return dev, [d.address for d in dev], [len(r) for r in raws]

    def timed_steps(step, steps, warmup, profile=False):
        for _ in range(warmup):
            step()
        if profile:
            ctx.set_profiling(True)
            ctx.kernel_time_reset()
        lat = []
        Z.hip().hipDeviceSynchronize()
        t_start = time.perf_counter()
        for _ in range(steps):
            s = time.perf_counter()
            step()
            lat.append(time.perf_counter() - s)
        Z.hip().hipDeviceSynchronize()
        elapsed = time.perf_counter() - t_start
        if profile:
            ctx.set_profiling(False)
        return elapsed, lat

    # plain path: one full-size blob per call (the headline configuration)
    raws = [Z.seeded_bytes(DATA_SEED, geo.max_raw_data_size)]
    keep, ptrs, lens = resident(raws)
    out1 = Z.EncodedBlobOutputs(geo)
    elapsed, lat = timed_steps(lambda: ctx.encode_device(ptrs[0], lens[0], out=out1), args.steps, args.warmup)
    value = args.steps / elapsed
    # per-kernel event times inside the plain pipeline (overlapping streams), from a
    # separate pass: the event timers sit between the launches (and keep the MSM stage
    # out of its captured graphs), so they are not in the timed steps above
    timed_steps(lambda: ctx.encode_device(ptrs[0], lens[0], out=out1), 3, 0, profile=True)
    kernels = {}
    for name in ("pack", "ntt", "coset_scale", "msm_sort", "msm_scatter", "msm_group_scatter", "msm_wide_sort",
                 "msm_accum", "msm_reduce", "msm_combine", "msm_tree", "msm_horner", "merkle_leaves", "merkle_level",
                 "file_root"):
        ms, cnt = ctx.kernel_time_ms(name)
        if cnt:
            kernels[name] = {"avg_ms": round(ms, 4), "launches": cnt}

    # RS recovery (recover_from_da_slice) of the plain blob from its encoded rows, host
    # rows in and raw bytes out: the 1024 rows of coset 2 (every coset-0 row erased,
    # the most work) and all 3072 rows (the consistency check over 2048 extra rows)
    w = 32 * geo.col_n
    rec_sets = {"coset2_1024_rows": list(range(2 * geo.row_n, 3 * geo.row_n)),
                "all_3072_rows": list(range(geo.row_encoded))}
    recovery = {}
    for label, ids in rec_sets.items():
        rows_b = b"".join(out1.data[i * w:(i + 1) * w].tobytes() for i in ids)
        assert ctx.recover(ids, rows_b) == raws[0]
        rows_a = np.frombuffer(rows_b, dtype=np.uint8)
        out_a = np.empty(geo.max_raw_data_size, dtype=np.uint8)
        pin_rows, pin_out = Z.PinnedHostArray(len(rows_b)), Z.PinnedHostArray(geo.max_raw_data_size)
        pin_rows.array[:] = rows_a
        timing = {}
        for kind, (ra, oa) in (("pageable", (rows_a, out_a)), ("pinned", (pin_rows.array, pin_out.array))):
            n = ctx.recover_into(ids, ra, oa)
            assert oa[:n].tobytes() == raws[0]
            ctx.set_profiling(True)
            ctx.kernel_time_reset()
            tt = []
            for _ in range(5):
                s0 = time.perf_counter()
                ctx.recover_into(ids, ra, oa)
                tt.append(time.perf_counter() - s0)
            ctx.set_profiling(False)
            timing[kind] = round(float(np.median(tt)) * 1e3, 2)
            if kind == "pageable":
                ks = {}
                for name in ("recover", "recover_rows", "recover_lagrange", "recover_planes", "recover_gemm",
                             "recover_fold"):
                    ms, cnt = ctx.kernel_time_ms(name)
                    if cnt:
                        ks[name] = round(ms * cnt / 5, 3)  # per call, summed over its launches
        pin_rows.free()
        pin_out.free()
        recovery[label] = {"ms": timing["pageable"], "ms_pinned_host_buffers": timing["pinned"],
                           "kernel_ms_per_call": ks}
        erased = sum(1 for i in range(geo.row_n) if i not in set(ids))
        if erased and ks.get("recover_gemm"):
            # the interpolation GEMM on the matrix cores: 32 x 32 int8 digit-plane products
            # of (erased rows x ROW) by (ROW x COL), padded to multiples of 32
            pad = lambda v: (v + 31) // 32 * 32  # noqa: E731
            macs = 32 * 32 * pad(erased) * pad(geo.row_n) * pad(geo.col_n)
            recovery[label]["mfma_gemm"] = {
                "int8_macs": macs, "ms": ks["recover_gemm"],
                "achieved_TMAC_s": round(macs / (ks["recover_gemm"] * 1e-3) / 1e12, 1),
                "frac_of_dense_int8_peak": round(macs / (ks["recover_gemm"] * 1e-3) / PEAK_INT8_MAC, 3)}
    recovery["note"] = ("zgr_recover over host arrays (rows in and raw bytes out cross PCIe inside the timed "
                        "call): pageable numpy arrays, and page-locked (hipHostMalloc) ones; median of 5")

    # deferred verification MSM (zgr_verify_msm): every row of the plain blob, one random
    # alpha each -- one 2^20-point MSM per coset over the resident basis
    vrng = np.random.default_rng(DATA_SEED + 3)
    v_ids = list(range(geo.row_encoded))
    v_rows = out1.data.tobytes()
    v_alphas = [int.from_bytes(vrng.bytes(32), "little") % Z.R_MOD for _ in v_ids]
    ctx.verify_msm(v_ids, v_rows, v_alphas)
    vt = []
    for _ in range(3):
        s0 = time.perf_counter()
        ctx.verify_msm(v_ids, v_rows, v_alphas)
        vt.append(time.perf_counter() - s0)
    pin_v = Z.PinnedHostArray(len(v_rows))
    pin_v.array[:] = np.frombuffer(v_rows, dtype=np.uint8)
    assert ctx.verify_msm(v_ids, pin_v.array, v_alphas) == ctx.verify_msm(v_ids, v_rows, v_alphas)
    vp = []
    for _ in range(3):
        s0 = time.perf_counter()
        ctx.verify_msm(v_ids, pin_v.array, v_alphas)
        vp.append(time.perf_counter() - s0)
    pin_v.free()
    verify = {"rows": len(v_ids), "ms": round(float(np.median(vt)) * 1e3, 2),
              "ms_pinned_host_rows": round(float(np.median(vp)) * 1e3, 2),
              "note": "zgr_verify_msm over 3 x 1024 rows (100 MB of host rows in: pageable bytes, and a "
                      "page-locked array), median of 3"}

    # G1 EC-FFT (zgr_ecfft) over 2^20 points: the powers [tau^i] G1 of the seeded tau,
    # as a ceremony gives them (distinct points; a tiled input would make most
    # butterflies' v infinity, which the stages skip by the wave)
    ctx.power_tau_setup(Z.tau_from_seed(TAU_SEED))
    ec_in = np.ascontiguousarray(ctx.power_tau_download()[0]).reshape(-1)
    ctx.set_profiling(True)
    ctx.kernel_time_reset()
    s0 = time.perf_counter()
    ctx.ecfft(ec_in, geo.log_n, inverse=False)
    ec_wall = time.perf_counter() - s0
    ec_ms, ec_cnt = ctx.kernel_time_ms("ecfft")
    ctx.set_profiling(False)
    butterflies = geo.log_n * geo.n // 2
    # stage 1's twiddles are all w^0 = 1 (no multiplication), so the scalar
    # multiplications are those of stages 2..log_n
    twiddled = (geo.log_n - 1) * geo.n // 2
    ec_s = ec_ms * ec_cnt * 1e-3
    ecfft = {"log_n": geo.log_n, "wall_ms": round(ec_wall * 1e3, 1), "kernel_ms": round(ec_ms * ec_cnt, 1),
             "butterflies_per_s": round(butterflies / ec_s) if ec_cnt else None,
             # reference-equivalent work: a 254-bit double-and-add scalar multiplication per
             # twiddled butterfly (254 doublings of 6 Fq multiplications, 127 additions of
             # 14), 136 MAC32 per Fq multiplication, against the 39.3 T MAC32/s issue peak
             "roofline": {"bound": "valu", "mac32_per_butterfly": (254 * 6 + 127 * 14) * 136,
                          "twiddled_butterflies": twiddled,
                          "frac": round(twiddled * (254 * 6 + 127 * 14) * 136 / ec_s / PEAK_MAC32, 4)
                          if ec_cnt else None},
             "note": "one forward DFT of 2^20 G1 points (Pairing::fast_fft); 20 x 2^19 butterflies, those of "
                     "stages 2..20 with a scalar multiplication by a GLV-split twiddle (stage 1's are all 1)"}

    # the isolated primitives at production shape, inputs resident in HBM where the entry
    # point allows it.
    # NTT: the batched 1024 rows x 2^10 (one launch, each row in LDS) and one 2^20-point
    # transform (two passes over HBM + the return to natural order); k_ntt_pass is the
    # "ntt" timer, the bit-reversal copy "ntt_unrev"
    nrng = np.random.default_rng(DATA_SEED + 11)
    nrows = nrng.integers(0, 2**63, (1024, 1 << 10, 4), dtype=np.uint64)
    nrows[:, :, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
    ndev = Z.DeviceBuffer(nrows.tobytes())
    ntt = {}
    for label, log_n, batch in (("batch_1024_rows_2pow10", 10, 1024), ("one_2pow20", 20, 1)):
        ctx.ntt_device(ndev.address, log_n, batch)  # warm
        ctx.set_profiling(True)
        ctx.kernel_time_reset()
        Z.hip().hipDeviceSynchronize()
        reps = 10
        s0 = time.perf_counter()
        for _ in range(reps):
            ctx.ntt_device(ndev.address, log_n, batch)
        wall = (time.perf_counter() - s0) / reps
        ctx.set_profiling(False)
        k_ms, k_cnt = ctx.kernel_time_ms("ntt")
        u_ms, u_cnt = ctx.kernel_time_ms("ntt_unrev")
        pass_ms = k_ms * k_cnt / reps  # k_ntt_pass time per call (1 or 2 launches)
        elems = batch << log_n
        butterflies = batch * log_n * (1 << (log_n - 1))
        ntt[label] = {
            "log_n": log_n, "rows": batch, "wall_ms": round(wall * 1e3, 4), "pass_kernel_ms": round(pass_ms, 4),
            "passes": k_cnt // reps, "unrev_ms": round(u_ms * u_cnt / reps, 4) if u_cnt else 0.0,
            # compulsory bytes of the passes: each pass reads and writes every element once
            "pass_GB_s": round(2 * 32 * elems * (k_cnt // reps) / (pass_ms * 1e-3) / 1e9, 1),
            "butterflies_per_s": round(butterflies / (pass_ms * 1e-3)),
            # VALU bound: one Fr Montgomery product (9 x 9 + 9 x 9 = 162 v_mad_u64_u32 in the
            # 29-bit form) per butterfly, against the 39.3 T lane-op/s issue peak
            "valu_mad_frac": round(butterflies * 162 / (pass_ms * 1e-3) / PEAK_MAC32, 4),
        }
    del ndev
    ntt["note"] = ("zgr_ntt_device (natural order in and out, forward), mean of 10 calls; the encode path runs "
                   "the same passes without the return to natural order (DIF iNTT, then DIT per coset)")

    # MSM of 2^20 (zgr_msm: host bases and scalars in, GLV halves through the wide sort and
    # bucket kernels) and the 2^10-point EC-FFT (zgr_ecfft)
    mbases = ctx.params_download(0).reshape(geo.h + 2, geo.n, 8)[geo.h].copy()
    mscal = roofline_scalars(geo)
    ctx.msm(mbases, mscal)
    ctx.set_profiling(True)
    ctx.kernel_time_reset()
    mt = []
    for _ in range(3):
        s0 = time.perf_counter()
        ctx.msm(mbases, mscal)
        mt.append(time.perf_counter() - s0)
    ctx.set_profiling(False)
    mk = {}
    for name in ("msm_inf", "msm_pm_table", "msm_wide_sort", "msm_accum", "msm_reduce", "msm_combine", "msm_root"):
        a, cnt = ctx.kernel_time_ms(name)
        if cnt:
            mk[name] = round(a * cnt / 3, 3)
    # the same call from page-locked host arrays (the copies at the link's rate)
    pin_b, pin_s = Z.PinnedHostArray(mbases.nbytes), Z.PinnedHostArray(mscal.nbytes)
    pin_b.array[:] = np.ascontiguousarray(mbases).view(np.uint8).reshape(-1)
    pin_s.array[:] = np.ascontiguousarray(mscal).view(np.uint8).reshape(-1)
    pb, ps = pin_b.array.view(np.uint64), pin_s.array.view(np.uint64)
    assert np.array_equal(ctx.msm(pb, ps), ctx.msm(mbases, mscal))
    mpt = []
    for _ in range(3):
        s0 = time.perf_counter()
        ctx.msm(pb, ps)
        mpt.append(time.perf_counter() - s0)
    pin_b.free()
    pin_s.free()
    msm = {"n": geo.n, "wall_ms": round(float(np.median(mt)) * 1e3, 2),
           "wall_ms_pinned_host": round(float(np.median(mpt)) * 1e3, 2), "kernel_ms": round(sum(mk.values()), 3),
           "kernels": mk,
           "note": "zgr_msm, one MSM of 2^20 (basis line bases x dense scalars); the call uploads 64 MB of bases "
                   "and 32 MB of scalars, splits each scalar by GLV into two 127-bit halves over P and phi(P), and "
                   "runs the wide sort and bucket kernels over 8 windows of 16 bits; wall: median of 3 from pageable "
                   "numpy arrays, wall_ms_pinned_host: the same from page-locked arrays, kernel_ms: mean per call of "
                   "the pageable calls (HIP events)"}
    ec10 = ec_in[:(1 << 10) * 64]
    ctx.ecfft(ec10, 10, inverse=False)
    ctx.set_profiling(True)
    ctx.kernel_time_reset()
    et = []
    for _ in range(3):
        s0 = time.perf_counter()
        ctx.ecfft(ec10, 10, inverse=False)
        et.append(time.perf_counter() - s0)
    e_ms, e_cnt = ctx.kernel_time_ms("ecfft")
    ctx.set_profiling(False)
    ecfft["log_n_10"] = {"wall_ms": round(float(np.median(et)) * 1e3, 3), "kernel_ms": round(e_ms * e_cnt / 3, 3)}

    # plain encode from host memory (zgr_encode: the 32.5 MB raw blob crosses PCIe inside
    # the call) -- the PCIe-inclusive rate, never the headline
    out_h = Z.EncodedBlobOutputs(geo)
    ctx.encode(raws[0])
    ht = []
    for _ in range(max(3, args.steps // 2)):
        s0 = time.perf_counter()
        lib_out = out_h.cstruct()
        Z._check(Z.lib().zgr_encode(ctx.h, Z._p(np.frombuffer(raws[0], dtype=np.uint8)), len(raws[0]),
                                    Z.ctypes.byref(lib_out)))
        ht.append(time.perf_counter() - s0)
    host_input = {"value": round(len(ht) / sum(ht), 4), "unit": "blobs/s", "ms_p50": round(float(np.median(ht)) * 1e3, 2),
                  "note": "zgr_encode from host memory, outputs to host: the PCIe-inclusive plain rate"}

    # batched path: args.batch full-size blobs per call (two blobs in flight)
    nb_ = max(2, args.batch)
    braws = [Z.seeded_bytes(DATA_SEED + 1 + i, geo.max_raw_data_size) for i in range(nb_)]
    bkeep, bptrs, blens = resident(braws)
    bouts = [Z.EncodedBlobOutputs(geo) for _ in braws]
    bsteps = max(1, args.steps // nb_ + 1)
    b_el, b_lat = timed_steps(lambda: ctx.encode_batch_device(bptrs, blens, bouts), bsteps, 1)
    batched = {"blobs_per_call": nb_, "calls": bsteps, "value": round(bsteps * nb_ / b_el, 4), "unit": "blobs/s",
               "ms_per_blob": round(b_el / (bsteps * nb_) * 1e3, 3)}
    del bkeep, bouts

    # uneven mix: 32 blobs, raw lengths log-uniform in [1 KB, MAX_RAW_DATA_SIZE] (seeded)
    rng = np.random.default_rng(DATA_SEED)
    mlens = np.exp(rng.uniform(np.log(1024), np.log(geo.max_raw_data_size), 32)).astype(np.int64)
    mraws = [Z.seeded_bytes(DATA_SEED + 100 + i, int(L)) for i, L in enumerate(mlens)]
    mkeep, mptrs, mlens_ = resident(mraws)
    mouts = [Z.EncodedBlobOutputs(geo) for _ in mraws]
    m_el, _ = timed_steps(lambda: ctx.encode_batch_device(mptrs, mlens_, mouts), 1, 1)
    # the same 32 blobs one call each (raw bytes resident too): the plain path on this mix
    m_pl, _ = timed_steps(lambda: [ctx.encode_device(p_, l_, out=o_) for p_, l_, o_ in zip(mptrs, mlens_, mouts)], 1, 0)
    mixed = {"blobs": 32, "value": round(32 / m_el, 4), "unit": "blobs/s",
             "raw_MB_total": round(float(mlens.sum()) / 1e6, 1),
             "plain_per_call_value": round(32 / m_pl, 4), "batched_over_plain": round(m_pl / m_el, 3)}
    del mkeep, mouts

    # served: Poisson arrivals at a fixed QPS with the same size mix, through the
    # continuous-batching server (host input, <= 64 in flight), against the plain
    # per-call path fed the same arrivals one at a time.  Two loads: 0.6x and 1.1x the
    # plain per-call capacity on this mix (measured first: the mix encoded one call at a
    # time from host memory).
    n_req = args.served_requests
    sraws = [mraws[i % len(mraws)] for i in range(n_req)]
    s0 = time.perf_counter()
    for r in mraws:
        ctx.encode(r)
    plain_capacity = len(mraws) / (time.perf_counter() - s0)

    def served(arr):
        srv = Z.Server(ctx, hbm_budget=4 << 30, max_inflight=64)
        lat_s, pend, i = [], {}, 0
        t0 = time.perf_counter()
        while i < n_req or pend:
            now = time.perf_counter() - t0
            while i < n_req and arr[i] <= now:
                pend[srv.submit(sraws[i])] = arr[i]
                i += 1
            for t in list(pend):
                if srv.poll(t, 0) is not None:
                    lat_s.append(time.perf_counter() - t0 - pend.pop(t))
            time.sleep(0.0005)
        wall = time.perf_counter() - t0
        srv.close()
        return wall, lat_s

    def plain_fifo(arr):
        lat_p = []
        t0 = time.perf_counter()
        for i in range(n_req):
            now = time.perf_counter() - t0
            if arr[i] > now:
                time.sleep(arr[i] - now)
            ctx.encode(sraws[i])
            lat_p.append(time.perf_counter() - t0 - arr[i])
        return time.perf_counter() - t0, lat_p

    pct = lambda v, q: round(float(np.percentile(np.array(v) * 1e3, q)), 1)  # noqa: E731
    serving = {"requests_per_load": n_req, "plain_capacity_on_mix": round(plain_capacity, 3), "loads": []}
    for k, load in enumerate((args.qps / plain_capacity,) if args.qps > 0 else (0.6, 1.1)):
        qps = load * plain_capacity
        arr = np.cumsum(np.random.default_rng(DATA_SEED + 7 + k).exponential(1.0 / qps, n_req))
        s_wall, s_lat = served(arr)
        p_wall, p_lat = plain_fifo(arr)
        serving["loads"].append({
            "load": round(load, 2), "qps": round(qps, 3), "value": round(n_req / s_wall, 4), "unit": "blobs/s",
            "p50_ms": pct(s_lat, 50), "p99_ms": pct(s_lat, 99),
            "plain_per_call_same_arrivals": {"value": round(n_req / p_wall, 4), "p50_ms": pct(p_lat, 50),
                                             "p99_ms": pct(p_lat, 99)}})

    # bin/build_params (crates/amt/src/bin/build_params.rs:24-53) in create mode on the
    # device: powers of tau, then per coset the prove bases and vanishes from them
    # (AMTParams::from_pp: 12 G1 EC-FFTs of 2^20 and the G2 vanishes DFTs) and the images
    # of the reference's four files (in host memory: no disk in the timed region)
    tb = {}
    t0 = time.perf_counter()
    ctx.power_tau_setup(Z.tau_from_seed(TAU_SEED))
    tb["power_tau_s"] = round(time.perf_counter() - t0, 3)
    img_bytes = 0
    for c in range(geo.coset_n):
        s0 = time.perf_counter()
        ctx.params_from_power_tau(c)
        tb.setdefault("from_pp_s", []).append(round(time.perf_counter() - s0, 3))
        s0 = time.perf_counter()
        for kind in (Z.FILE_AMT, Z.FILE_AMT_MONT, Z.FILE_AMT_VERIFY) + ((Z.FILE_PTAU,) if c == 0 else ()):
            img_bytes += len(ctx.write_params_file(kind, c))
        tb.setdefault("file_images_s", []).append(round(time.perf_counter() - s0, 3))
    tb["total_s"] = round(time.perf_counter() - t0, 3)
    tb["file_bytes"] = img_bytes
    tb["note"] = ("depth 20, prove/verify depth 10, 3 cosets; the reference quotes about 30 minutes for its "
                  "params build on an AWS g5.4xlarge (README.md:121)")
    params_build = tb

    # ceremony import (crates/ppot2ark/src/main.rs:45-119): the production read of
    # 2^20 G1 and 2^20 G2 powers from a response file (compressed points), decoded and
    # checked on the device -- square roots, and [r] P == O for every G2 point
    g1pp, g2pp, _, _ = ctx.power_tau_download()
    ppot = {}
    for label, kind, enc in (("g1_compressed_2pow20", Z.PPOT_G1_COMPRESSED, ppot_encode(g1pp, False)),
                             ("g2_compressed_2pow20", Z.PPOT_G2_COMPRESSED, ppot_encode(g2pp, True))):
        ctx.set_profiling(True)
        ctx.kernel_time_reset()
        s0 = time.perf_counter()
        out = ctx.ppot_decode(kind, enc, geo.n)
        wall = time.perf_counter() - s0
        ctx.set_profiling(False)
        k_ms, k_cnt = ctx.kernel_time_ms("ppot")
        assert np.array_equal(out, g2pp if kind == Z.PPOT_G2_COMPRESSED else g1pp)
        ppot[label] = {"wall_ms": round(wall * 1e3, 1), "kernel_ms": round(k_ms * k_cnt, 1)}
    ppot["note"] = ("zgr_ppot_decode of 2^20 compressed points from host bytes (copies included in wall_ms); "
                    "ppot2ark reads 6 x 2^20 G1 and 2 x 2^20 G2 points this way for a depth-20 PowerTau")
    del g1pp, g2pp

    # The encode pipeline overlaps kernels of different cosets on two streams, which
    # stretches event-timed durations; the roofline uses isolated launches of the same
    # bucket kernels (amt_prove of one coset: the encode path's multi-MSM over 12 lines
    # x 2^20 dense scalars).
    scal = roofline_scalars(geo)
    ctx.amt_prove(1, scal)
    ctx.set_profiling(True)
    ctx.kernel_time_reset()
    for _ in range(3):
        ctx.amt_prove(1, scal)
    ctx.set_profiling(False)
    def per_call(name):  # total time of the kernels named `name` per amt_prove call
        ms, cnt = ctx.kernel_time_ms(name)
        return ms * cnt / 3
    sort_ms = per_call("msm_sort") + per_call("msm_group_scatter") + per_call("msm_wide_sort")
    sums_ms, red_ms, comb_ms = per_call("msm_accum"), per_call("msm_reduce"), per_call("msm_combine")
    acc_cnt = 3
    acc_ms = sort_ms + sums_ms + red_ms + comb_ms
    work = 12 * geo.n * 32 * MAC32_PER_PAIR_WINDOW  # per launch pair (one coset)
    achieved = work / (acc_ms * 1e-3) / 1e12 if acc_ms else None
    peak = PEAK_MAC32 / 1e12
    p50 = float(np.percentile(np.array(lat) * 1e3, 50))
    p99 = float(np.percentile(np.array(lat) * 1e3, 99))

    res = {
        "metric": "encoded 1024x1024 blobs/s on 1 GPU, batched vs plain; p50/p99 encode latency",
        "value": round(value, 4),
        "unit": "blobs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (exact BN254 Fr/Fq Montgomery integer arithmetic)",
        "data": "synthetic: seeded uniform bytes, 32,505,852 B per blob (MAX_RAW_DATA_SIZE); "
                "params from a seeded tau (insecure test setup)",
        "config": {"workload": "plain single-blob encode, 1024x1024 Fr, 3 cosets, AMT depth 10, one blob per call",
                   "rows": geo.row_n, "cols": geo.col_n, "cosets": geo.coset_n, "amt_depth": geo.h,
                   "blobs_per_step": 1},
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3)},
        "batched": batched,
        "uneven_mix_32": mixed,
        "served": serving,
        "params_generate_s": round(t_params, 3),
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 4) if achieved else None,
            "peak": round(peak, 3),
            "unit": "TOP/s (32x32->64 integer MAC)",
            "frac": round(achieved / peak, 4) if achieved else None,
            "traffic": round(MSM_TRAFFIC_BYTES),
            "traffic_unit": "bytes per bucket stage (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
            "kernel": "bucket stage of one coset's AMT multi-MSM: the digit sorts (k_fb_count, k_fb_scatter, "
                      "k_fw_*), bucket sums (k_fb_sums), bucket reductions (k_fb_reduce) and wide-group "
                      "combine (k_fw_combine)",
            "avg_ms": round(acc_ms, 3),
            "kernel_ms": {"sorts": round(sort_ms, 3), "k_fb_sums": round(sums_ms, 3), "k_fb_reduce": round(red_ms, 3),
                          "k_fw_combine": round(comb_ms, 3)},
            "launches": acc_cnt,
            # VALU lane-instructions the stage actually issues (rocprofv3 SQ_INSTS_VALU x 64)
            # over the time, against the same 39.3 T lane-op/s issue peak
            "valu_issue_frac": round(VALU_LANE_OPS / (acc_ms * 1e-3) / PEAK_MAC32, 4) if acc_ms else None,
        },
        "kernels": kernels,
        "recovery": recovery,
        "verify_msm": verify,
        "ecfft": ecfft,
        "ntt": ntt,
        "msm_2pow20": msm,
        "plain_host_input": host_input,
        "params_build": params_build,
        "ppot_import": ppot,
        "autotune": tune,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(ctx, geo, raws[0], args.cpu_threads)
        # the oracle's recovery (the reference's 4N-domain algorithm) on the coset-2 rows
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.set_threads(args.cpu_threads)
        ids = rec_sets["coset2_1024_rows"]
        t0 = time.time()
        rc, got = O.recover({i: out1.data[i * w:(i + 1) * w].tobytes() for i in ids}, geo.log_col, geo.log_row)
        recovery["coset2_1024_rows"]["cpu_oracle_ms"] = round((time.time() - t0) * 1e3, 1)
        recovery["coset2_1024_rows"]["cpu_oracle_threads"] = args.cpu_threads
        assert rc == 0 and got == raws[0]
        # the oracle's deferred-verification MSM over coset 0's 1024 rows (a third)
        bases = [ctx.params_download(c) for c in range(geo.coset_n)]
        w1 = 32 * geo.col_n
        t0 = time.time()
        O.verify_msm(bases, geo.log_col, geo.log_row, v_ids[:geo.row_n], v_rows[:geo.row_n * w1],
                     v_alphas[:geo.row_n])
        verify["cpu_oracle_ms_1024_rows"] = round((time.time() - t0) * 1e3, 1)
        verify["cpu_oracle_threads"] = args.cpu_threads
    ctx.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
