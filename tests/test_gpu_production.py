"""HIP path at the production geometry (1024 x 1024, 3 cosets, AMT depth 10).

Full-size parity is checked three ways: the generated params against the oracle's
closed form at sampled positions; the encoded blob against size-independent
identities (commitment = [f(tau)] G1, high = [tau^(3N) f(tau)] G1, coset 0 rows
equal the raw data); and the whole EncodedBlob bit-exact against the oracle's
encode run over the GPU's own params."""
import sys
import time

import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import CPU_THREADS, DATA_SEED, seeded_tau

pytestmark = pytest.mark.gpu

GEO = Z.PRODUCTION


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    t = time.time()
    c.params_generate(seeded_tau())
    log("production params generated in %.2f s" % (time.time() - t))
    yield c
    c.close()


@pytest.fixture(scope="module")
def blob(ctx):
    raw = Z.seeded_bytes(DATA_SEED, GEO.max_raw_data_size)
    ctx.encode(raw)  # warm
    t = time.time()
    out = ctx.encode(raw)
    log("production encode (host in/out) %.3f s" % (time.time() - t))
    return raw, out


def test_params_sampled_points(ctx):
    rng = np.random.default_rng(17)
    tau = seeded_tau()
    for c in range(GEO.coset_n):
        got = ctx.params_download(c).reshape(GEO.h + 2, GEO.n, 8)
        for _ in range(6):
            line = int(rng.integers(0, GEO.h + 2))
            pos = int(rng.integers(0, GEO.n))
            exp = O.params_point(tau, GEO.log_n, GEO.h, c, line, pos)
            assert np.array_equal(got[line, pos], exp), (c, line, pos)


def test_params_from_pp_production(ctx):
    """zgr_params_from_pp over 2^20 powers of tau gives, for coset 2, exactly the
    buffer the closed-form generator left (and the shifted-copy tables are rebuilt)."""
    O.set_threads(CPU_THREADS)
    tau = seeded_tau()
    t = time.time()
    g1pp = O.power_tau(tau, 0, GEO.log_n)
    high = O.power_tau(tau, 3 * GEO.n, GEO.log_n)
    log("powers of tau (oracle, %d threads): %.1f s" % (CPU_THREADS, time.time() - t))
    before = ctx.params_download(2)
    t = time.time()
    ctx.params_from_pp(2, g1pp, high)
    log("params_from_pp, one coset at 2^20: %.2f s" % (time.time() - t))
    assert np.array_equal(ctx.params_download(2), before)


def test_build_params_production(ctx):
    """bin/build_params at 1024 x 1024 on the device, without tau after the powers: the
    powers of tau in G1 and G2, then for every coset the prove bases and the vanishes from
    them (AMTParams::from_pp), each equal to the closed-form generator's; the bamt image
    of coset 0 loads back into the same bases."""
    tau = seeded_tau()
    ctx.params_generate(tau)  # the closed form, verifier half included (a G1-only from_pp clears it)
    closed = [ctx.params_g2(c) for c in range(GEO.coset_n)]
    before = [ctx.params_download(c) for c in range(GEO.coset_n)]
    t = time.time()
    ctx.power_tau_setup(tau)
    log("power tau (2^20 G1, 2^20 high G1, 2^20 G2) on device: %.2f s" % (time.time() - t))
    for c in range(GEO.coset_n):
        t = time.time()
        ctx.params_from_power_tau(c)
        log("params from powers of tau, coset %d (12 EC-FFTs of 2^20 + vanishes): %.2f s" % (c, time.time() - t))
        assert np.array_equal(ctx.params_download(c), before[c]), c
        for a, b in zip(ctx.params_g2(c), closed[c]):
            assert np.array_equal(a, b), c
    for kind, name in ((Z.FILE_AMT_MONT, "bamt"), (Z.FILE_AMT, "ark AMTParams"), (Z.FILE_AMT_VERIFY, "verify"),
                       (Z.FILE_PTAU, "power tau")):
        t = time.time()
        img = ctx.write_params_file(kind, 0)
        log("%s image of coset 0: %.1f MB in %.2f s" % (name, len(img) / 1e6, time.time() - t))
        if kind == Z.FILE_AMT_MONT:
            ctx.params_load_bamt(0, img)
            assert np.array_equal(ctx.params_download(0), before[0])
            assert ctx.write_params_file(Z.FILE_AMT_MONT, 0) == img
        del img


def test_encode_closed_form(blob):
    raw, out = blob
    O.set_threads(CPU_THREADS)
    tau = seeded_tau()
    rowmaj = O.raw_to_blob(raw, GEO.log_n).reshape(GEO.row_n, GEO.col_n, 4)
    colmaj = np.ascontiguousarray(rowmaj.transpose(1, 0, 2)).reshape(-1, 4)
    ftau = O.eval_interp(colmaj, GEO.log_n, tau)
    assert O.point_to_ints(out.commitment) == O.point_to_ints(O.g1_mul_gen(ftau))
    high = O.g1_mul_gen(ftau * pow(tau, 3 * GEO.n, O.R_MOD))
    for c in range(GEO.coset_n):
        assert np.array_equal(out.high[64 * c:64 * c + 64], high), c
    # coset 0 is the raw blob itself (encode.rs:138-140), canonical
    canon = np.array(O.fr_from_mont_array(rowmaj.reshape(-1, 4)[:GEO.col_n]), dtype=object)
    first_row = [int.from_bytes(bytes(out.data[32 * i:32 * i + 32]), "little") for i in range(GEO.col_n)]
    assert list(canon) == first_row


def test_encode_bit_exact_vs_oracle(ctx, blob):
    raw, out = blob
    O.set_threads(CPU_THREADS)
    bases = [ctx.params_download(c) for c in range(GEO.coset_n)]
    t = time.time()
    exp = O.encode(bases, GEO.log_col, GEO.log_row, raw)
    log("oracle production encode %.1f s on %d threads" % (time.time() - t, CPU_THREADS))
    assert np.array_equal(out.commitment, exp.commitment)
    assert np.array_equal(out.file_root, exp.root)
    assert np.array_equal(out.high, exp.high)
    assert np.array_equal(out.comm_tree, exp.comm_tree)
    assert np.array_equal(out.proof_tree, exp.proof_tree)
    assert np.array_equal(out.merkle, exp.merkle)
    assert np.array_equal(out.data, exp.data)


def test_ecfft_2pow10(ctx):
    # the isolated 2^10-point G1 EC-FFT over seeded multiples of G1, both directions
    rng = np.random.default_rng(1010)
    pts = np.zeros((1 << 10, 64), dtype=np.uint8)
    for i in range(1 << 10):
        pts[i] = O.g1_mul_gen(int.from_bytes(rng.bytes(32), "little"))
    for inverse in (False, True):
        t = time.time()
        got = ctx.ecfft(pts, 10, inverse)
        log("ecfft 2^10 inverse=%d: %.3f s" % (inverse, time.time() - t))
        assert np.array_equal(got, O.ecfft(pts, 10, inverse))


def test_msm_2pow20(ctx):
    # the isolated MSM: 2^20 resident-style bases (the basis line of coset 0)
    # with uniform random scalars, bit-exact against the oracle's Pippenger
    rng = np.random.default_rng(2020)
    bases = ctx.params_download(0).reshape(GEO.h + 2, GEO.n, 8)[GEO.h].copy()
    scal = rng.integers(0, 2**63, (GEO.n, 4), dtype=np.uint64)
    scal[:, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)  # < 2^252 < r: a valid Fr value
    t = time.time()
    got = ctx.msm(bases, scal)
    log("msm 2^20: %.3f s" % (time.time() - t))
    canon = [O.from_mont(0, O.limbs_to_int(scal[i])) for i in range(GEO.n)]
    exp = O.msm(bases, O.ints_to_limbs(canon))
    assert np.array_equal(got, exp)


def _fr_rows(rng, rows, log_n):
    """uniform values below 2^252 < r as Montgomery limbs (every such value is a valid
    Montgomery representative)"""
    a = rng.integers(0, 2**63, (rows, 1 << log_n, 4), dtype=np.uint64)
    a[:, :, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
    return a


@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_batch_1024_rows_2pow10(ctx, inverse):
    """the isolated batched Fr NTT of 1024 rows x 2^10 (one launch, each row in LDS),
    every row bit-exact against the oracle's radix-2 fft / ifft, through both the
    host-memory and the device-memory entry points."""
    rng = np.random.default_rng(3030 + inverse)
    rows = _fr_rows(rng, 1024, 10)
    t = time.time()
    got = ctx.ntt(rows, 10, inverse).reshape(1024, 1 << 10, 4)
    log("ntt 1024 x 2^10 inverse=%d (host in/out): %.4f s" % (inverse, time.time() - t))
    f = O.ifft if inverse else O.fft
    for r in range(1024):
        assert np.array_equal(got[r], f(rows[r], 10)), r
    dev = Z.DeviceBuffer(rows.tobytes())
    ctx.ntt_device(dev.address, 10, 1024, inverse)
    assert np.array_equal(np.frombuffer(dev.to_host(), dtype=np.uint64).reshape(got.shape), got)


@pytest.mark.parametrize("log_n,batch", [(12, 3), (13, 2), (17, 1), (20, 1), (20, 2)])
def test_ntt_two_pass(ctx, log_n, batch):
    """Transforms above 2^11 points run as two passes (four-step, n = ceil(n/2) +
    floor(n/2)) and come back to natural order; both directions bit-exact against the
    oracle, at even and odd splits and at the full 2^20 domain."""
    rng = np.random.default_rng(4000 + log_n)
    rows = _fr_rows(rng, batch, log_n)
    for inverse in (False, True):
        got = ctx.ntt(rows, log_n, inverse).reshape(rows.shape)
        f = O.ifft if inverse else O.fft
        for r in range(batch):
            assert np.array_equal(got[r], f(rows[r], log_n)), (inverse, r)
    dev = Z.DeviceBuffer(rows.tobytes())
    ctx.ntt_device(dev.address, log_n, batch, False)
    assert np.array_equal(np.frombuffer(dev.to_host(), dtype=np.uint64).reshape(rows.shape),
                          ctx.ntt(rows, log_n, False).reshape(rows.shape))


def test_coset_extend_production(ctx):
    """to_coset_blob at 2^20 (four-step iNTT, coset scaling fused into the first DIT
    pass): every coset bit-exact against the oracle."""
    rng = np.random.default_rng(5050)
    data = _fr_rows(rng, 1, GEO.log_n)[0]
    for c in range(GEO.coset_n):
        assert np.array_equal(ctx.coset_extend(data, c).reshape(-1, 4), O.to_coset_blob(data, GEO.log_n, c)), c


def test_batch_equals_single(ctx, blob):
    # the uneven-batch shape: uneven full-geometry blobs in one batched call, each bit-equal
    # to the same blob encoded alone
    raw0, out0 = blob
    rng = np.random.default_rng(44)
    lengths = [GEO.max_raw_data_size, 1024, int(rng.integers(1, GEO.max_raw_data_size)), 0]
    raws = [raw0] + [Z.seeded_bytes(300 + i, L) for i, L in enumerate(lengths[1:])]
    dev = [Z.DeviceBuffer(r) for r in raws]
    outs = [Z.EncodedBlobOutputs(GEO) for _ in raws]
    t = time.time()
    ctx.encode_batch_device([d.address for d in dev], [len(r) for r in raws], outs)
    log("batch of %d blobs: %.3f s" % (len(raws), time.time() - t))
    singles = [out0] + [ctx.encode(r) for r in raws[1:]]
    for a, b in zip(outs, singles):
        for f in ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_uneven_mix_32(ctx):
    """the uneven batch: the seeded 32-blob mix (raw lengths log-uniform in [1 KB, 32.5 MB], the
    same seeds bench.py times) in one batched call.  Every blob: commitment = [f(tau)] G1
    and high = [tau^(3N) f(tau)] G1 for the interpolant f of its column-major data,
    coset 0 equal to its raw blob, and every output equal to the same blob encoded alone."""
    tau = seeded_tau()
    rng = np.random.default_rng(DATA_SEED)
    lens = np.exp(rng.uniform(np.log(1024), np.log(GEO.max_raw_data_size), 32)).astype(np.int64)
    raws = [Z.seeded_bytes(DATA_SEED + 100 + i, int(L)) for i, L in enumerate(lens)]
    dev = [Z.DeviceBuffer(r) for r in raws]
    outs = [Z.EncodedBlobOutputs(GEO) for _ in raws]
    t = time.time()
    ctx.encode_batch_device([d.address for d in dev], [len(r) for r in raws], outs)
    log("32-blob uneven mix (%.1f MB raw): %.3f s" % (lens.sum() / 1e6, time.time() - t))
    del dev
    O.set_threads(CPU_THREADS)
    t3n = pow(tau, 3 * GEO.n, O.R_MOD)
    for k, (raw, out) in enumerate(zip(raws, outs)):
        rowmaj = O.raw_to_blob(raw, GEO.log_n).reshape(GEO.row_n, GEO.col_n, 4)
        colmaj = np.ascontiguousarray(rowmaj.transpose(1, 0, 2)).reshape(-1, 4)
        ftau = O.eval_interp(colmaj, GEO.log_n, tau)
        assert O.point_to_ints(out.commitment) == O.point_to_ints(O.g1_mul_gen(ftau)), k
        high = O.g1_mul_gen(ftau * t3n % O.R_MOD)
        assert all(np.array_equal(out.high[64 * c:64 * c + 64], high) for c in range(GEO.coset_n)), k
        canon = O.fr_canonical_array(rowmaj.reshape(-1, 4)).view(np.uint8).reshape(-1)
        assert np.array_equal(out.data[:GEO.n * 32], canon), k
        single = ctx.encode(raw)
        for f in ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data"):
            assert np.array_equal(getattr(out, f), getattr(single, f)), (k, f)


def test_autotune_production(ctx, blob):
    """The tuner at 1024 x 1024: a plan per density class and the NTT split; the full
    blob encodes to the same outputs afterwards."""
    raw, out = blob
    t = time.time()
    rep = ctx.autotune()
    log("autotune %.2f s: %s" % (time.time() - t, rep))
    assert len(rep["msm_plan"]) == 4 and rep["ntt_split"]["chosen"] in ("10+10", "11+9")
    got = ctx.encode(raw)
    for f in ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data"):
        assert np.array_equal(getattr(got, f), getattr(out, f)), f


def test_served_equals_single(ctx, blob):
    # the served shape: requests of mixed sizes submitted together to the serving
    # layer; each result equals the same blob encoded alone
    raw0, out0 = blob
    lengths = [GEO.max_raw_data_size, 5000, 3_000_000, 1, 20_000_000]
    raws = [raw0] + [Z.seeded_bytes(400 + i, L) for i, L in enumerate(lengths[1:])]
    srv = Z.Server(ctx, hbm_budget=1 << 30, max_inflight=64)
    try:
        t = time.time()
        tickets = [srv.submit(r) for r in raws]
        outs = [srv.poll(tk) for tk in tickets]
        log("served %d mixed blobs: %.3f s" % (len(raws), time.time() - t))
    finally:
        srv.close()
    singles = [out0] + [ctx.encode(r) for r in raws[1:]]
    for a, b in zip(outs, singles):
        for f in ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_windowed_path_equals_fixed_base(ctx, blob):
    # the windowed MSM kernels (ZGR_FIXED_BASE=0) over the same params give the same
    # multiple_multiexp points and the same encoded blob as the shifted-copy tables
    raw0, out0 = blob
    w = Z.make_context(GEO, fixed_base=False)
    try:
        w.params_generate(seeded_tau())
        rng = np.random.default_rng(77)
        ri = rng.integers(0, 2**63, (GEO.n, 4), dtype=np.uint64)
        ri[:, 3] &= np.uint64(0x0FFFFFFFFFFFFFFF)
        for c in (0, 2):
            assert np.array_equal(w.multiple_multiexp(c, ri), ctx.multiple_multiexp(c, ri)), "coset %d" % c
        got = w.encode(raw0)
        for f in ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data"):
            assert np.array_equal(getattr(got, f), getattr(out0, f)), f
    finally:
        w.close()


def test_recover_production(ctx, blob):
    """recover_from_da_slice at 1024 x 1024: the raw data back from a third of the
    rows (no coset-0 row among them), from a random 1100, and from all 3072; a
    corrupted extra row is caught."""
    raw, out = blob
    w = 32 * GEO.col_n
    rows = {i: out.data[i * w:(i + 1) * w].tobytes() for i in range(GEO.row_encoded)}
    rng = np.random.default_rng(41)
    sets = [list(range(GEO.row_encoded - GEO.row_n, GEO.row_encoded)),
            sorted(int(i) for i in rng.choice(GEO.row_encoded, 1100, replace=False)),
            list(range(GEO.row_encoded))]
    for ids in sets:
        t = time.time()
        got = Z.recover_from_da_slice(ctx, {i: rows[i] for i in ids})
        log("production recovery from %d rows: %.3f s" % (len(ids), time.time() - t))
        assert got == raw
    s = {i: rows[i] for i in sets[1]}
    k = sets[1][-1]
    b = bytearray(s[k])
    b[77] ^= 4
    s[k] = bytes(b)
    with pytest.raises(Z.RecoveryPanic):
        Z.recover_from_da_slice(ctx, s)


def test_deferred_verify_production(ctx, blob):
    """The deferred verifier's MSM over all 3072 rows of the production blob (one
    2^20-point MSM per coset over the resident basis) accepts, and rejects one changed
    element."""
    from zg_rocm import wire

    raw, out = blob
    eb = Z.EncodedBlob(GEO, out)
    rows = eb.encoded_slices()
    v = Z.DeferredVerifier(ctx, bytes(out.commitment))
    for i in range(GEO.row_encoded):
        v.record_slice(rows[i].tobytes())
    t = time.time()
    assert v.fast_check_msm()
    log("deferred MSM check over 3072 rows: %.3f s" % (time.time() - t))
    f = wire.parse_encoded_slice(rows[1500].tobytes())
    r = bytearray(b"".join(f["row"]))
    r[32 * 17] ^= 1
    f["row"] = bytes(r)
    v = Z.DeferredVerifier(ctx, bytes(out.commitment))
    v.record_slice(wire.serialize_encoded_slice(f))
    v.record_slice(rows[0].tobytes())
    assert not v.fast_check_msm()
