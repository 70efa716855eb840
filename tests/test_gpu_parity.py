"""HIP path vs the CPU oracle at the reference's test geometry (2^6 rows x 2^5
columns, 3 cosets; crates/encoder/src/constants.rs:11-15).  Bit-exact."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, bamt_image, oracle_params, random_fr_mont, ri_from_y, seeded_tau

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING


PLANS = {"fixed_base": None, "fixed_base_narrow": None, "fixed_base_wide_segments": None,
         "fixed_base_wide_fine": "wide-fine", "fixed_base_narrow_plan": "narrow", "fixed_base_per_chunk": "none",
         "fixed_base_mixed_plan": "1,0,2,1,0,2,1,0,2,1,0,2", "windowed": None}


@pytest.fixture(scope="module", params=list(PLANS))
def ctx(request):
    """Every MSM path over the resident bases: the shifted-copy tables (default: the
    AMT lines share bucket sets over chunk groups; quotient lines 0 and 1 and the high
    line, whose groups hold at least 16 chunks, in 16-bit windows), the same with 12-bit
    windows on every line (ZGR_WIDE=0), with the wide lines split into 7 bucket-segment
    waves each (ZGR_WIDE_SEGS=7; the default at this size is two), the tuner's candidates
    "wide-fine" (twice the default segments) and "narrow" (12-bit windows everywhere, the
    reduction two groups per wave) alone (ZGR_GROUP_PLAN=<name>), with one bucket set
    per chunk on every line (ZGR_GROUP_PLAN=none) and with an explicit per-line plan that
    mixes the group levels (each capped at what the line's outputs allow), and the
    windowed kernels (ZGR_FIXED_BASE=0); the variables are read at context creation."""
    fixed = request.param != "windowed"
    c = Z.make_context(GEO, fixed_base=fixed, group_plan=PLANS[request.param], wide=request.param != "fixed_base_narrow",
                       wide_segs=7 if request.param == "fixed_base_wide_segments" else None)
    c.set_profiling(True)
    c.params_generate(seeded_tau())
    # the table build runs only on the fixed-base path
    assert (c.kernel_time_ms("fb_build")[1] > 0) == fixed
    c.set_profiling(False)
    c.kernel_time_reset()
    c.mode = request.param
    yield c
    c.close()


def test_msm_path_in_use(ctx):
    # chunk-group entry lists are written only when some AMT line shares bucket sets
    ctx.set_profiling(True)
    ctx.kernel_time_reset()
    ctx.encode(Z.seeded_bytes(DATA_SEED, 1000))
    merges = ctx.kernel_time_ms("msm_group_scatter")[1]
    ctx.set_profiling(False)
    ctx.kernel_time_reset()
    assert (merges > 0) == (ctx.mode in ("fixed_base", "fixed_base_narrow", "fixed_base_wide_segments",
                                         "fixed_base_wide_fine", "fixed_base_narrow_plan", "fixed_base_mixed_plan"))


def test_wide_lines_in_use(ctx):
    # the 16-bit-window sort runs exactly when the default plan has wide lines
    ctx.set_profiling(True)
    ctx.kernel_time_reset()
    ctx.encode(Z.seeded_bytes(DATA_SEED, 1000))
    wide = ctx.kernel_time_ms("msm_wide_sort")[1]
    ctx.set_profiling(False)
    ctx.kernel_time_reset()
    assert (wide > 0) == (ctx.mode in ("fixed_base", "fixed_base_wide_segments", "fixed_base_wide_fine"))


def test_group_plan_rejected():
    """A malformed ZGR_GROUP_PLAN fails context creation with the variable in the message,
    and so does a wide candidate where no line is wide."""
    with pytest.raises(Z.ZgrError, match="ZGR_GROUP_PLAN"):
        Z.make_context(GEO, group_plan="2,x,1")
    with pytest.raises(Z.ZgrError, match="ZGR_GROUP_PLAN=wide-fine"):
        Z.make_context(GEO, group_plan="wide-fine", wide=False)


@pytest.fixture(scope="module")
def params():
    return oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n)


def test_params_generate_matches_oracle(ctx, params):
    # generate.rs:139-232 (closed form in both; the oracle's closed form is pinned
    # against its EC-FFT restatement in test_oracle_params.py)
    for c in range(GEO.coset_n):
        got = ctx.params_download(c)
        assert np.array_equal(got, params[c]), "coset %d" % c


def test_params_load_bamt(ctx, params):
    # read_amt_params layout (fast_serde_bn254.rs:58-88): the prove lines land in the
    # device order quotients || basis || high, G2 sections skipped
    for c in range(GEO.coset_n):
        img = bamt_image(params[c], GEO.log_n, GEO.h)
        ctx.params_load_bamt(c, img)
        assert np.array_equal(ctx.params_download(c), params[c]), "coset %d" % c
    img = bamt_image(params[0], GEO.log_n, GEO.h)
    bad = [b"bamx" + img[4:],                                    # header
           img[:4] + bytes([GEO.log_n + 1]) + img[5:],           # degree
           img[:5] + bytes([GEO.h - 1]) + img[6:],               # sub-degree
           img[:-64],                                            # truncated
           img + bytes(1)]                                       # trailing byte
    for b in bad:
        with pytest.raises(Z.ZgrError) as e:
            ctx.params_load_bamt(0, b)
        assert e.value.code == Z.ERR_INVALID_ARG
    assert np.array_equal(ctx.params_download(0), params[0])  # still resident


def test_coset_extend_matches_oracle(ctx):
    rng = np.random.default_rng(11)
    data = random_fr_mont(rng, GEO.n)
    for c in range(GEO.coset_n):
        exp = O.to_coset_blob(data, GEO.log_n, c)
        got = ctx.coset_extend(data, c)
        assert np.array_equal(got.reshape(-1), exp.reshape(-1)), "coset %d" % c


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("log_n", [1, 4, 10, 11])
def test_ntt_matches_oracle(ctx, log_n, inverse):
    rng = np.random.default_rng(100 + log_n)
    rows = np.stack([random_fr_mont(rng, 1 << log_n) for _ in range(2)])
    got = ctx.ntt(rows, log_n, inverse)
    for r in range(2):
        exp = (O.ifft if inverse else O.fft)(rows[r], log_n)
        assert np.array_equal(got[r].reshape(-1), exp.reshape(-1))


@pytest.mark.parametrize("kind", ["random", "zeros", "small", "sparse", "equal", "negative"])
def test_multiple_multiexp_matches_oracle(ctx, params, kind):
    rng = np.random.default_rng(5)
    if kind == "random":
        ri = random_fr_mont(rng, GEO.n)
    elif kind == "zeros":
        ri = np.zeros((GEO.n, 4), dtype=np.uint64)
    elif kind == "small":
        ri = O.fr_to_mont_array([int(x) for x in rng.integers(0, 300, GEO.n)])
    elif kind == "equal":
        # one scalar everywhere: every chunk's digits land in at most 22 buckets, each
        # far larger than a lane's even share of the entries
        ri = O.fr_to_mont_array([int.from_bytes(rng.bytes(32), "little") % O.R_MOD] * GEO.n)
    elif kind == "negative":
        # -k for small k: top digits at their extremes with carries through every window
        ri = O.fr_to_mont_array([O.R_MOD - int(x) for x in rng.integers(1, 5000, GEO.n)])
    else:
        vals = [0] * GEO.n
        for i in rng.choice(GEO.n, 40, replace=False):
            vals[int(i)] = int.from_bytes(rng.bytes(32), "little") % O.R_MOD
        vals[3] = O.R_MOD - 1  # -1: every digit at its extreme
        ri = O.fr_to_mont_array(vals)
    for c in range(GEO.coset_n):
        exp = O.multiple_multiexp(params[c], GEO.log_n, GEO.h, ri)
        got = ctx.multiple_multiexp(c, ri)
        assert np.array_equal(got, exp), "coset %d" % c


def test_amt_prove_matches_oracle(ctx, params):
    rng = np.random.default_rng(9)
    ri = random_fr_mont(rng, GEO.n)
    for c in range(GEO.coset_n):
        exp = O.amt_prove(params[c], GEO.log_n, GEO.h, ri)
        got = ctx.amt_prove(c, ri)
        for a, b, name in zip(got, exp, ("commitment", "comm_tree", "proof_tree", "high")):
            assert np.array_equal(np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)), (c, name)


# encoder/blob.rs:282-286 test sizes: 0, 1, 1234, MAX_RAW_DATA_SIZE
@pytest.mark.parametrize("length", [0, 1, 1234, GEO.max_raw_data_size])
def test_encode_matches_oracle(ctx, params, length):
    raw = Z.seeded_bytes(DATA_SEED, length)
    exp = O.encode(list(params), GEO.log_col, GEO.log_row, raw)
    got = ctx.encode(raw)
    assert np.array_equal(got.commitment, exp.commitment)
    assert np.array_equal(got.file_root, exp.root)
    assert np.array_equal(got.high, exp.high)
    assert np.array_equal(got.comm_tree, exp.comm_tree)
    assert np.array_equal(got.proof_tree, exp.proof_tree)
    assert np.array_equal(got.merkle, exp.merkle)
    assert np.array_equal(got.data, exp.data)


def test_encode_too_large(ctx):
    # encoder/blob.rs:286: MAX_RAW_DATA_SIZE + 1 -> TooLargeBlob
    with pytest.raises(Z.TooLargeBlob):
        ctx.encode(b"\x00" * (GEO.max_raw_data_size + 1))
    # the C ABI itself rejects the size too (no Python pre-check)
    out = Z.EncodedBlobOutputs(GEO)
    buf = np.zeros(GEO.max_raw_data_size + 1, dtype=np.uint8)
    cs = out.cstruct()
    import ctypes
    rc = Z.lib().zgr_encode(ctx.h, ctypes.c_void_p(buf.ctypes.data), buf.size, ctypes.byref(cs))
    assert rc == Z.ERR_TOO_LARGE
    assert "TooLargeBlob" in Z.lib().zgr_last_error().decode()


def test_merkle_primitive_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, GEO.coset_n * GEO.n * 32, dtype=np.uint8)
    nl, exp_layers, exp_root = O.merkle(data, GEO.coset_n, GEO.log_row)
    layers, root = ctx.merkle(data, GEO.log_row)
    assert root == exp_root
    assert np.array_equal(layers.reshape(-1), exp_layers.reshape(-1))


def _random_points(rng, n, with_inf=True):
    pts = np.zeros((n, 64), dtype=np.uint8)
    for i in range(n):
        pts[i] = O.g1_mul_gen(int.from_bytes(rng.bytes(32), "little"))
    if with_inf:
        pts[n // 3] = 0  # the point at infinity
    return pts


def _canon_to_mont_bases(pts):
    out = np.zeros((len(pts), 8), dtype=np.uint64)
    for i, p in enumerate(pts):
        xy = O.point_to_ints(p)
        out[i, :4] = O.int_to_limbs(O.to_mont(1, xy[0]))
        out[i, 4:] = O.int_to_limbs(O.to_mont(1, xy[1]))
    return out


@pytest.mark.parametrize("log_n", [1, 3, 6])
@pytest.mark.parametrize("inverse", [False, True])
def test_ecfft_matches_oracle(ctx, log_n, inverse):
    # Pairing::fast_fft / fast_ifft (crates/amt/src/ec_algebra.rs:34-95)
    rng = np.random.default_rng(40 + log_n)
    pts = _random_points(rng, 1 << log_n)
    exp = O.ecfft(pts, log_n, inverse)
    got = ctx.ecfft(pts, log_n, inverse)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("n", [1, 5, 1000, 2048])
def test_msm_matches_oracle(ctx, n):
    rng = np.random.default_rng(n)
    pts = _random_points(rng, n, with_inf=False)
    bases = _canon_to_mont_bases(pts)
    scal = random_fr_mont(rng, n)
    if n > 4:
        scal[1] = 0
        scal[2] = scal[3]
    exp = O.msm(bases, np.array([O.int_to_limbs(v) for v in O.fr_from_mont_array(scal)]))
    got = ctx.msm(bases, scal)
    assert np.array_equal(got, exp)


def test_encode_batch_matches_oracle(ctx, params):
    # the batched entry point pipelines blobs over two slots and two coset lanes;
    # every blob must come out exactly as the oracle encodes it alone
    lengths = [0, 1, 1234, GEO.max_raw_data_size, 777, 31 * 64]
    raws = [Z.seeded_bytes(DATA_SEED + 1 + i, L) for i, L in enumerate(lengths)]
    dev = [Z.DeviceBuffer(r) for r in raws]
    outs = [Z.EncodedBlobOutputs(GEO) for _ in raws]
    ctx.encode_batch_device([d.address for d in dev], [len(r) for r in raws], outs)
    for r, got in zip(raws, outs):
        exp = O.encode(list(params), GEO.log_col, GEO.log_row, r)
        assert np.array_equal(got.commitment, exp.commitment)
        assert np.array_equal(got.file_root, exp.root)
        assert np.array_equal(got.proof_tree, exp.proof_tree)
        assert np.array_equal(got.comm_tree, exp.comm_tree)
        assert np.array_equal(got.high, exp.high)
        assert np.array_equal(got.merkle, exp.merkle)
        assert np.array_equal(got.data, exp.data)


def test_autotune_keeps_outputs(params):
    """zgr_autotune picks a chunk-group plan per density class of a coset's scalars; the
    encode outputs stay bit-exact with the oracle for blobs of every class."""
    c = Z.make_context(GEO)
    try:
        c.params_generate(seeded_tau())
        rep = c.autotune()
        classes = rep["msm_plan"]
        assert len(classes) == 4
        for v in classes.values():
            assert v["chosen"] in ("default", "narrow", "wide-coarse", "wide-fine", "max", "half", "chunk") and v["chosen"] in v
        for length in (0, 100, GEO.max_raw_data_size // 9, GEO.max_raw_data_size // 3, GEO.max_raw_data_size):
            raw = Z.seeded_bytes(DATA_SEED + 17, length)
            got = c.encode(raw)
            exp = O.encode(list(params), GEO.log_col, GEO.log_row, raw)
            assert np.array_equal(got.commitment, exp.commitment), length
            assert np.array_equal(got.proof_tree, exp.proof_tree), length
            assert np.array_equal(got.comm_tree, exp.comm_tree), length
            assert np.array_equal(got.file_root, exp.root), length
    finally:
        c.close()


def test_graph_replay_equals_launches(params):
    """The MSM stage replayed from captured graphs (the default) and launched kernel by
    kernel (ZGR_GRAPHS=0) give the oracle's outputs, over repeated encodes of blobs of
    different density (several captured graphs, each replayed more than once)."""
    import os

    ctxs = []
    try:
        for mode in ("1", "0"):
            old = os.environ.get("ZGR_GRAPHS")
            os.environ["ZGR_GRAPHS"] = mode
            try:
                c = Z.Context(GEO)
            finally:
                if old is None:
                    os.environ.pop("ZGR_GRAPHS", None)
                else:
                    os.environ["ZGR_GRAPHS"] = old
            c.params_generate(seeded_tau())
            ctxs.append(c)
        for rep in range(3):
            for length in (0, 5000, GEO.max_raw_data_size):
                raw = Z.seeded_bytes(DATA_SEED + 40 + rep, length)
                a, b = ctxs[0].encode(raw), ctxs[1].encode(raw)
                exp = O.encode(list(params), GEO.log_col, GEO.log_row, raw)
                for got in (a, b):
                    assert np.array_equal(got.commitment, exp.commitment), (rep, length)
                    assert np.array_equal(got.proof_tree, exp.proof_tree), (rep, length)
                    assert np.array_equal(got.comm_tree, exp.comm_tree), (rep, length)
                    assert np.array_equal(got.data, exp.data), (rep, length)
    finally:
        for c in ctxs:
            c.close()
