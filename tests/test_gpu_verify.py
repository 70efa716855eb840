"""The GPU half of the deferred AMT verification (zgr_verify_msm, DeferredVerifier::
fast_check_msm_gpu, crates/amt/src/deferred_verification.rs:203-222) at the test
geometry: the device MSM equals the oracle's restatement, and the Python
DeferredVerifier accepts every honest slice of a device-encoded blob and rejects
tampered ones (verify_amt_proof's MSM check, crates/amt/src/verify_params.rs:176-186)."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, oracle_params, seeded_tau
from zg_rocm import wire

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    c.params_generate(seeded_tau())
    yield c
    c.close()


@pytest.fixture(scope="module")
def slices(ctx):
    out = ctx.encode(Z.seeded_bytes(DATA_SEED + 500, 40000))
    blob = Z.EncodedBlob(GEO, out)
    return [blob.encoded_slice(i) for i in range(GEO.row_encoded)]


def _auth(slices):
    """The authoritative commitment (canonical x || y) from a slice's wire form."""
    b = bytearray(wire.parse_encoded_slice(slices[0])["commitment"])
    b[63] &= 0x3F
    return bytes(b)


def test_verify_msm_matches_oracle(ctx, slices):
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    rng = np.random.default_rng(8)
    rows = [b"".join(wire.parse_encoded_slice(s)["row"]) for s in slices]
    for ids in [[0], [GEO.row_encoded - 1], list(range(GEO.row_n, 2 * GEO.row_n)), list(range(GEO.row_encoded)),
                sorted(int(i) for i in rng.choice(GEO.row_encoded, 57, replace=False))]:
        alphas = [int.from_bytes(rng.bytes(32), "little") % O.R_MOD for _ in ids]
        got = ctx.verify_msm(ids, [rows[i] for i in ids], alphas)
        want = O.verify_msm(params, GEO.log_col, GEO.log_row, ids, b"".join(rows[i] for i in ids), alphas)
        assert got == want, ids[:3]
    assert ctx.verify_msm([], b"", []) == bytes(64)


def test_deferred_verifier_accepts_honest_rows(ctx, slices):
    v = Z.DeferredVerifier(ctx, _auth(slices))
    for s in slices:
        v.record_slice(s)
    assert v.fast_check_msm()
    v = Z.DeferredVerifier(ctx, _auth(slices))
    for i in range(5, GEO.row_encoded, 11):
        v.record_slice(slices[i])
    assert v.fast_check_msm()


def _tamper_row(s, pos):
    f = wire.parse_encoded_slice(s)
    row = bytearray(b"".join(f["row"]))
    row[pos] ^= 1
    f["row"] = bytes(row)
    return wire.serialize_encoded_slice(f)


def test_deferred_verifier_rejects_tampering(ctx, slices):
    v = Z.DeferredVerifier(ctx, _auth(slices))
    for i, s in enumerate(slices):
        v.record_slice(_tamper_row(s, 70) if i == 77 else s)
    assert not v.fast_check_msm()
    # a proof commitment taken from another row
    f = wire.parse_encoded_slice(slices[20])
    g = wire.parse_encoded_slice(slices[21])
    f["amt_proof"] = [g["amt_proof"][0]] + list(f["amt_proof"][1:])
    v = Z.DeferredVerifier(ctx, _auth(slices))
    v.record_slice(wire.serialize_encoded_slice(f))
    v.record_slice(slices[30])
    assert not v.fast_check_msm()


def test_verify_msm_argument_errors(ctx, slices):
    row = b"".join(wire.parse_encoded_slice(slices[0])["row"])
    with pytest.raises(Z.ZgrError):
        ctx.verify_msm([GEO.row_encoded], row, [1])
    with pytest.raises(Z.ZgrError):
        ctx.verify_msm([0], row, [O.R_MOD])  # alpha not canonical
    with pytest.raises(Z.ZgrError):
        ctx.verify_msm([2, 1], row + row, [1, 1])


def test_device_slices_pass_reference_verification(ctx, slices):
    """Slices of the device-encoded blob pass the restated EncodedSlice::verify with
    its pairing checks (crates/encoder/src/encoder/slice.rs:72-95), one by one and as a
    deferred batch whose MSM half runs on the device."""
    import random

    import verify as V

    tau = seeded_tau()
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    vps = []
    for c in range(GEO.coset_n):
        canon = O.fr_canonical_array(params[c].reshape(-1, 4), which=1).reshape(GEO.h + 2, GEO.n, 2, 4)
        vps.append(V.VerifyParams(tau, GEO.log_col, GEO.log_row, c,
                                  [V.g1_from_canon(canon[GEO.h][i].tobytes()) for i in range(GEO.n)]))
    first = wire.parse_encoded_slice(slices[0])
    com = V.g1_from_wire(first["commitment"])
    root = V.file_root([bytes(x) for x in first["merkle_root"]])
    for i in (3, GEO.row_n + 40, GEO.row_encoded - 2):
        assert V.verify_slice(wire.parse_encoded_slice(slices[i]), vps, com, root, GEO) is None, i
    col = V.Collector(True)
    dv = Z.DeferredVerifier(ctx, _auth(slices))
    for i in (0, 17, GEO.row_n + 1, 2 * GEO.row_n + 50):
        assert V.verify_slice(wire.parse_encoded_slice(slices[i]), vps, com, root, GEO, col) is None
        dv.record_slice(slices[i])
    col.msm = []  # the MSM equations go to the device instead
    assert col.fast_check(random.Random(9))
    assert dv.fast_check_msm()
