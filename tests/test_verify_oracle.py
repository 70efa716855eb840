"""The deferred verification MSM (DeferredVerifier::record_msm + fast_check_msm,
crates/amt/src/deferred_verification.rs:71-92, 133-150) in the oracle, no device work.

Pinned by the verification identity itself: for every row of a blob the oracle
encoded, sum_i basis[bitrev(l) COL + i] index_reverse(row)[i] equals the commitment
minus the row's proof commitments (verify_amt_proof, crates/amt/src/verify_params.rs:
176-186), so alpha-weighted sums over any set of rows agree; a changed element, a
changed proof point or a row recorded under another index breaks it."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, OracleOutputs, oracle_params
from zg_rocm import wire

GEO = Z.TEST_SETTING


def g1_from_wire(b):
    raw = bytearray(b)
    inf = raw[63] & 0x40
    raw[63] &= 0x3F
    return bytes(64) if inf else bytes(raw)


def neg(p):
    if not any(p):
        return p
    y = int.from_bytes(p[32:], "little")
    return p[:32] + ((O.P_MOD - y) % O.P_MOD).to_bytes(32, "little")


def _a(b):
    return np.frombuffer(bytes(b), dtype=np.uint8)


def answer(slices, alphas):
    """sum_k alpha_k (commitment - sum_d proof_k[d].commitment) with the oracle's G1 ops"""
    acc = bytes(64)
    for s, a in zip(slices, alphas):
        d = s["commitment"]
        d = g1_from_wire(d)
        for com, _ in s["amt_proof"]:
            d = bytes(O.g1_add(_a(d), _a(neg(g1_from_wire(com)))))
        acc = bytes(O.g1_add(_a(acc), O.g1_mul(_a(d), a)))
    return acc


@pytest.fixture(scope="module")
def blob():
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    e = O.encode(params, GEO.log_col, GEO.log_row, Z.seeded_bytes(DATA_SEED + 400, 30000))
    b = Z.EncodedBlob(GEO, OracleOutputs(e))
    return params, [wire.parse_encoded_slice(b.encoded_slice(i)) for i in range(GEO.row_encoded)]


def test_identity_over_row_sets(blob):
    params, sl = blob
    rng = np.random.default_rng(3)
    for ids in [[0], [GEO.row_n + 5], list(range(GEO.row_encoded)),
                sorted(int(i) for i in rng.choice(GEO.row_encoded, 40, replace=False))]:
        alphas = [int(rng.integers(1, 2**62)) * 7919 + 1 for _ in ids]
        got = O.verify_msm(params, GEO.log_col, GEO.log_row, ids, b"".join(b"".join(sl[i]["row"]) for i in ids), alphas)
        assert got == answer([sl[i] for i in ids], alphas), ids[:3]


def test_identity_breaks_on_tampering(blob):
    params, sl = blob
    ids = [3, GEO.row_n + 9, 2 * GEO.row_n + 1]
    alphas = [11, 13, 17]
    rows = [bytearray(b"".join(sl[i]["row"])) for i in ids]
    want = answer([sl[i] for i in ids], alphas)
    assert O.verify_msm(params, GEO.log_col, GEO.log_row, ids, b"".join(map(bytes, rows)), alphas) == want
    rows[1][64] ^= 1  # one element of one row
    assert O.verify_msm(params, GEO.log_col, GEO.log_row, ids, b"".join(map(bytes, rows)), alphas) != want
    rows[1][64] ^= 1
    shifted = [4, GEO.row_n + 9, 2 * GEO.row_n + 1]  # row 3's data under index 4
    assert O.verify_msm(params, GEO.log_col, GEO.log_row, shifted, b"".join(map(bytes, rows)), alphas) != want
    bad = [dict(sl[i]) for i in ids]
    bad[2]["amt_proof"] = [(sl[ids[0]]["amt_proof"][0][0], None)] + list(sl[ids[2]]["amt_proof"][1:])
    assert answer(bad, alphas) != want


def test_bad_row_index(blob):
    params, sl = blob
    with pytest.raises(ValueError):
        O.verify_msm(params, GEO.log_col, GEO.log_row, [GEO.row_encoded], b"".join(sl[0]["row"]), [1])
