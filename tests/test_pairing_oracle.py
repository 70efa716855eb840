"""The verifier-side restatement (oracle/pairing.py, oracle/verify.py), no device work.

The pairing is pinned by what the mathematics fixes: the G2 generator lies on the
twist and has order r, e is bilinear in each argument, non-degenerate, and its values
have order r.  The closed form of gen_vanishes is checked against the inverse DFT the
reference computes (crates/amt/src/prove_params/generate.rs:234-263).  Slices of a
blob the oracle encoded pass EncodedSlice::verify (crates/encoder/src/encoder/
slice.rs:72-95) one by one and in one deferred batch, and the fault injections of the
reference's test_verify (crates/encoder/src/encoder/blob.rs:84-196) give its errors."""
import random

import numpy as np
import pytest

import oracle as O
import pairing as E
import verify as V
import zg_rocm as Z
from helpers import DATA_SEED, OracleOutputs, oracle_params, seeded_tau
from zg_rocm import wire

GEO = Z.TEST_SETTING


def test_g2_generator():
    assert E.g2_on_curve(E.G2_GEN)
    assert E.g2_mul(E.G2_GEN, E.R) is None
    assert E.g2_mul(E.G2_GEN, E.R - 1) == E.g2_neg(E.G2_GEN)


def test_pairing_bilinear_nondegenerate():
    rng = random.Random(1)
    g, q = E.G1_GEN, E.G2_GEN
    e = E.pairing(g, q)
    assert e != E.F12_ONE
    assert E.f12_pow(e, E.R) == E.F12_ONE
    a, b = rng.randrange(1, E.R), rng.randrange(1, E.R)
    assert E.pairing(E.g1_mul(g, a), E.g2_mul(q, b)) == E.f12_pow(e, a * b % E.R)
    # linear in each argument
    p1, p2 = E.g1_mul(g, 5), E.g1_mul(g, 11)
    assert E.pairing(E.g1_add(p1, p2), q) == E.f12_mul(E.pairing(p1, q), E.pairing(p2, q))
    q1, q2 = E.g2_mul(q, 3), E.g2_mul(q, 8)
    assert E.pairing(g, E.g2_add(q1, q2)) == E.f12_mul(E.pairing(g, q1), E.pairing(g, q2))
    # the multi-pairing check used by the verifier
    assert E.multi_pairing_is_one([(E.g1_mul(g, 6), q), (E.g1_neg(g), E.g2_mul(q, 6))])
    assert not E.multi_pairing_is_one([(E.g1_mul(g, 6), q), (E.g1_neg(g), E.g2_mul(q, 7))])


def test_vanish_closed_form_matches_gen_vanishes():
    """V[t] = sum_i coeff[i] w_D^(-i t) with coeff[0] = [tau'^(N-m)], coeff[i] =
    [tau'^((i-1) m)] (gen_vanishes), against the closed form, for depths 1 and 2."""
    tau = seeded_tau()
    vp = V.VerifyParams(tau, GEO.log_col, GEO.log_row, 1, [])
    n = GEO.n
    for d in (0, 1):
        D = 2 << d
        m = n // D
        wD = pow(V.G28, 1 << (28 - (d + 1)), E.R)
        coeff = [pow(vp.tau_c, n - m, E.R)] + [pow(vp.tau_c, (i - 1) * m, E.R) for i in range(1, D)]
        for t in range(D):
            s = sum(c * pow(wD, (E.R - 1 - t * i % (E.R - 1)) % (E.R - 1), E.R) for i, c in enumerate(coeff)) % E.R
            k = V._bitrev(t, d + 1)
            assert vp.vanish(d, k) == E.g2_mul(E.G2_GEN, s), (d, t)


@pytest.fixture(scope="module")
def setup():
    tau = seeded_tau()
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    e = O.encode(params, GEO.log_col, GEO.log_row, Z.seeded_bytes(DATA_SEED + 600, GEO.max_raw_data_size))
    blob = Z.EncodedBlob(GEO, OracleOutputs(e))
    vps = []
    for c in range(GEO.coset_n):
        canon = O.fr_canonical_array(params[c].reshape(-1, 4), which=1).reshape(GEO.h + 2, GEO.n, 2, 4)  # Fq
        line = canon[GEO.h]
        basis = [V.g1_from_canon(line[i].tobytes()) for i in range(GEO.n)]
        vps.append(V.VerifyParams(tau, GEO.log_col, GEO.log_row, c, basis))
    return blob, vps, V.g1_from_canon(e.commitment.tobytes()), e.root.tobytes()


def test_slices_verify(setup):
    blob, vps, com, root = setup
    for i in (0, GEO.row_n + 17, GEO.row_encoded - 1):
        assert V.verify_slice(wire.parse_encoded_slice(blob.encoded_slice(i)), vps, com, root, GEO) is None, i


def test_deferred_batch(setup):
    blob, vps, com, root = setup
    rows = [0, 9, GEO.row_n + 33, 2 * GEO.row_n + 2, GEO.row_encoded - 1]
    col = V.Collector(True)
    for i in rows:
        assert V.verify_slice(wire.parse_encoded_slice(blob.encoded_slice(i)), vps, com, root, GEO, col) is None
    assert col.fast_check(random.Random(2))
    # the deepest-level quotient of one row swapped with its neighbour's: the batch fails
    col = V.Collector(True)
    for i in rows[:3]:
        s = wire.parse_encoded_slice(blob.encoded_slice(i))
        if i == 9:
            other = wire.parse_encoded_slice(blob.encoded_slice(8))
            last = GEO.log_row - 1
            assert other["amt_proof"][last] != s["amt_proof"][last]
            s["amt_proof"] = list(s["amt_proof"][:last]) + [(s["amt_proof"][last][0], other["amt_proof"][last][1])]
        assert V.verify_slice(s, vps, com, root, GEO, col) is None  # deferred: nothing checked yet
    assert not col.fast_check(random.Random(3))


def _plus_gen(b):
    """ark-wire G1 point + generator (the reference's `+ G1A::generator()`)"""
    pt = E.g1_add(V.g1_from_wire(b), E.G1_GEN)
    canon = bytes(64) if pt is None else pt[0].to_bytes(32, "little") + pt[1].to_bytes(32, "little")
    return Z.ark_g1_uncompressed(canon)


def _low_u64_plus_one(b):
    """H256::from_low_u64_be(H256(b).to_low_u64_be() + 1)"""
    return bytes(24) + ((int.from_bytes(bytes(b[24:]), "big") + 1) % (1 << 64)).to_bytes(8, "big")


def test_fault_injection(setup):
    """get_invalid_row + gen_err_signer_map (crates/encoder/src/encoder/blob.rs:84-196,
    amt/blob.rs:62-87, merkle/blob.rs:62-97) for one row."""
    blob, vps, com, root = setup
    index = GEO.row_n + 5
    coset, local = divmod(index, GEO.row_n)
    good = lambda i: wire.parse_encoded_slice(blob.encoded_slice(i))  # noqa: E731
    amt_keys = ("amt_index", "commitment", "row_index", "row", "amt_proof", "high_commitment")
    mk_keys = ("merkle_root", "merkle_proof", "leaf_index", "leaf")

    def compose(g_index, amt, merkle):
        s = {"index": g_index}
        s.update({k: amt[k] for k in amt_keys})
        s.update({k: merkle[k] for k in mk_keys})
        return s

    cases = {}
    cases["WrongIndex"] = (compose(index + 1, good(index), good(index)),
                           "UnmatchedAMTIndex { row_index: %d, amt_index: %d }" % (index + 1, index))
    cases["WrongAmtIndex"] = (compose(index, good(index + 1), good(index)),
                              "UnmatchedAMTIndex { row_index: %d, amt_index: %d }" % (index, index + 1))
    cases["WrongMerkleIndex"] = (compose(index, good(index), good(index + 1)),
                                 "UnmatchedMerkleIndex { row_index: %d, merkle_index: %d }" % (index, index + 1))
    a = good(index)
    row = [bytes(x) for x in a["row"]]
    row[0] = ((int.from_bytes(row[0], "little") + 1) % E.R).to_bytes(32, "little")
    a["row"] = row
    cases["AMT WrongRow"] = (a, "AMT(IncorrectProof { coset_index: %d, amt_index: %d, error: InconsistentCommitment })"
                             % (coset, local))
    a = good(index)
    a["row_index"] += 1
    cases["AMT WrongIndex"] = (a, "AMT(UnmatchedCosetIndex { coset_index: %d, local_index: %d, amt_index: %d })"
                               % (coset, local, local + 1))
    a = good(index)
    a["commitment"] = _plus_gen(a["commitment"])
    cases["AMT WrongCommitment"] = (a, "AMT(IncorrectCommitment)")
    a = good(index)
    a["high_commitment"] = _plus_gen(a["high_commitment"])
    cases["AMT IncorrectHighCommitment"] = (a, "AMT(IncorrectProof { coset_index: %d, amt_index: %d, error: "
                                               "FailedLowDegreeTest })" % (coset, local))
    m = good(index)
    m["leaf_index"] += 1
    cases["Merkle WrongIndex"] = (m, "UnmatchedMerkleIndex { row_index: %d, merkle_index: %d }" % (index, index + 1))
    m = good(index)
    m["leaf"] = _low_u64_plus_one(m["leaf"])
    cases["Merkle WrongLocalRoot"] = (m, "Merkle(IncorrectLocalRoot { row_index: %d })" % index)
    m = good(index)
    m["merkle_proof"] = [_low_u64_plus_one(m["merkle_proof"][0])] + list(m["merkle_proof"][1:])
    cases["Merkle WrongProof"] = (m, "Merkle(IncorrectProof { row_index: %d })" % index)
    m = good(index)
    m["merkle_root"] = [_low_u64_plus_one(x) for x in m["merkle_root"]]
    cases["Merkle WrongRoot"] = (m, "Merkle(IncorrectRoot)")
    for name, (s, want) in cases.items():
        assert V.verify_slice(s, vps, com, root, GEO) == want, name
