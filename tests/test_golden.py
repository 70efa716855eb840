"""The oracle against the committed encode fixtures (tests/golden/encode_test_geometry.json,
written by tests/golden/make_golden.py and pinned there by the AMT closed form, the raw
coset-0 rows and a textbook-Keccak file root): a change to the oracle that moves any
output of the four blobs fails here."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
import zg_rocm as Z

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "encode_test_geometry.json")


def load_golden():
    with open(GOLDEN) as f:
        return json.load(f)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_against_golden(g, out_commitment, out_high, out_root, proof_tree, comm_tree, merkle, data):
    """Compare one blob's outputs (library array layouts) with its fixture entry."""
    assert bytes(out_commitment).hex() == g["commitment"]
    assert bytes(np.asarray(out_high).reshape(-1)[:64]).hex() == g["high_commitment"]
    assert bytes(out_root).hex() == g["file_root"]
    flat = np.ascontiguousarray(proof_tree).view(np.uint8).reshape(-1)
    assert [bytes(flat[64 * i:64 * i + 64]).hex() for i in range(4)] == g["proof_tree_first_points"]
    assert digest(proof_tree) == g["proof_tree_sha256"]
    assert digest(comm_tree) == g["comm_tree_sha256"]
    assert digest(merkle) == g["merkle_sha256"]
    assert digest(data) == g["data_sha256"]


@pytest.fixture(scope="module")
def golden():
    return load_golden()


@pytest.fixture(scope="module")
def bases():
    geo = Z.TEST_SETTING
    tau = Z.tau_from_seed(7)
    return [O.params_from_tau(tau, geo.log_n, geo.log_row, c, 0) for c in range(geo.coset_n)]


def test_golden_geometry(golden):
    geo = Z.TEST_SETTING
    assert golden["geometry"] == {"log_col": geo.log_col, "log_row": geo.log_row, "cosets": geo.coset_n}
    assert [b["length"] for b in golden["blobs"]] == [0, 1, 1234, geo.max_raw_data_size]


@pytest.mark.parametrize("i", range(4))
def test_oracle_matches_golden(golden, bases, i):
    geo = Z.TEST_SETTING
    g = golden["blobs"][i]
    raw = Z.seeded_bytes(222, g["length"])
    assert hashlib.sha256(raw).hexdigest() == g["raw_sha256"]
    exp = O.encode(bases, geo.log_col, geo.log_row, raw)
    check_against_golden(g, exp.commitment, exp.high, exp.root, exp.proof_tree, exp.comm_tree, exp.merkle, exp.data)
