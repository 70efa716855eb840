"""The slice checks DeferredVerifier.record_slice makes before it records a row, on the
CPU (no device work): each of the reference's rejections in its order, with its text
(EncodedSlice::verify, crates/encoder/src/encoder/slice.rs:72-95; EncodedSliceAMT::
verify, crates/encoder/src/amt/slice.rs:51-88; verify_amt_proof's length and position
checks, crates/amt/src/verify_params.rs:149-159).  Slices come from the oracle's encode
at the test geometry."""
import functools

import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, OracleOutputs, oracle_params
from zg_rocm import wire

GEO = Z.TEST_SETTING


class _Geo:
    """Stands in for a Context: record_slice reads only the geometry."""

    geo = GEO


@functools.lru_cache(maxsize=None)
def _blob():
    e = O.encode(list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n)), GEO.log_col, GEO.log_row,
                 Z.seeded_bytes(DATA_SEED + 3, 5000))
    return Z.EncodedBlob(GEO, OracleOutputs(e))


def _verifier():
    return Z.DeferredVerifier(_Geo(), _blob().get_affine_commitment())


def _slice(i, **changes):
    f = wire.parse_encoded_slice(_blob().encoded_slice(i))
    f.update(changes)
    return wire.serialize_encoded_slice(f)


def test_honest_slices_are_recorded():
    v = _verifier()
    for i in (0, 5, GEO.row_n + 3, GEO.row_encoded - 1):
        v.record_slice(_blob().encoded_slice(i))
    assert sorted(v._rows) == [0, 5, GEO.row_n + 3, GEO.row_encoded - 1]
    with pytest.raises(ValueError, match="recorded twice"):
        v.record_slice(_blob().encoded_slice(5))


def _rejects(slice_bytes, text):
    with pytest.raises(Z.VerifierError) as e:
        _verifier().record_slice(slice_bytes)
    assert str(e.value) == text


def test_unmatched_amt_index():
    _rejects(_slice(7, amt_index=8), "UnmatchedAMTIndex { row_index: 7, amt_index: 8 }")


def test_unmatched_merkle_index():
    _rejects(_slice(7, leaf_index=9), "UnmatchedMerkleIndex { row_index: 7, merkle_index: 9 }")


def test_incorrect_commitment():
    other = wire.parse_encoded_slice(_blob().encoded_slice(3))["amt_proof"][0][0]  # some other G1 point
    _rejects(_slice(7, commitment=other), "AMT(IncorrectCommitment)")


def test_incorrect_row_size():
    f = wire.parse_encoded_slice(_blob().encoded_slice(7))
    _rejects(_slice(7, row=f["row"][:-1]), "AMT(IncorrectRowSize { actual: %d, expected: %d })" % (GEO.col_n - 1,
                                                                                                  GEO.col_n))


def test_row_index_overflow():
    i = GEO.row_encoded
    _rejects(_slice(3, index=i, amt_index=i, leaf_index=i),
             "AMT(RowIndexOverflow { actual: %d, expected_max: %d })" % (i, GEO.row_encoded))


def test_unmatched_coset_index():
    i = GEO.row_n + 5
    _rejects(_slice(i, row_index=4),
             "AMT(UnmatchedCosetIndex { coset_index: 1, local_index: 5, amt_index: 4 })")


def test_unexpected_data_length():
    f = wire.parse_encoded_slice(_blob().encoded_slice(9))
    _rejects(_slice(9, amt_proof=f["amt_proof"][:-1]),
             "AMT(IncorrectProof { coset_index: 0, amt_index: 9, error: UnexpectedDataLength })")


def test_commitment_argument():
    with pytest.raises(ValueError):
        Z.DeferredVerifier(_Geo(), bytes(63))
