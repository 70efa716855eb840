"""RS recovery on the device (zgr_recover, recover_from_da_slice of
crates/recovery/src/lib.rs:15-43) against the oracle's restatement of the reference
algorithm, at the test geometry: the same bytes or the same outcome for the row sets
of the reference's own test (data_poly.rs:183-216), random row sets, rows of the
wrong length, out-of-range elements, a corrupted extra row and malformed RawData."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, oracle_params
from test_recovery_oracle import _blob_with, degree_n_rows, encoded_rows, reference_row_sets, rows_of_blob

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING
R, RE = GEO.row_n, GEO.row_encoded

OUTCOME = {1: "Not enough valid lines", 2: "Cannot recover data: RowIdOverflow", 3: "Incorrect scalar",
           4: "Incorrect length field", 5: "Non zero in padding range"}


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)  # recovery needs no params
    yield c
    c.close()


@pytest.fixture(scope="module")
def blobs():
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    out = []
    for k, length in enumerate([0, 1, 777, GEO.max_raw_data_size]):
        raw = Z.seeded_bytes(DATA_SEED + 300 + k, length)
        out.append((raw, encoded_rows(O.encode(params, GEO.log_col, GEO.log_row, raw))))
    return out


def same_outcome(ctx, slices):
    """GPU and oracle agree on slices; returns the oracle's status."""
    rc, want = O.recover(slices, GEO.log_col, GEO.log_row)
    if rc == 0:
        assert Z.recover_from_da_slice(ctx, slices) == want
    elif rc in OUTCOME:
        with pytest.raises(Z.RecoveryError) as ei:
            Z.recover_from_da_slice(ctx, slices)
        assert str(ei.value) == OUTCOME[rc]
    else:  # 6, 7: the reference panics
        with pytest.raises(Z.RecoveryPanic):
            Z.recover_from_da_slice(ctx, slices)
    return rc


def test_reference_row_sets(ctx, blobs):
    for raw, rows in blobs:
        for ids in reference_row_sets():
            assert same_outcome(ctx, {i: rows[i] for i in ids}) == 0
            assert Z.recover_from_da_slice(ctx, {i: rows[i] for i in ids}) == raw


def test_random_row_sets(ctx, blobs):
    rng = np.random.default_rng(29)
    raw, rows = blobs[3]
    for n in [R, R + 1, R + 5, R + 33, 2 * R - 1, 2 * R, RE - 2, RE]:
        for _ in range(2):
            ids = rng.choice(RE, n, replace=False)
            s = {int(i): rows[int(i)] for i in ids}
            assert same_outcome(ctx, s) == 0, n
            assert Z.recover_from_da_slice(ctx, s) == raw


def test_reference_error_row_sets(ctx, blobs):
    _, rows = blobs[2]
    assert same_outcome(ctx, {i: rows[i] for i in range(RE - R + 2, RE)}) == 1
    assert same_outcome(ctx, {i: rows.get(i, rows[0]) for i in range(RE - R + 1, RE + 1)}) == 2
    assert same_outcome(ctx, {}) == 1


def test_wrong_length_rows(ctx, blobs):
    raw, rows = blobs[2]
    s = {i: rows[i] for i in range(R)}
    s[R] = rows[R][:-1]
    assert same_outcome(ctx, s) == 0
    s = {i: rows[i] for i in range(R - 1)}
    s[R] = rows[R] + b"\x00"
    assert same_outcome(ctx, s) == 1


def test_non_canonical_element(ctx, blobs):
    _, rows = blobs[2]
    bad = bytearray(rows[5])
    bad[96:128] = O.R_MOD.to_bytes(32, "little")
    s = {i: rows[i] for i in range(RE)}
    s[5] = bytes(bad)
    assert same_outcome(ctx, s) == 6
    few = {0: rows[0], 1: bytes(bad)}
    assert same_outcome(ctx, few) == 6


def test_corrupted_extra_row(ctx, blobs):
    raw, rows = blobs[3]
    for where in [R + 9, 2, 2 * R + 40]:
        s = {i: rows[i] for i in range(RE)}
        b = bytearray(s[where])
        b[40] ^= 1
        s[where] = bytes(b)
        assert same_outcome(ctx, s) == 7, where
    # one row beyond the minimum is enough redundancy
    s = {i: rows[i] for i in range(R + 1)}
    b = bytearray(s[R])
    b[0] ^= 2
    s[R] = bytes(b)
    assert same_outcome(ctx, s) == 7


def test_malformed_raw_data(ctx):
    max_raw = GEO.max_raw_data_size
    ids = list(range(R + 3, 2 * R + 3)) + [RE - 1]
    img = bytearray(31 * GEO.n)
    img[max_raw:] = (max_raw + 1).to_bytes(4, "little")
    rows = rows_of_blob(_blob_with(img))
    assert same_outcome(ctx, {i: rows[i] for i in ids}) == 4
    img = bytearray(31 * GEO.n)
    img[:5] = b"hello"
    img[max_raw:] = (5).to_bytes(4, "little")
    img[9] = 1
    rows = rows_of_blob(_blob_with(img))
    assert same_outcome(ctx, {i: rows[i] for i in ids}) == 5
    vals = _blob_with(bytearray(31 * GEO.n))
    vals[7] = 1 << 250
    rows = rows_of_blob(vals)
    assert same_outcome(ctx, {i: rows[i] for i in ids}) == 3


def test_abi_argument_errors(ctx):
    with pytest.raises(Z.ZgrError):
        ctx.recover([3, 2], bytes(2 * 32 * GEO.col_n))  # not increasing


def test_recover_into_host_arrays(ctx, blobs):
    """recover_into over caller-owned arrays, pageable and page-locked (hipHostMalloc):
    the same bytes as the oracle; undersized arrays are refused before the call."""
    raw, rows = blobs[3]
    ids = list(range(R + 2, 2 * R + 2))
    buf = b"".join(rows[i] for i in ids)
    pin_rows, pin_out = Z.PinnedHostArray(len(buf)), Z.PinnedHostArray(GEO.max_raw_data_size)
    try:
        pin_rows.array[:] = np.frombuffer(buf, dtype=np.uint8)
        for ra, oa in ((np.frombuffer(buf, dtype=np.uint8), np.empty(GEO.max_raw_data_size, dtype=np.uint8)),
                       (pin_rows.array, pin_out.array)):
            n = ctx.recover_into(ids, ra, oa)
            assert oa[:n].tobytes() == raw
    finally:
        pin_rows.free()
        pin_out.free()
    with pytest.raises(ValueError):
        ctx.recover_into(ids, np.frombuffer(buf[:-1], dtype=np.uint8), np.empty(GEO.max_raw_data_size, np.uint8))
    with pytest.raises(Z.ZgrError):  # raw_out below MAX_RAW_DATA_SIZE: the C ABI refuses it
        ctx.recover_into(ids, np.frombuffer(buf, dtype=np.uint8), np.empty(GEO.max_raw_data_size - 1, np.uint8))


def test_gemm_paths(ctx, blobs):
    """The interpolation GEMM on the matrix cores (default) and on the VALU
    (ZGR_REC_GEMM=valu, read at every call) recover the same bytes, over row sets that
    erase every coset-0 row, some of them, or none."""
    import os
    raw, rows = blobs[3]
    sets = [list(range(2 * R, 3 * R)), list(range(5, R)) + list(range(R, R + 5)), list(range(R))]
    old = os.environ.get("ZGR_REC_GEMM")
    try:
        for path in ("valu", "mfma"):
            os.environ["ZGR_REC_GEMM"] = path
            for ids in sets:
                assert Z.recover_from_da_slice(ctx, {i: rows[i] for i in ids}) == raw, (path, ids[0])
    finally:
        if old is None:
            os.environ.pop("ZGR_REC_GEMM", None)
        else:
            os.environ["ZGR_REC_GEMM"] = old


def test_degree_n_rows_match_oracle(ctx, blobs):
    """Rows on a polynomial of degree exactly N (the reference folds its top coefficient
    onto coset 0) decode to the oracle's bytes; degree N + 1, or extra rows that disagree
    on the coefficient, give the reference's panic."""
    raw, rows = blobs[3]
    for coeff in (1, 5):
        bent = degree_n_rows(rows, coeff)
        for ids in (list(range(R, RE)), list(range(2 * R, RE)) + [R + 5], sorted(set(range(0, RE, 2)) | {1}),
                    list(range(RE))):
            assert same_outcome(ctx, {i: bent[i] for i in ids}) in (0, 3, 4, 5), (coeff, len(ids))
    bent = degree_n_rows(rows, 1)
    assert same_outcome(ctx, {i: bent[i] for i in range(R, RE)}) == 0
    above = degree_n_rows(rows, 1, extra=1)
    assert same_outcome(ctx, {i: above[i] for i in range(R, RE)}) == 7
    mixed = degree_n_rows(rows, 1)
    mixed.update({i: v for i, v in degree_n_rows(rows, 2).items() if i == RE - 1})
    assert same_outcome(ctx, {i: mixed[i] for i in range(R, RE)}) == 7
