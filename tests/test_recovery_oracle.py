"""The oracle's RS recovery (oracle/zg_oracle.c zgo_recover, a restatement of
recover_from_da_slice, crates/recovery/src/lib.rs:15-43 and data_poly.rs:72-118), no
device work.

Pinned by what the mathematics fixes: a blob encoded by the oracle's encoder (itself
pinned in test_oracle_pins.py) and cut down to any ROW of its rows comes back byte
for byte, over the row sets of the reference's own test (data_poly.rs:183-231) and
random ones; erasures that leave too few rows, row ids past the encoded blob, a
corrupted extra row, and blobs whose decoded RawData is malformed end with the
reference's outcomes (lib.rs:34-39, raw_blob.rs:43-57, raw_data.rs:22-38)."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, oracle_params

GEO = Z.TEST_SETTING
R, RE, COL = GEO.row_n, GEO.row_encoded, GEO.col_n


def encoded_rows(e):
    w = 32 * COL
    return {i: e.data[i * w:(i + 1) * w].tobytes() for i in range(RE)}


def rows_of_blob(scalars_row_major):
    """Encoded rows of an arbitrary RawBlob (canonical ints, row-major): the coset
    extension of process_blob (crates/amt/src/blob/encode.rs:91-108) by the oracle."""
    g = GEO
    pts = [0] * g.n
    for i, v in enumerate(scalars_row_major):  # change_matrix_direction(COL_LOG, ROW_LOG)
        r, c = divmod(i, g.col_n)
        pts[c * g.row_n + r] = v
    mont = O.fr_to_mont_array(pts)
    out = {}
    for coset in range(g.coset_n):
        y = O.fr_from_mont_array(O.to_coset_blob(mont, g.log_n, coset))
        for r in range(g.row_n):
            out[coset * g.row_n + r] = b"".join(y[c * g.row_n + r].to_bytes(32, "little") for c in range(g.col_n))
    return out


@pytest.fixture(scope="module")
def blobs():
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    out = []
    for k, length in enumerate([0, 777, GEO.max_raw_data_size]):
        raw = Z.seeded_bytes(DATA_SEED + 200 + k, length)
        out.append((raw, encoded_rows(O.encode(params, GEO.log_col, GEO.log_row, raw))))
    return out


def reference_row_sets():
    """The row sets of test_data_poly_with_data (data_poly.rs:183-206)."""
    return [list(range(1, RE)), list(range(R)) + list(range(R + 1, RE)), list(range(R + 1, RE)),
            list(range(RE - R, RE)), list(range(R)), list(range(1, RE, 2))]


def test_round_trip_reference_row_sets(blobs):
    for raw, rows in blobs:
        for ids in reference_row_sets():
            rc, out = O.recover({i: rows[i] for i in ids}, GEO.log_col, GEO.log_row)
            assert (rc, out) == (0, raw), "rows %s..%s (%d)" % (ids[0], ids[-1], len(ids))


def test_round_trip_random_row_sets(blobs):
    rng = np.random.default_rng(17)
    raw, rows = blobs[2]
    for n in [R, R + 1, R + 7, 2 * R, RE - 1, RE]:
        ids = rng.choice(RE, n, replace=False)
        rc, out = O.recover({int(i): rows[int(i)] for i in ids}, GEO.log_col, GEO.log_row)
        assert (rc, out) == (0, raw), n


def test_reference_error_row_sets(blobs):
    _, rows = blobs[1]
    # (RE - R + 2 .. RE): R - 1 rows -> TooFewRowIds, which lib.rs reports first as
    # "Not enough valid lines"; (RE - R + 1 ..= RE): id RE overflows (data_poly.rs:207-216)
    few = {i: rows[i] for i in range(RE - R + 2, RE)}
    assert O.recover(few, GEO.log_col, GEO.log_row)[0] == 1
    over = {i: rows.get(i, rows[0]) for i in range(RE - R + 1, RE + 1)}
    assert O.recover(over, GEO.log_col, GEO.log_row)[0] == 2


def test_wrong_length_rows_are_dropped(blobs):
    raw, rows = blobs[1]
    s = {i: rows[i] for i in range(R)}
    s[R] = rows[R][:-1]  # dropped: the R rows left still suffice
    assert O.recover(s, GEO.log_col, GEO.log_row) == (0, raw)
    s = {i: rows[i] for i in range(R - 1)}
    s[R] = rows[R] + b"\x00"
    assert O.recover(s, GEO.log_col, GEO.log_row)[0] == 1


def test_non_canonical_element(blobs):
    _, rows = blobs[1]
    s = {i: rows[i] for i in range(RE)}
    bad = bytearray(s[5])
    bad[32 * 3:32 * 4] = O.R_MOD.to_bytes(32, "little")  # r itself
    s[5] = bytes(bad)
    assert O.recover(s, GEO.log_col, GEO.log_row)[0] == 6
    few = {i: rows[i] for i in range(3)}  # checked before the row count
    few[1] = bytes(bad)
    assert O.recover(few, GEO.log_col, GEO.log_row)[0] == 6


def test_inconsistent_rows(blobs):
    raw, rows = blobs[2]
    s = {i: rows[i] for i in range(R, RE)}
    b = bytearray(s[R + 9])
    b[40] ^= 1
    s[R + 9] = bytes(b)
    assert O.recover(s, GEO.log_col, GEO.log_row)[0] == 7  # redundancy: the degree assert fires
    minimal = {i: s[i] for i in range(R, 2 * R)}
    rc, _ = O.recover(minimal, GEO.log_col, GEO.log_row)  # no redundancy: some other blob decodes
    assert rc in (0, 3, 4, 5)


def test_arbitrary_blob_rows_round_trip():
    rng = np.random.default_rng(23)
    # random elements, the last one 0: its bytes 27..30 are the length field, so length 0
    vals = [int.from_bytes(rng.bytes(31), "little") for _ in range(GEO.n - 1)] + [0]
    rows = rows_of_blob(vals)
    ids = list(range(R, 2 * R))
    rc, out = O.recover({i: rows[i] for i in ids}, GEO.log_col, GEO.log_row)
    # the length field (last 4 of the 31 * N bytes) is 0: RawData of length 0 needs
    # an all-zero padding range, which random data is not
    assert rc == 5 and out == b""


def _blob_with(raw_padded):
    """Row-major canonical scalars of a padded RawData image (31 * N bytes)."""
    return [int.from_bytes(raw_padded[31 * i:31 * i + 31], "little") for i in range(GEO.n)]


def test_malformed_raw_data_outcomes():
    max_raw = GEO.max_raw_data_size
    ids = list(range(R + 3, 2 * R + 3))
    # a length field above MAX_RAW_DATA_SIZE
    img = bytearray(31 * GEO.n)
    img[max_raw:] = (max_raw + 1).to_bytes(4, "little")
    rows = rows_of_blob(_blob_with(img))
    assert O.recover({i: rows[i] for i in ids}, GEO.log_col, GEO.log_row)[0] == 4
    # a non-zero byte after the data
    img = bytearray(31 * GEO.n)
    img[:5] = b"hello"
    img[max_raw:] = (5).to_bytes(4, "little")
    img[9] = 1
    rows = rows_of_blob(_blob_with(img))
    assert O.recover({i: rows[i] for i in ids}, GEO.log_col, GEO.log_row)[0] == 5
    img[9] = 0
    rows = rows_of_blob(_blob_with(img))
    assert O.recover({i: rows[i] for i in ids}, GEO.log_col, GEO.log_row) == (0, b"hello")
    # an element with a non-zero top byte
    vals = _blob_with(bytearray(31 * GEO.n))
    vals[7] = 1 << 250
    rows = rows_of_blob(vals)
    assert O.recover({i: rows[i] for i in ids}, GEO.log_col, GEO.log_row)[0] == 3


def degree_n_rows(rows, coeff, extra=0):
    """The encoded rows of D'(x) = D(x) + coeff x^(N + extra): row l of coset k holds
    D(x) at x = cf_k w_N^(l + ROW j), j < COL, so every element gains coeff x^(N + extra)."""
    g = GEO
    w = O.domain_root(g.log_n)
    out = {}
    for i, row in rows.items():
        k, l = divmod(i, g.row_n)
        cf = O.coset_factor(g.log_n, k)
        vals = []
        for j in range(g.col_n):
            x = cf * pow(w, l + g.row_n * j, O.R_MOD) % O.R_MOD
            v = int.from_bytes(row[32 * j:32 * j + 32], "little")
            vals.append(((v + coeff * pow(x, g.n + extra, O.R_MOD)) % O.R_MOD).to_bytes(32, "little"))
        out[i] = b"".join(vals)
    return out


def test_degree_n_rows_fold_onto_coset0(blobs):
    """Rows on a polynomial of degree exactly N pass the reference's degree assert
    (data_coeffs.len() <= RAW_BLOB_SIZE + 1, data_poly.rs:112) and coeffs_to_evals folds the
    top coefficient c onto coset 0 (every coset-0 value gains c, utils.rs:81-105: the
    2N-point FFT taken at every second point).  With c = 1 and a full-length blob the
    recovered bytes are the original 31-byte chunks each plus one."""
    raw, rows = blobs[2]
    assert len(raw) == GEO.max_raw_data_size
    padded = raw + len(raw).to_bytes(4, "little")
    chunks = [int.from_bytes(padded[31 * i:31 * i + 31], "little") + 1 for i in range(GEO.n)]
    assert all(ch < 1 << 248 for ch in chunks)
    want = b"".join(ch.to_bytes(31, "little") for ch in chunks)[:len(raw)]
    bent = degree_n_rows(rows, 1)
    for ids in (list(range(R, RE)), list(range(2 * R, RE)) + [R + 5], sorted(set(range(0, RE, 2)) | {1})):
        rc, out = O.recover({i: bent[i] for i in ids}, GEO.log_col, GEO.log_row)
        assert (rc, out) == (0, want), len(ids)
    # exactly ROW rows fix a polynomial of degree < N: the bent rows of coset 2 alone
    # decode to other data (no fold is visible from ROW rows)
    rc, out = O.recover({i: bent[i] for i in range(2 * R, RE)}, GEO.log_col, GEO.log_row)
    assert out != want


def test_degree_above_n_panics(blobs):
    """Degree N + 1, or rows whose extra terms do not agree on one coefficient: the
    degree assert fires (the reference panics)."""
    raw, rows = blobs[2]
    bent = degree_n_rows(rows, 1, extra=1)
    rc, _ = O.recover({i: bent[i] for i in range(R, RE)}, GEO.log_col, GEO.log_row)
    assert rc == 7
    mixed = degree_n_rows(rows, 1)
    mixed.update({i: v for i, v in degree_n_rows(rows, 2).items() if i == RE - 1})
    rc, _ = O.recover({i: mixed[i] for i in range(R, RE)}, GEO.log_col, GEO.log_row)
    assert rc == 7
