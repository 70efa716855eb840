"""EncodeBlobReply.encoded_slice bytes (serialize_uncompressed of EncodedBlob::get_row,
grpc/src/service.rs:93-95, 162-166) from the oracle's outputs, no device work.

The field order is the reference's struct declarations (encoder/slice.rs:15-20,
amt/slice.rs:13-18, amt/src/blob/encode.rs:203-209, merkle/slice.rs:12-20).  The
ark-serialize 0.4 encodings assumed for its leaves (u64 LE usize and Vec lengths, no
length for arrays, 32 B canonical Fr, G1 uncompressed x || y with the SW flags in y's
top bits) cannot be imported here: byte-level parity with ark is unpinned.  What pins
the layout: the slice sizes it gives (2,296 B per row at this geometry, 34,680 B at
1024 x 1024) and a field-by-field parse of every row against the oracle's arrays."""
import struct

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, OracleOutputs, oracle_params

GEO = Z.TEST_SETTING


def parse_slice(b, g):
    """Independent reader of the layout: returns the fields and checks it consumes b."""
    pos = 0

    def take(n):
        nonlocal pos
        v = b[pos:pos + n]
        assert len(v) == n
        pos += n
        return v

    u64 = lambda: struct.unpack("<Q", take(8))[0]  # noqa: E731

    def g1():
        raw = bytearray(take(64))
        flags = raw[63] & 0xC0
        raw[63] &= 0x3F
        return bytes(raw), flags

    f = {"index": u64(), "amt_index": u64(), "commitment": g1(), "row_index": u64()}
    n = u64()
    f["row"] = [take(32) for _ in range(n)]
    n = u64()
    f["proof"] = [(g1(), g1()) for _ in range(n)]
    f["high"] = g1()
    f["root"] = [take(32) for _ in range(g.coset_n)]
    n = u64()
    f["merkle_proof"] = [take(32) for _ in range(n)]
    f["leaf_index"] = u64()
    f["leaf"] = take(32)
    assert pos == len(b)
    return f


def check_flags(point_flags):
    p, flags = point_flags
    if not any(p):
        assert flags == 0x40, "infinity"
    else:
        y = int.from_bytes(p[32:], "little")
        assert flags == (0x80 if y > (Z.P_MOD - 1) // 2 else 0)


def slice_size(g):
    return (8 + 8 + 64 + 8 + (8 + 32 * g.col_n) + (8 + 128 * g.log_row) + 64 + 32 * g.coset_n
            + (8 + 32 * g.log_row) + 8 + 32)


def test_slice_sizes_match_the_reference_figures():
    assert slice_size(Z.TEST_SETTING) == 2296
    assert slice_size(Z.PRODUCTION) == 34680


def _blob(length):
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    e = O.encode(params, GEO.log_col, GEO.log_row, Z.seeded_bytes(DATA_SEED + 60, length))
    return e, Z.EncodedBlob(GEO, OracleOutputs(e))


def test_every_slice_parses_back_to_the_oracle_fields():
    e, blob = _blob(1234)
    g = GEO
    data = e.data.reshape(g.coset_n, g.row_n, g.col_n, 32)
    ct = e.comm_tree.reshape(g.coset_n, -1, 64)
    pt = e.proof_tree.reshape(g.coset_n, -1, 64)
    m = e.merkle.reshape(-1, 32)
    layer = lambda d: m[sum(g.coset_n << k for k in range(d)):][:g.coset_n << d]  # noqa: E731
    for index in range(g.row_encoded):
        b = blob.encoded_slice(index)
        assert len(b) == slice_size(g)
        f = parse_slice(b, g)
        coset, local = divmod(index, g.row_n)
        assert f["index"] == f["amt_index"] == f["leaf_index"] == index
        assert f["row_index"] == local
        assert f["commitment"][0] == e.commitment.tobytes()
        assert b"".join(f["row"]) == data[coset, local].tobytes()
        rev = int(format(local, "0%db" % g.log_row)[::-1], 2)
        for d, (com, prf) in enumerate(f["proof"]):
            at = (2 << d) - 2 + ((rev >> (g.log_row - 1 - d)) ^ 1)
            assert com[0] == ct[coset, at].tobytes() and prf[0] == pt[coset, at].tobytes()
            check_flags(com)
            check_flags(prf)
        check_flags(f["commitment"])
        check_flags(f["high"])
        assert f["high"][0] == e.high.reshape(-1, 64)[coset].tobytes()
        assert b"".join(f["root"]) == layer(0).tobytes()
        assert f["merkle_proof"] == [layer(d)[(index >> (g.log_row - d)) ^ 1].tobytes()
                                     for d in range(g.log_row, 0, -1)]
        assert f["leaf"] == layer(g.log_row)[index].tobytes()


def test_empty_blob_slices_flag_infinity():
    # RawData of length 0 pads to an all-zero blob: every AMT point is the identity
    e, blob = _blob(0)
    assert not e.commitment.any() and not e.comm_tree.any() and not e.proof_tree.any()
    f = parse_slice(blob.encoded_slice(GEO.row_encoded - 1), GEO)
    for pf in [f["commitment"], f["high"]] + [x for pair in f["proof"] for x in pair]:
        assert pf == (bytes(64), 0x40)


def test_flag_rule_on_both_halves():
    # y and -y of the generator (1, 2): exactly one carries 0x80
    g_aff = (1).to_bytes(32, "little") + (2).to_bytes(32, "little")
    neg = (1).to_bytes(32, "little") + (Z.P_MOD - 2).to_bytes(32, "little")
    assert Z.ark_g1_uncompressed(g_aff)[63] & 0xC0 == 0
    assert Z.ark_g1_uncompressed(neg)[63] & 0xC0 == 0x80
    assert Z.ark_g1_uncompressed(bytes(64)) == bytes(63) + b"\x40"
    half = ((Z.P_MOD - 1) // 2).to_bytes(32, "little")
    assert Z.ark_g1_uncompressed(bytes(31) + b"\x01" + half)[63] & 0x80 == 0  # y = (p-1)/2: y <= -y
