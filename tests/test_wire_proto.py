"""EncodeBlob wire format, no device work: the protobuf codec of grpc/proto/encoder.proto
checked against the protobuf runtime (google.protobuf, messages built from a
descriptor of the same .proto), LightEncodedSlice (encoder/light_slice.rs:11-79)
against the EncodedSlice bytes, the all-rows slice gather against the per-row
serializer, and the gRPC front end over loopback with the oracle standing in for
the device (grpc/src/service.rs:47-106, grpc/src/lib.rs:11-33)."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, OracleOutputs, oracle_params
from zg_rocm import wire

GEO = Z.TEST_SETTING


# ---------------------------------------------------------------------------
# protobuf runtime messages for encoder.proto
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def pb():
    from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

    F = descriptor_pb2.FieldDescriptorProto
    fdp = descriptor_pb2.FileDescriptorProto(name="encoder_test.proto", package="encoder", syntax="proto3")
    req = fdp.message_type.add(name="EncodeBlobRequest")
    req.field.add(name="data", number=1, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)
    req.field.add(name="require_data", number=2, type=F.TYPE_BOOL, label=F.LABEL_OPTIONAL)
    rep = fdp.message_type.add(name="EncodeBlobReply")
    rep.field.add(name="version", number=1, type=F.TYPE_UINT32, label=F.LABEL_OPTIONAL)
    rep.field.add(name="erasure_commitment", number=2, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)
    rep.field.add(name="storage_root", number=3, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)
    rep.field.add(name="encoded_data", number=4, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)
    rep.field.add(name="encoded_slice", number=5, type=F.TYPE_BYTES, label=F.LABEL_REPEATED)
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fdp)
    get = message_factory.GetMessageClass
    return get(pool.FindMessageTypeByName("encoder.EncodeBlobRequest")), \
        get(pool.FindMessageTypeByName("encoder.EncodeBlobReply"))


REQUESTS = [(b"", False), (b"", True), (b"\x00", False), (bytes(range(256)) * 3, True), (b"x" * 200000, False)]


@pytest.mark.parametrize("data,require", REQUESTS)
def test_request_codec_matches_protobuf(pb, data, require):
    Req, _ = pb
    m = Req(data=data, require_data=require)
    assert wire.encode_request(data, require) == m.SerializeToString(deterministic=True)
    assert wire.decode_request(m.SerializeToString()) == (data, require)


def _replies(rng):
    yield {"version": 0, "erasure_commitment": b"", "storage_root": b"", "encoded_data": b"", "encoded_slice": []}
    yield {"version": 0, "erasure_commitment": rng.bytes(64), "storage_root": rng.bytes(32),
           "encoded_data": b"", "encoded_slice": [rng.bytes(2296) for _ in range(5)]}
    yield {"version": 7, "erasure_commitment": rng.bytes(64), "storage_root": rng.bytes(32),
           "encoded_data": rng.bytes(70000), "encoded_slice": [b"", rng.bytes(1), b"", rng.bytes(130)]}
    yield {"version": 0xFFFFFFFF, "erasure_commitment": b"\x00" * 64, "storage_root": b"\x00" * 32,
           "encoded_data": b"\x00" * 300, "encoded_slice": [b"\x00" * 127, b"\x01" * 128, b"\x02" * 16384]}


def test_reply_codec_matches_protobuf(pb):
    _, Rep = pb
    for r in _replies(np.random.default_rng(5)):
        m = Rep(**r)
        ref = m.SerializeToString(deterministic=True)
        assert wire.encode_reply(r) == ref
        back = Rep()
        back.ParseFromString(wire.encode_reply(r))
        assert back == m
        d = wire.decode_reply(ref)
        assert d == {k: (list(v) if k == "encoded_slice" else v) for k, v in r.items()}


def test_reply_slice_array_form_equals_list_form(pb):
    rng = np.random.default_rng(6)
    arr = rng.integers(0, 256, (9, 300), dtype=np.uint8)
    r = {"version": 0, "erasure_commitment": rng.bytes(64), "storage_root": rng.bytes(32), "encoded_data": b""}
    a = wire.encode_reply(dict(r, encoded_slice=arr))
    b = wire.encode_reply(dict(r, encoded_slice=[x.tobytes() for x in arr]))
    assert a == b
    _, Rep = pb
    assert Rep(**dict(r, encoded_slice=[x.tobytes() for x in arr])).SerializeToString(deterministic=True) == a


def test_decoder_skips_unknown_fields_and_keeps_last_scalar(pb):
    Req, _ = pb
    body = wire.encode_request(b"abc", True)
    # unknown varint field 9, unknown fixed64 field 10, unknown fixed32 field 11, unknown bytes field 12
    extra = b"\x48\x05" + b"\x51" + bytes(8) + b"\x5d" + bytes(4) + b"\x62\x02hi"
    msg = extra + body + b"\x10\x00"  # require_data repeated: last value wins
    m = Req()
    m.ParseFromString(msg)
    assert wire.decode_request(msg) == (m.data, m.require_data) == (b"abc", False)


@pytest.mark.parametrize("bad", [b"\x0a\x05ab", b"\x0a", b"\x10", b"\x08" + b"\xff" * 11, b"\x0b", b"\x00\x01",
                                 b"\x12\x02", b"\x0d\x01\x02"])
def test_decoder_rejects_malformed(bad):
    with pytest.raises(wire.WireError):
        wire.decode_request(bad)


def test_decoder_rejects_wrong_wire_type():
    with pytest.raises(wire.WireError):
        wire.decode_request(b"\x08\x01")  # field 1 (bytes) as a varint
    with pytest.raises(wire.WireError):
        wire.decode_reply(b"\x2d" + bytes(4))  # field 5 (bytes) as fixed32


# ---------------------------------------------------------------------------
# slices
# ---------------------------------------------------------------------------
def _blob(length, seed=DATA_SEED + 70):
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    e = O.encode(params, GEO.log_col, GEO.log_row, Z.seeded_bytes(seed, length))
    return e, Z.EncodedBlob(GEO, OracleOutputs(e))


def light_size(g):
    return 8 + 64 + (8 + 128 * g.log_row) + 64 + 32 * g.coset_n + (8 + 32 * g.log_row) + 32


@pytest.mark.parametrize("length", [0, 1234, GEO.max_raw_data_size])
def test_all_rows_gather_equals_per_row_serializer(length):
    _, blob = _blob(length)
    arr = blob.encoded_slices()
    assert arr.shape == (GEO.row_encoded, 2296)
    for i in range(GEO.row_encoded):
        assert arr[i].tobytes() == blob.encoded_slice(i), "row %d" % i


def test_light_slice_round_trip():
    _, blob = _blob(1234)
    g = GEO
    for i in range(g.row_encoded):
        full = blob.encoded_slice(i)
        light = blob.light_slice(i)
        assert len(light) == light_size(g)
        assert wire.light_from_slice(full) == light
        f = wire.parse_encoded_slice(full)
        # into_slice puts the row back: the same EncodedSlice bytes
        assert wire.light_into_slice(light, f["row"], g.row_n) == full
        lf = wire.parse_light_slice(light)
        assert lf["index"] == i and lf["leaf"] == f["leaf"] and lf["amt_proof"] == f["amt_proof"]


def test_light_slice_sizes():
    # the full slice (2,296 / 34,680 B) less the two extra indices, the row vector and
    # the leaf index that only EncodedSlice carries
    for g, full in ((Z.TEST_SETTING, 2296), (Z.PRODUCTION, 34680)):
        assert light_size(g) == full - (8 + 8 + 8 + 32 * g.col_n + 8)
    assert (light_size(Z.TEST_SETTING), light_size(Z.PRODUCTION)) == (1240, 1880)


def test_light_from_slice_rejects_mismatched_indices():
    _, blob = _blob(1234)
    f = wire.parse_encoded_slice(blob.encoded_slice(3))
    for key, name in (("amt_index", "UnmatchedAMTIndex"), ("leaf_index", "UnmatchedMerkleIndex")):
        bad = dict(f, **{key: 4})
        with pytest.raises(wire.WireError, match=name):
            wire.light_from_slice(wire.serialize_encoded_slice(bad))


def test_slice_parser_rejects_truncation_and_trailing_bytes():
    _, blob = _blob(1234)
    full = blob.encoded_slice(0)
    with pytest.raises(wire.WireError):
        wire.parse_encoded_slice(full[:-1])
    with pytest.raises(wire.WireError):
        wire.parse_encoded_slice(full + b"\x00")
    with pytest.raises(wire.WireError):
        wire.parse_light_slice(blob.light_slice(0)[:-5])


# ---------------------------------------------------------------------------
# gRPC front end over loopback, the oracle in place of the device
# ---------------------------------------------------------------------------
class OracleCtx:
    """Test stand-in for Context.encode: the oracle's outputs in zgr_encoded shape."""

    geo = GEO

    def encode(self, raw, want_data=True):
        params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
        return OracleOutputs(O.encode(params, GEO.log_col, GEO.log_row, bytes(raw)))


@pytest.fixture(scope="module")
def front():
    pytest.importorskip("grpc")
    from zg_rocm import grpc_front

    service = Z.EncoderService(OracleCtx())
    server, port = grpc_front.serve(service, "127.0.0.1:0", max_workers=2)
    client = grpc_front.EncoderClient("127.0.0.1:%d" % port)
    yield service, client
    client.close()
    server.stop(0)


@pytest.mark.parametrize("length,require", [(1, True), (1234, False), (GEO.max_raw_data_size, True)])
def test_grpc_encode_blob(front, length, require):
    service, client = front
    data = Z.seeded_bytes(DATA_SEED + length, length)
    got = client.encode_blob(data, require)
    want = service.process_data(data, require)
    e = OracleCtx().encode(data)
    assert got["version"] == 0
    assert got["erasure_commitment"] == want["erasure_commitment"] == e.commitment.tobytes()
    assert got["storage_root"] == want["storage_root"] == e.file_root.tobytes()
    assert got["encoded_data"] == (e.data.tobytes() if require else b"")
    assert got["encoded_slice"] == want["encoded_slice"]
    assert len(got["encoded_slice"]) == GEO.row_encoded
    assert all(s == want["blob"].encoded_slice(i) for i, s in enumerate(got["encoded_slice"]))
    assert client.encode_blob_raw(data, require) == service.encode_blob(data, require)


def test_grpc_too_large_blob(front):
    import grpc

    _, client = front
    n = GEO.max_raw_data_size + 1
    with pytest.raises(grpc.RpcError) as ei:
        client.encode_blob(bytes(n), True)
    assert ei.value.code() == grpc.StatusCode.UNKNOWN
    assert ei.value.details() == "TooLargeBlob { actual: %d, expected_max: %d }" % (n, GEO.max_raw_data_size)


def test_grpc_message_limit():
    grpc = pytest.importorskip("grpc")
    from zg_rocm import grpc_front

    server, port = grpc_front.serve(Z.EncoderService(OracleCtx()), "127.0.0.1:0", max_workers=1,
                                    message_limit=100000)
    try:
        with grpc_front.EncoderClient("127.0.0.1:%d" % port, message_limit=100000) as c:
            with pytest.raises(grpc.RpcError) as ei:
                c.encode_blob(b"abc", True)  # the reply (~450 KB) exceeds the limit
            assert ei.value.code() == grpc.StatusCode.RESOURCE_EXHAUSTED
    finally:
        server.stop(0)
