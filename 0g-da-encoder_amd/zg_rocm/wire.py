"""Wire formats of the encoding service.

* The protobuf messages of the ``encoder.Encoder`` service (grpc/proto/encoder.proto:5-21),
  proto3 encoding: a field with its default value (0, false, empty bytes) is not
  written; every element of ``repeated bytes`` is written, empty ones included.
* ``EncodedSlice`` and ``LightEncodedSlice`` as ark-serialize 0.4 writes them
  uncompressed (encoder/slice.rs:15-20, encoder/light_slice.rs:11-20): fields in
  declaration order, ``usize`` as u64 LE, ``Vec`` with a u64 LE length, arrays with
  none, Fr as 32 B canonical LE, G1 affine as x || y with the SW flags in y's top
  byte (ark_g1_uncompressed).  G1 points pass through these parsers as their 64 wire
  bytes; no curve check is made.
"""

MESSAGE_SIZE_LIMIT = 1024 * 1024 * 1024  # grpc/src/lib.rs:11, both directions

_WT_VARINT, _WT_I64, _WT_LEN, _WT_I32 = 0, 1, 2, 5


class WireError(ValueError):
    pass


# ---------------------------------------------------------------------------
# protobuf
# ---------------------------------------------------------------------------
def _varint(v):
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def _read_varint(buf, i):
    v = shift = 0
    while True:
        if i >= len(buf):
            raise WireError("truncated varint")
        b = buf[i]
        i += 1
        v |= (b & 0x7F) << shift
        if not b & 0x80:
            return v & ((1 << 64) - 1), i
        shift += 7
        if shift >= 70:
            raise WireError("varint too long")


def _len_field(num, payload):
    return _varint((num << 3) | _WT_LEN) + _varint(len(payload)) + bytes(payload)


def _fields(buf):
    """Yields (field number, wire type, value) over a message; value is an int for
    varint / fixed fields and a memoryview for length-delimited ones."""
    buf = memoryview(buf)
    i = 0
    while i < len(buf):
        key, i = _read_varint(buf, i)
        num, wt = key >> 3, key & 7
        if num == 0:
            raise WireError("field number 0")
        if wt == _WT_VARINT:
            v, i = _read_varint(buf, i)
        elif wt == _WT_I64:
            if i + 8 > len(buf):
                raise WireError("truncated fixed64")
            v, i = int.from_bytes(buf[i:i + 8], "little"), i + 8
        elif wt == _WT_I32:
            if i + 4 > len(buf):
                raise WireError("truncated fixed32")
            v, i = int.from_bytes(buf[i:i + 4], "little"), i + 4
        elif wt == _WT_LEN:
            n, i = _read_varint(buf, i)
            if i + n > len(buf):
                raise WireError("truncated length-delimited field")
            v, i = buf[i:i + n], i + n
        else:
            raise WireError("unsupported wire type %d" % wt)
        yield num, wt, v


def _expect(num, wt, want):
    if wt != want:
        raise WireError("field %d has wire type %d, expected %d" % (num, wt, want))


def encode_request(data, require_data=False):
    """EncodeBlobRequest { bytes data = 1; bool require_data = 2; }"""
    out = b""
    if len(data):
        out += _len_field(1, data)
    if require_data:
        out += _varint((2 << 3) | _WT_VARINT) + b"\x01"
    return out


def decode_request(buf):
    """-> (data, require_data); unknown fields are skipped, a repeated scalar keeps
    its last value (proto3 merge rules)."""
    data, req = b"", False
    for num, wt, v in _fields(buf):
        if num == 1:
            _expect(num, wt, _WT_LEN)
            data = bytes(v)
        elif num == 2:
            _expect(num, wt, _WT_VARINT)
            req = v != 0
    return data, req


def encode_reply(reply):
    """EncodeBlobReply { uint32 version = 1; bytes erasure_commitment = 2;
    bytes storage_root = 3; bytes encoded_data = 4; repeated bytes encoded_slice = 5; }
    from a dict with those keys (EncoderService.process_data)."""
    parts = []
    if reply.get("version", 0):
        parts.append(_varint((1 << 3) | _WT_VARINT) + _varint(reply["version"] & 0xFFFFFFFF))
    for num, key in ((2, "erasure_commitment"), (3, "storage_root"), (4, "encoded_data")):
        v = reply.get(key, b"")
        if len(v):
            parts.append(_len_field(num, v))
    slices = reply.get("encoded_slice", ())
    if getattr(slices, "ndim", 1) == 2:  # (rows, bytes) uint8 array: one shared header
        import numpy as np

        head = np.frombuffer(_varint((5 << 3) | _WT_LEN) + _varint(slices.shape[1]), dtype=np.uint8)
        body = np.concatenate([np.broadcast_to(head, (slices.shape[0], len(head))), slices], axis=1)
        parts.append(body.tobytes())
    else:
        for s in slices:
            parts.append(_len_field(5, s))
    return b"".join(bytes(p) if not isinstance(p, bytes) else p for p in parts)


def decode_reply(buf):
    r = {"version": 0, "erasure_commitment": b"", "storage_root": b"", "encoded_data": b"", "encoded_slice": []}
    for num, wt, v in _fields(buf):
        if num == 1:
            _expect(num, wt, _WT_VARINT)
            r["version"] = v & 0xFFFFFFFF
        elif num in (2, 3, 4):
            _expect(num, wt, _WT_LEN)
            r[("erasure_commitment", "storage_root", "encoded_data")[num - 2]] = bytes(v)
        elif num == 5:
            _expect(num, wt, _WT_LEN)
            r["encoded_slice"].append(bytes(v))
    return r


# ---------------------------------------------------------------------------
# EncodedSlice / LightEncodedSlice
# ---------------------------------------------------------------------------
def _u64(v):
    return int(v).to_bytes(8, "little")


class _Reader:
    def __init__(self, buf):
        self.b = bytes(buf)
        self.i = 0

    def take(self, n):
        if self.i + n > len(self.b):
            raise WireError("truncated slice at byte %d" % self.i)
        v = self.b[self.i:self.i + n]
        self.i += n
        return v

    def u64(self):
        return int.from_bytes(self.take(8), "little")

    def vec(self, elem, limit):
        n = self.u64()
        if n > limit:
            raise WireError("vector length %d exceeds %d" % (n, limit))
        return [self.take(elem) for _ in range(n)]

    def done(self):
        if self.i != len(self.b):
            raise WireError("%d trailing bytes" % (len(self.b) - self.i))


def serialize_encoded_slice(s):
    """EncodedSlice { index, amt: EncodedSliceAMT { index, commitment, row: BlobRow {
    index, row, proof, high_commitment } }, merkle: EncodedSliceMerkle { root, proof,
    leaf_index, leaf } } (encoder/slice.rs:15-20, amt/slice.rs:13-18,
    amt/src/blob/encode.rs:203-209, merkle/slice.rs:12-20) from a dict with keys
    index, amt_index, commitment, row_index, row (list of 32 B or one bytes object),
    amt_proof [(commitment, proof)], high_commitment, merkle_root, merkle_proof,
    leaf_index, leaf; G1 values are 64 B wire form."""
    row = s["row"]
    rows = [row[k:k + 32] for k in range(0, len(row), 32)] if isinstance(row, (bytes, bytearray)) else list(row)
    parts = [_u64(s["index"]), _u64(s["amt_index"]), s["commitment"], _u64(s["row_index"]), _u64(len(rows))]
    parts += rows
    parts.append(_u64(len(s["amt_proof"])))
    for com, prf in s["amt_proof"]:
        parts += [com, prf]
    parts.append(s["high_commitment"])
    parts += list(s["merkle_root"])
    parts += [_u64(len(s["merkle_proof"]))] + list(s["merkle_proof"])
    parts += [_u64(s["leaf_index"]), s["leaf"]]
    return b"".join(bytes(p) for p in parts)


def parse_encoded_slice(buf, coset_n=3, max_row=1 << 20, max_depth=64):
    """Inverse of serialize_encoded_slice (CanonicalDeserialize field order)."""
    r = _Reader(buf)
    s = {"index": r.u64(), "amt_index": r.u64(), "commitment": r.take(64), "row_index": r.u64()}
    s["row"] = r.vec(32, max_row)
    n = r.u64()
    if n > max_depth:
        raise WireError("proof depth %d exceeds %d" % (n, max_depth))
    s["amt_proof"] = [(r.take(64), r.take(64)) for _ in range(n)]
    s["high_commitment"] = r.take(64)
    s["merkle_root"] = [r.take(32) for _ in range(coset_n)]
    s["merkle_proof"] = r.vec(32, max_depth)
    s["leaf_index"] = r.u64()
    s["leaf"] = r.take(32)
    r.done()
    return s


def serialize_light_slice(s):
    """LightEncodedSlice { index, amt_commitment, amt_proof, amt_high_commitment,
    merkle_root, merkle_proof, merkle_leaf } (encoder/light_slice.rs:11-20)."""
    parts = [_u64(s["index"]), s["commitment"], _u64(len(s["amt_proof"]))]
    for com, prf in s["amt_proof"]:
        parts += [com, prf]
    parts.append(s["high_commitment"])
    parts += list(s["merkle_root"])
    parts += [_u64(len(s["merkle_proof"]))] + list(s["merkle_proof"])
    parts.append(s["leaf"])
    return b"".join(bytes(p) for p in parts)


def parse_light_slice(buf, coset_n=3, max_depth=64):
    r = _Reader(buf)
    s = {"index": r.u64(), "commitment": r.take(64)}
    n = r.u64()
    if n > max_depth:
        raise WireError("proof depth %d exceeds %d" % (n, max_depth))
    s["amt_proof"] = [(r.take(64), r.take(64)) for _ in range(n)]
    s["high_commitment"] = r.take(64)
    s["merkle_root"] = [r.take(32) for _ in range(coset_n)]
    s["merkle_proof"] = r.vec(32, max_depth)
    s["leaf"] = r.take(32)
    r.done()
    return s


def light_from_slice(slice_bytes, coset_n=3):
    """LightEncodedSlice::from_slice (encoder/light_slice.rs:59-79): the slice's
    index must equal its AMT and Merkle indices (check_amt_idx / check_merkle_idx,
    encoder/slice.rs:41-70; the reference unwraps, so a mismatch is an error here)."""
    s = parse_encoded_slice(slice_bytes, coset_n)
    if s["index"] != s["amt_index"]:
        raise WireError("UnmatchedAMTIndex { row_index: %d, amt_index: %d }" % (s["index"], s["amt_index"]))
    if s["index"] != s["leaf_index"]:
        raise WireError("UnmatchedMerkleIndex { row_index: %d, merkle_index: %d }" % (s["index"], s["leaf_index"]))
    return serialize_light_slice(s)


def light_into_slice(light_bytes, row, row_n, coset_n=3):
    """LightEncodedSlice::into_slice (encoder/light_slice.rs:22-40): the row data comes
    back in; BlobRow.index = index % row_n, both indices = index."""
    s = parse_light_slice(light_bytes, coset_n)
    s.update(amt_index=s["index"], row_index=s["index"] % row_n, row=row, leaf_index=s["index"])
    return serialize_encoded_slice(s)
