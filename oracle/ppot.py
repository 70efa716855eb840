"""CPU restatement of the perpetual-powers-of-tau file format (TEST INFRASTRUCTURE: only
tests/ import this module; it writes synthetic ceremony files and checks the device
decoder, and is never the thing measured or shipped).

crates/ppot2ark/src/main.rs:45-119 reads ceremony files through the powersoftau crate of
kobigurk/phase2-bn254 (rev dd6b966), whose point encodings are pairing_ce's bn256
EncodedPoint (G1Uncompressed / G1Compressed / G2Uncompressed / G2Compressed):
  * coordinates as canonical integers, 32 bytes big-endian each; G2 coordinates as
    c1 then c0;
  * uncompressed G1 = x || y, compressed G1 = x; uncompressed G2 = x || y, compressed
    G2 = x;
  * first byte: 0x40 = point at infinity (all other bits zero); in a compressed point
    0x80 = "y is the larger of y and -y" (canonical integers; Fq2 compared c1 first);
    0x80 on an uncompressed point is an error (wrong compression mode);
  * a checked decode (into_affine) rejects coordinates >= p, points off the curve and G2
    points outside the order-r subgroup; BatchedAccumulator::read_points_chunk with
    CheckForCorrectness::Yes also rejects the point at infinity.
A file: 64-byte hash, then tau_powers_g1 (2 * 2^power - 1 points), tau_powers_g2,
alpha_tau_powers_g1, beta_tau_powers_g1 (2^power each) and beta_g2 (one point), all
uncompressed in a challenge and compressed in a response (which then carries the
contributor's public key).

Parity unpinned: the crates are absent and no ceremony file or fixture exists offline.
What pins this restatement is mathematics: the generator's encodings worked by hand,
round trips over both signs of y, the curve and subgroup rules (pairing.py), and the
layout re-derived by concatenation rather than from offsets.
"""
import numpy as np

import pairing as E

P = E.P
R = E.R
HALF = (P - 1) // 2

OK, COMPRESSION_MODE, UNEXPECTED_INFO, INFINITY, COORDINATE, NOT_ON_CURVE, NOT_IN_SUBGROUP = range(7)


class DecodeError(ValueError):
    def __init__(self, status):
        super().__init__("ppot decode status %d" % status)
        self.status = status


def _be(x):
    return x.to_bytes(32, "big")


def _f2_larger(y):
    """y > -y with Fq2 ordered by c1, then c0 (canonical integers)."""
    c0, c1 = y
    if c1 != 0:
        return c1 > HALF
    return c0 > HALF


def g1_encode(pt, compressed):
    size = 32 if compressed else 64
    if pt is None:
        return bytes([0x40]) + bytes(size - 1)
    x, y = pt
    if compressed:
        b = bytearray(_be(x))
        if y > HALF:
            b[0] |= 0x80
        return bytes(b)
    return _be(x) + _be(y)


def g2_encode(q, compressed):
    size = 64 if compressed else 128
    if q is None:
        return bytes([0x40]) + bytes(size - 1)
    (x0, x1), (y0, y1) = q
    if compressed:
        b = bytearray(_be(x1) + _be(x0))
        if _f2_larger((y0, y1)):
            b[0] |= 0x80
        return bytes(b)
    return _be(x1) + _be(x0) + _be(y1) + _be(y0)


def _flags(b, compressed):
    if not compressed and b[0] & 0x80:
        raise DecodeError(COMPRESSION_MODE)
    if b[0] & 0x40:
        rest = bytes([b[0] & 0x3f]) + bytes(b[1:])
        raise DecodeError(UNEXPECTED_INFO if any(rest) else INFINITY)
    return bool(b[0] & 0x80)


def _coord(b, first):
    v = int.from_bytes(bytes([b[0] & 0x3f]) + bytes(b[1:32]) if first else bytes(b[:32]), "big")
    if v >= P:
        raise DecodeError(COORDINATE)
    return v


def fq_sqrt(a):
    s = pow(a, (P + 1) // 4, P)
    return s if s * s % P == a % P else None


def f2_sqrt(a):
    """A square root in Fq2 (any of the two), or None."""
    a0, a1 = a
    if a1 == 0:
        s = fq_sqrt(a0)
        if s is not None:
            return (s, 0)
        s = fq_sqrt(-a0 % P)
        return None if s is None else (0, s)
    t = fq_sqrt((a0 * a0 + a1 * a1) % P)
    if t is None:
        return None
    inv2 = pow(2, P - 2, P)
    for d in ((a0 + t) * inv2 % P, (a0 - t) * inv2 % P):
        x0 = fq_sqrt(d)
        if x0 is not None and x0 != 0:
            x = (x0, a1 * pow(2 * x0, P - 2, P) % P)
            if E.f2_sqr(x) == (a0 % P, a1 % P):
                return x
    return None


def g1_decode(b, compressed):
    """pairing_ce G1 into_affine, then the accumulator's infinity check."""
    b = bytes(b)
    greatest = _flags(b, compressed)
    x = _coord(b, True)
    rhs = (x * x * x + 3) % P
    if compressed:
        y = fq_sqrt(rhs)
        if y is None:
            raise DecodeError(NOT_ON_CURVE)
        if (y > HALF) != greatest:
            y = (P - y) % P
    else:
        y = _coord(b[32:], False)
        if y * y % P != rhs:
            raise DecodeError(NOT_ON_CURVE)
    return (x, y)


def g2_decode(b, compressed):
    b = bytes(b)
    greatest = _flags(b, compressed)
    x1 = _coord(b, True)
    x0 = _coord(b[32:], False)
    x = (x0, x1)
    rhs = E.f2_add(E.f2_mul(E.f2_sqr(x), x), E.B2)
    if compressed:
        y = f2_sqrt(rhs)
        if y is None:
            raise DecodeError(NOT_ON_CURVE)
        if _f2_larger(y) != greatest:
            y = E.f2_neg(y)
    else:
        y = (_coord(b[96:], False), _coord(b[64:], False))
        if E.f2_sqr(y) != rhs:
            raise DecodeError(NOT_ON_CURVE)
    q = (x, y)
    if E.g2_mul(q, R) is not None:
        raise DecodeError(NOT_IN_SUBGROUP)
    return q


# ---- arrays in the library's memory form (canonical affine, little-endian) ----
_HALF_LIMBS = np.array([(HALF >> (64 * k)) & (2**64 - 1) for k in range(4)], dtype=np.uint64)


def _gt_half(le):
    """(n, 32) little-endian canonical integers -> bool, value > (p - 1) / 2."""
    w = np.ascontiguousarray(le).view("<u8")
    gt = np.zeros(len(w), dtype=bool)
    eq = np.ones(len(w), dtype=bool)
    for k in (3, 2, 1, 0):
        gt |= eq & (w[:, k] > _HALF_LIMBS[k])
        eq &= w[:, k] == _HALF_LIMBS[k]
    return gt


def encode_g1_array(pts, compressed):
    """(n, 64) canonical affine G1 (all-zero rows = infinity) -> n encodings, bytes."""
    pts = np.ascontiguousarray(pts, dtype=np.uint8).reshape(-1, 64)
    x = pts[:, 0:32][:, ::-1]
    y = pts[:, 32:64][:, ::-1]
    inf = ~pts.any(axis=1)
    if compressed:
        out = x.copy()
        out[:, 0] |= np.where(_gt_half(pts[:, 32:64]), 0x80, 0).astype(np.uint8)
    else:
        out = np.concatenate([x, y], axis=1)
    out[inf] = 0
    out[inf, 0] = 0x40
    return out.tobytes()


def encode_g2_array(pts, compressed):
    """(n, 128) canonical affine G2 (x.c0, x.c1, y.c0, y.c1) -> n encodings, bytes."""
    pts = np.ascontiguousarray(pts, dtype=np.uint8).reshape(-1, 128)
    be = lambda k: pts[:, 32 * k:32 * k + 32][:, ::-1]
    inf = ~pts.any(axis=1)
    if compressed:
        out = np.concatenate([be(1), be(0)], axis=1)
        c1_zero = ~pts[:, 96:128].any(axis=1)
        larger = np.where(c1_zero, _gt_half(pts[:, 64:96]), _gt_half(pts[:, 96:128]))
        out[:, 0] |= np.where(larger, 0x80, 0).astype(np.uint8)
    else:
        out = np.concatenate([be(1), be(0), be(3), be(2)], axis=1)
    out[inf] = 0
    out[inf, 0] = 0x40
    return out.tobytes()


def accumulator_bytes(power, tau_g1, tau_g2, alpha_g1, beta_g1, beta_g2, compressed, prev_hash=None,
                      public_key=b""):
    """A whole challenge (compressed = False) or response file from canonical arrays:
    tau_g1 (2 * 2^power - 1, 64), tau_g2 / alpha_g1 / beta_g1 (2^power, ...), beta_g2 (1, 128)."""
    n = 1 << power
    assert len(tau_g1) == 2 * n - 1 and len(tau_g2) == n and len(alpha_g1) == n and len(beta_g1) == n
    assert len(beta_g2) == 1
    h = bytes(64) if prev_hash is None else prev_hash
    assert len(h) == 64
    return b"".join([h, encode_g1_array(tau_g1, compressed), encode_g2_array(tau_g2, compressed),
                     encode_g1_array(alpha_g1, compressed), encode_g1_array(beta_g1, compressed),
                     encode_g2_array(beta_g2, compressed), public_key])


def g1_to_canon(pt):
    if pt is None:
        return bytes(64)
    return pt[0].to_bytes(32, "little") + pt[1].to_bytes(32, "little")


def g2_to_canon(q):
    if q is None:
        return bytes(128)
    (x0, x1), (y0, y1) = q
    return b"".join(v.to_bytes(32, "little") for v in (x0, x1, y0, y1))
