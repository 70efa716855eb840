"""Ceremony import on the CPU: the oracle's restatement of the perpetual-powers-of-tau
point encodings and file layout (oracle/ppot.py), pinned by mathematics, and the host
side of zg_rocm.ppot (crates/ppot2ark/src/main.rs) that needs no device.

The format itself (powersoftau / pairing_ce of kobigurk/phase2-bn254 rev dd6b966) is
parity unpinned: the crates are absent and no ceremony file exists offline.  The pins
below are the generator's encodings worked by hand, round trips over both signs of y,
the curve and subgroup rules, and the product's section offsets against a file built by
concatenation."""
import random

import numpy as np
import pytest

import pairing as E
import ppot as PO
import zg_rocm.ppot as ZP

P, R = PO.P, PO.R


def g1_points(k, seed=5):
    rng = random.Random(seed)
    return [E.g1_mul(E.G1_GEN, rng.randrange(1, R)) for _ in range(k)]


def g2_points(k, seed=6):
    rng = random.Random(seed)
    return [E.g2_mul(E.G2_GEN, rng.randrange(1, R)) for _ in range(k)]


def test_generator_encodings_by_hand():
    # G1 generator (1, 2): y = 2 is the smaller of 2 and p - 2, so no flag
    assert PO.g1_encode(E.G1_GEN, False) == (1).to_bytes(32, "big") + (2).to_bytes(32, "big")
    assert PO.g1_encode(E.G1_GEN, True) == (1).to_bytes(32, "big")
    neg = E.g1_neg(E.G1_GEN)
    assert neg == (1, P - 2)
    assert PO.g1_encode(neg, True) == bytes([0x80]) + (1).to_bytes(32, "big")[1:]
    assert PO.g1_decode(bytes([0x80]) + bytes(30) + b"\x01", True) == (1, P - 2)
    assert PO.g1_decode((1).to_bytes(32, "big"), True) == (1, 2)
    # infinity: 0x40 then zeros, in both widths
    assert PO.g1_encode(None, True) == b"\x40" + bytes(31)
    assert PO.g2_encode(None, False) == b"\x40" + bytes(127)
    # G2 generator: x.c1 first, big-endian
    (x0, x1), (y0, y1) = E.G2_GEN
    enc = PO.g2_encode(E.G2_GEN, False)
    assert enc[:32] == x1.to_bytes(32, "big") and enc[32:64] == x0.to_bytes(32, "big")
    assert enc[64:96] == y1.to_bytes(32, "big") and enc[96:] == y0.to_bytes(32, "big")
    comp = PO.g2_encode(E.G2_GEN, True)
    assert bool(comp[0] & 0x80) == (y1 > (P - 1) // 2)


@pytest.mark.parametrize("compressed", [False, True])
def test_g1_roundtrip_both_signs(compressed):
    for pt in g1_points(12):
        for q in (pt, E.g1_neg(pt)):
            assert PO.g1_decode(PO.g1_encode(q, compressed), compressed) == q


@pytest.mark.parametrize("compressed", [False, True])
def test_g2_roundtrip_both_signs(compressed):
    for q in g2_points(4):
        for s in (q, E.g2_neg(q)):
            assert PO.g2_decode(PO.g2_encode(s, compressed), compressed) == s


def test_f2_sqrt_is_a_root():
    rng = random.Random(9)
    found = 0
    for _ in range(20):
        a = (rng.randrange(P), rng.randrange(P))
        s = PO.f2_sqrt(a)
        if s is not None:
            assert E.f2_sqr(s) == a
            found += 1
        else:
            # a non-square: its norm is a non-residue in Fq
            assert pow((a[0] ** 2 + a[1] ** 2) % P, (P - 1) // 2, P) == P - 1
    assert found > 3
    # c1 = 0: both a residue and a non-residue c0
    assert E.f2_sqr(PO.f2_sqrt((4, 0))) == (4, 0)
    assert E.f2_sqr(PO.f2_sqrt((P - 4, 0))) == (P - 4, 0)


def test_decode_rejections():
    g = PO.g1_encode(E.G1_GEN, False)
    with pytest.raises(PO.DecodeError) as e:
        PO.g1_decode(bytes([g[0] | 0x80]) + g[1:], False)
    assert e.value.status == PO.COMPRESSION_MODE
    with pytest.raises(PO.DecodeError) as e:
        PO.g1_decode(b"\x40" + bytes(63), False)
    assert e.value.status == PO.INFINITY
    with pytest.raises(PO.DecodeError) as e:  # compressed infinity ignores 0x80
        PO.g1_decode(b"\xc0" + bytes(31), True)
    assert e.value.status == PO.INFINITY
    with pytest.raises(PO.DecodeError) as e:
        PO.g1_decode(b"\x40" + bytes(30) + b"\x01", True)
    assert e.value.status == PO.UNEXPECTED_INFO
    with pytest.raises(PO.DecodeError) as e:
        PO.g1_decode(P.to_bytes(32, "big") + (2).to_bytes(32, "big"), False)
    assert e.value.status == PO.COORDINATE
    with pytest.raises(PO.DecodeError) as e:
        PO.g1_decode((1).to_bytes(32, "big") + (3).to_bytes(32, "big"), False)
    assert e.value.status == PO.NOT_ON_CURVE
    # an x with x^3 + 3 a non-residue
    x = next(x for x in range(2, 100) if PO.fq_sqrt((x ** 3 + 3) % P) is None)
    with pytest.raises(PO.DecodeError) as e:
        PO.g1_decode(x.to_bytes(32, "big"), True)
    assert e.value.status == PO.NOT_ON_CURVE


def twist_point_outside_subgroup(seed=3):
    """A point of the twist y^2 = x^3 + 3 / (9 + u) that r does not annihilate (the
    twist has order r (2p - r), so almost every point qualifies)."""
    rng = random.Random(seed)
    while True:
        x = (rng.randrange(P), rng.randrange(P))
        y = PO.f2_sqrt(E.f2_add(E.f2_mul(E.f2_sqr(x), x), E.B2))
        if y is not None:
            q = (x, y)
            assert E.g2_on_curve(q)
            if E.g2_mul(q, R) is not None:
                return q


def test_g2_subgroup_rejection():
    q = twist_point_outside_subgroup()
    for compressed in (False, True):
        with pytest.raises(PO.DecodeError) as e:
            PO.g2_decode(PO.g2_encode(q, compressed), compressed)
        assert e.value.status == PO.NOT_IN_SUBGROUP


@pytest.mark.parametrize("compressed", [False, True])
def test_array_encoders_match_point_encoders(compressed):
    g1 = g1_points(9) + [None]
    g2 = g2_points(3) + [None, E.g2_neg(E.G2_GEN)]
    a1 = np.frombuffer(b"".join(PO.g1_to_canon(p) for p in g1), dtype=np.uint8).reshape(-1, 64)
    a2 = np.frombuffer(b"".join(PO.g2_to_canon(q) for q in g2), dtype=np.uint8).reshape(-1, 128)
    assert PO.encode_g1_array(a1, compressed) == b"".join(PO.g1_encode(p, compressed) for p in g1)
    assert PO.encode_g2_array(a2, compressed) == b"".join(PO.g2_encode(q, compressed) for q in g2)


@pytest.mark.parametrize("compressed", [False, True])
def test_section_offsets_against_concatenated_file(compressed):
    """zg_rocm.ppot's offsets (calculate_mmap_position) index the oracle's file, which is
    built by concatenating the sections: every section's first and last points."""
    power = 2
    n = 1 << power
    pts1 = g1_points(2 * n - 1 + 2 * n, seed=11)
    pts2 = g2_points(n + 1, seed=12)
    c1 = lambda ps: np.frombuffer(b"".join(PO.g1_to_canon(p) for p in ps), dtype=np.uint8).reshape(-1, 64)
    c2 = lambda qs: np.frombuffer(b"".join(PO.g2_to_canon(q) for q in qs), dtype=np.uint8).reshape(-1, 128)
    tau_g1, alpha, beta = pts1[:2 * n - 1], pts1[2 * n - 1:3 * n - 1], pts1[3 * n - 1:]
    tau_g2, beta_g2 = pts2[:n], pts2[n:]
    blob = PO.accumulator_bytes(power, c1(tau_g1), c2(tau_g2), c1(alpha), c1(beta), c2(beta_g2), compressed,
                                prev_hash=bytes(range(64)), public_key=b"\x07" * 100)
    params = ZP.CeremonyParams(power, power)
    assert params.file_length(compressed) == len(blob) - 100
    sections = [(ZP.ElementType.TauG1, tau_g1, False), (ZP.ElementType.TauG2, tau_g2, True),
                (ZP.ElementType.AlphaG1, alpha, False), (ZP.ElementType.BetaG1, beta, False),
                (ZP.ElementType.BetaG2, beta_g2, True)]
    for el, pts, g2 in sections:
        size = params.point_size(el, compressed)
        assert params.section_length(el) == len(pts)
        for i in (0, len(pts) - 1):
            pos = params.position(i, el, compressed)
            dec = (PO.g2_decode if g2 else PO.g1_decode)(blob[pos:pos + size], compressed)
            assert dec == pts[i], (el, i)


def test_file_names():
    assert ZP.InputType.Challenge.file_name(12) == "challenge_12"
    assert ZP.InputType.Response.file_name(28) == "response_28"
    # crates/amt/src/utils.rs:31-34 with the Bn254 type hash of README.md:73-82
    assert ZP.ptau_file_name(20) == "power-tau-5DWgDV-20.bin"
    assert ZP.ptau_file_name(20, True) == "power-tau-mont-5DWgDV-20.bin"


class _NoDevice:
    """Stands in for a context: the checks below come before any device call."""

    def ppot_decode(self, *a):
        raise AssertionError("no device call expected")


def test_too_long_to_read(tmp_path):
    # the reference's test_load_from_challenge_12_too_long (main.rs:333-353)
    with pytest.raises(ZP.PpotError, match="^too long to read$"):
        ZP.from_ppot_file(_NoDevice(), str(tmp_path), ZP.InputType.Challenge, 12, 3841, 8, 10)


def test_missing_file(tmp_path):
    with pytest.raises(ZP.PpotError, match="Cannot open .*challenge_12"):
        ZP.from_ppot_file(_NoDevice(), str(tmp_path), ZP.InputType.Challenge, 12, 0, 8, 10)


def test_short_file(tmp_path):
    (tmp_path / "challenge_4").write_bytes(bytes(64 + 100))
    with pytest.raises(ZP.PpotError, match="must read point data from file"):
        ZP.from_ppot_file(_NoDevice(), str(tmp_path), ZP.InputType.Challenge, 4, 0, 2, 2)
