"""Ceremony import on the device (crates/ppot2ark/src/main.rs): zgr_ppot_decode and the
zg_rocm.ppot mirror of from_ppot_file / from_ppot_file_ldt / load_save_power_tau, over
synthetic challenge and response files of 2^12 powers (the reference tests' size,
main.rs:285-381) written by the oracle's restatement of the format (oracle/ppot.py).

The powers in the files are the device's own [tau^i] G1 / G2 for a seeded tau (pinned
against the oracle and textbook multiples in test_gpu_params_io.py); what is checked
here is the decoding: every point read back equals the point written, each rejection
of a checked read is reported with its index, and the pairing identities of the
reference tests hold.  The file format itself is parity unpinned (oracle/ppot.py)."""
import numpy as np
import pytest

import pairing as E
import ppot as PO
import zg_rocm as Z
import zg_rocm.ppot as ZP
from helpers import seeded_tau

pytestmark = pytest.mark.gpu

FILE_POW = 12
N = 1 << FILE_POW


@pytest.fixture(scope="module")
def powers():
    """[tau^i] G1 for i < 2^13 and [tau^i] G2 for i < 2^13 (a 2^13-point context)."""
    c = Z.Context(Z.Geometry(7, 6))
    c.power_tau_setup(seeded_tau())
    g1, g2, _, _ = c.power_tau_download()
    c.close()
    return g1, g2


@pytest.fixture(scope="module")
def files(powers, tmp_path_factory):
    g1, g2 = powers
    d = tmp_path_factory.mktemp("ppot")
    rng = np.random.default_rng(84)
    sections = (g1[:2 * N - 1], g2[:N], g1[1:N + 1], g1[2:N + 2], g2[1:2])
    (d / ("challenge_%d" % FILE_POW)).write_bytes(
        PO.accumulator_bytes(FILE_POW, *sections, compressed=False, prev_hash=rng.bytes(64)))
    (d / ("response_%d" % FILE_POW)).write_bytes(
        PO.accumulator_bytes(FILE_POW, *sections, compressed=True, prev_hash=rng.bytes(64),
                             public_key=rng.bytes(6 * 64 + 3 * 128)))
    return d


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(Z.TEST_SETTING)
    yield c
    c.close()


def g1_of(row):
    b = bytes(row)
    return int.from_bytes(b[:32], "little"), int.from_bytes(b[32:], "little")


def g2_of(row):
    b = bytes(row)
    v = [int.from_bytes(b[32 * k:32 * k + 32], "little") for k in range(4)]
    return (v[0], v[1]), (v[2], v[3])


@pytest.mark.parametrize("kind", [Z.PPOT_G1_UNCOMPRESSED, Z.PPOT_G1_COMPRESSED, Z.PPOT_G2_UNCOMPRESSED,
                                  Z.PPOT_G2_COMPRESSED])
def test_decode_matches_written_points(ctx, powers, kind):
    g1, g2 = powers
    g2k = kind >= Z.PPOT_G2_UNCOMPRESSED
    comp = kind % 2 == 1
    pts = g2[:700] if g2k else g1[:3000]
    # both signs of y: negate every third point through the oracle
    pts = pts.copy()
    for i in range(0, len(pts), 3):
        pts[i] = np.frombuffer(PO.g2_to_canon(E.g2_neg(g2_of(pts[i]))) if g2k
                               else PO.g1_to_canon(E.g1_neg(g1_of(pts[i]))), dtype=np.uint8)
    enc = (PO.encode_g2_array if g2k else PO.encode_g1_array)(pts, comp)
    out = ctx.ppot_decode(kind, enc, len(pts))
    assert np.array_equal(out, pts)
    # the oracle decoder agrees on a sample
    per = len(enc) // len(pts)
    for i in (0, 1, 2, len(pts) - 1):
        dec = (PO.g2_decode if g2k else PO.g1_decode)(enc[i * per:(i + 1) * per], comp)
        assert (PO.g2_to_canon(dec) if g2k else PO.g1_to_canon(dec)) == bytes(out[i])


def planted(enc, per, at, new):
    b = bytearray(enc)
    b[at * per:(at + 1) * per] = new
    return bytes(b)


def expect_reject(ctx, kind, data, count, at, status):
    with pytest.raises(Z.ZgrError) as e:
        ctx.ppot_decode(kind, data, count)
    assert e.value.first_bad == at
    reasons = {PO.COMPRESSION_MODE: "compression mode", PO.UNEXPECTED_INFO: "unexpected information",
               PO.INFINITY: "point at infinity", PO.COORDINATE: "coordinate", PO.NOT_ON_CURVE: "not on the curve",
               PO.NOT_IN_SUBGROUP: "subgroup"}
    assert reasons[status] in e.value.msg


def test_rejections_g1(ctx, powers):
    g1, _ = powers
    pts = g1[:64]
    unc = PO.encode_g1_array(pts, False)
    comp = PO.encode_g1_array(pts, True)
    k_u, k_c = Z.PPOT_G1_UNCOMPRESSED, Z.PPOT_G1_COMPRESSED
    first = bytearray(unc[5 * 64:6 * 64])
    first[0] |= 0x80
    expect_reject(ctx, k_u, planted(unc, 64, 5, bytes(first)), 64, 5, PO.COMPRESSION_MODE)
    expect_reject(ctx, k_u, planted(unc, 64, 9, b"\x40" + bytes(63)), 64, 9, PO.INFINITY)
    expect_reject(ctx, k_c, planted(comp, 32, 10, b"\xc0" + bytes(31)), 64, 10, PO.INFINITY)
    expect_reject(ctx, k_c, planted(comp, 32, 11, b"\x40" + bytes(30) + b"\x02"), 64, 11, PO.UNEXPECTED_INFO)
    expect_reject(ctx, k_u, planted(unc, 64, 12, PO.P.to_bytes(32, "big") + (2).to_bytes(32, "big")), 64, 12,
                  PO.COORDINATE)
    expect_reject(ctx, k_u, planted(unc, 64, 13, (1).to_bytes(32, "big") + (3).to_bytes(32, "big")), 64, 13,
                  PO.NOT_ON_CURVE)
    x = next(x for x in range(2, 100) if PO.fq_sqrt((x ** 3 + 3) % PO.P) is None)
    expect_reject(ctx, k_c, planted(comp, 32, 63, x.to_bytes(32, "big")), 64, 63, PO.NOT_ON_CURVE)
    # the first rejected point is reported
    two = planted(planted(unc, 64, 40, b"\x40" + bytes(63)), 64, 20, (1).to_bytes(32, "big") * 2)
    expect_reject(ctx, k_u, two, 64, 20, PO.NOT_ON_CURVE)


def test_rejections_g2(ctx, powers):
    from test_ppot import twist_point_outside_subgroup
    _, g2 = powers
    pts = g2[:16]
    q = twist_point_outside_subgroup()
    for comp, kind in ((False, Z.PPOT_G2_UNCOMPRESSED), (True, Z.PPOT_G2_COMPRESSED)):
        per = 64 if comp else 128
        enc = PO.encode_g2_array(pts, comp)
        expect_reject(ctx, kind, planted(enc, per, 7, PO.g2_encode(q, comp)), 16, 7, PO.NOT_IN_SUBGROUP)
        expect_reject(ctx, kind, planted(enc, per, 3, b"\x40" + bytes(per - 1)), 16, 3, PO.INFINITY)
        bad = bytearray(enc[2 * per:3 * per])
        bad[32:64] = PO.P.to_bytes(32, "big")  # x.c0 = p
        expect_reject(ctx, kind, planted(enc, per, 2, bytes(bad)), 16, 2, PO.COORDINATE)
    unc = PO.encode_g2_array(pts, False)
    off = bytearray(unc[4 * 128:5 * 128])
    off[127] ^= 1  # y.c0 + 1: off the curve
    expect_reject(ctx, Z.PPOT_G2_UNCOMPRESSED, planted(unc, 128, 4, bytes(off)), 16, 4, PO.NOT_ON_CURVE)


def test_from_ppot_file_challenge(ctx, powers, files):
    """test_load_from_challenge_12_nomal (main.rs:307-331)."""
    g1, g2 = powers
    pot = ZP.from_ppot_file(ctx, str(files), ZP.InputType.Challenge, FILE_POW, 3840, 8, 10)
    assert len(pot.g1) == 256 and len(pot.g2) == 256
    assert np.array_equal(pot.g1, g1[3840:4096]) and np.array_equal(pot.g2, g2[3840:4096])
    assert E.pairing(g1_of(pot.g1[0]), g2_of(pot.g2[4])) == E.pairing(g1_of(pot.g1[1]), g2_of(pot.g2[3]))


def test_from_ppot_file_too_long(ctx, files):
    with pytest.raises(ZP.PpotError, match="^too long to read$"):
        ZP.from_ppot_file(ctx, str(files), ZP.InputType.Challenge, FILE_POW, 3841, 8, 10)


def test_from_ppot_file_high_response(ctx, powers, files):
    """test_load_from_high_deg_response_nomal (main.rs:355-381): compressed points."""
    g1, g2 = powers
    start = N - 256
    pot = ZP.from_ppot_file(ctx, str(files), ZP.InputType.Response, FILE_POW, start, 8, 6)
    assert np.array_equal(pot.g1, g1[start:N]) and np.array_equal(pot.g2, g2[start:N])


def test_from_ppot_file_ldt(ctx, powers, files):
    """test_from_ppot_file_ldt (main.rs:285-305) with PowerTau::check_ldt
    (crates/amt/src/power_tau.rs:210-225): e(g1pp[i], high_g2) = e(high_g1pp[i], g2pp[0]),
    here through tau: high_g1pp[i] = [tau^H] g1pp[i], H = 2^11 - 2^8."""
    g1, g2 = powers
    pt = ZP.from_ppot_file_ldt(ctx, str(files), ZP.InputType.Challenge, FILE_POW, 8, 11, 10)
    H = (1 << 11) - (1 << 8)
    assert len(pt.g1pp) == len(pt.g2pp) == len(pt.high_g1pp) == 256
    assert np.array_equal(pt.g1pp, g1[:256]) and np.array_equal(pt.g2pp, g2[:256])
    assert np.array_equal(pt.high_g1pp, g1[H:H + 256]) and np.array_equal(pt.high_g2, g2[H])
    assert E.pairing(g1_of(pt.g1pp[1]), g2_of(pt.high_g2)) == E.pairing(g1_of(pt.high_g1pp[1]), g2_of(pt.g2pp[0]))


def test_corrupt_section_fails_the_read(ctx, files, tmp_path):
    """A checked chunk read covers the alpha and beta sections too."""
    blob = bytearray((files / ("challenge_%d" % FILE_POW)).read_bytes())
    params = ZP.CeremonyParams(FILE_POW, FILE_POW)
    pos = params.position(17, ZP.ElementType.AlphaG1, False)
    blob[pos + 63] ^= 1
    (tmp_path / ("challenge_%d" % FILE_POW)).write_bytes(bytes(blob))
    with pytest.raises(ZP.PpotError, match=r"AlphaG1\[17\]: .*not on the curve"):
        ZP.from_ppot_file(ctx, str(tmp_path), ZP.InputType.Challenge, FILE_POW, 0, 5, 5)
    ZP.from_ppot_file(ctx, str(tmp_path), ZP.InputType.Challenge, FILE_POW, 32, 5, 5)  # range clear of it


def test_load_save_power_tau(ctx, powers, files, tmp_path):
    """load_save_power_tau (main.rs:166-191) into a 2^11 context: the powers stay loaded
    and the compressed PowerTau file is written under its reference name."""
    g1, g2 = powers
    path = ZP.load_save_power_tau(ctx, str(files), ZP.InputType.Response, FILE_POW, 11, 12, 9, str(tmp_path))
    assert path.endswith("power-tau-5DWgDV-11.bin")
    a, b, hi, hg2 = ctx.power_tau_download()
    H = (1 << 12) - (1 << 11)
    assert np.array_equal(a, g1[:2048]) and np.array_equal(b, g2[:2048])
    assert np.array_equal(hi, g1[H:H + 2048]) and np.array_equal(hg2, g2[H])
    with open(path, "rb") as f:
        assert f.read() == ctx.write_params_file(Z.FILE_PTAU)
    ctx.params_from_power_tau(0)  # the imported powers feed from_pp
