"""The public-parameter pipeline on the device at the reference's test geometry:
powers of tau in G1 and G2 (PowerTau::setup_inner, crates/amt/src/power_tau.rs:54-80),
the verifier's G2 data of every coset (gen_vanishes, crates/amt/src/prove_params/
generate.rs:234-263; g2, high_g2), and the parameter files of bin/build_params.rs:24-53.

Pins: G1 powers against the oracle's power_tau; G2 powers, g2 and high_g2 against
textbook G2 scalar multiples (oracle/pairing.py); vanishes against the closed form of
oracle/verify.py (itself pinned against the reference's inverse DFT); the bamt image
against the reference writer's layout (write_amt_params) and read back by the loader
byte for byte.  The ark-serialize 0.4 encodings inside the files (u64 lengths, point
flags) cannot be imported here and are parity unpinned; the independent readers below
check their self-consistency: every coordinate and flag against the device's points."""
import numpy as np
import pytest

import oracle as O
import pairing as E
import verify as V
import zg_rocm as Z
from helpers import DATA_SEED, bamt_image, seeded_tau

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING
R, P = O.R_MOD, Z.P_MOD


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    c.params_generate(seeded_tau())
    yield c
    c.close()


def g2_from_canon(b):
    """canonical affine G2 bytes -> pairing.py tuple (None for infinity)"""
    b = bytes(b)
    if not any(b):
        return None
    v = [int.from_bytes(b[32 * k:32 * k + 32], "little") for k in range(4)]
    return ((v[0], v[1]), (v[2], v[3]))


def g1_from_canon(b):
    b = bytes(b)
    if not any(b):
        return None
    return int.from_bytes(b[:32], "little"), int.from_bytes(b[32:64], "little")


def test_power_tau_matches_oracle(ctx):
    tau = seeded_tau()
    ctx.power_tau_setup(tau)
    g1, g2, hi, hg2 = ctx.power_tau_download()
    assert np.array_equal(g1.reshape(-1), O.power_tau(tau, 0, GEO.log_n))
    assert np.array_equal(hi.reshape(-1), O.power_tau(tau, 3 * GEO.n, GEO.log_n))
    for i in (0, 1, 2, 3, 100, GEO.n - 1):
        assert g2_from_canon(g2[i]) == E.g2_mul(E.G2_GEN, pow(tau, i, R)), i
    assert g2_from_canon(hg2) == E.g2_mul(E.G2_GEN, pow(tau, 3 * GEO.n, R))
    assert g2_from_canon(g2[0]) == E.G2_GEN


def test_vanishes_match_closed_form(ctx):
    """Every vanish point of every coset, from the closed-form generator and from the
    G2 DFT over the powers (the reference's gen_vanishes), against oracle/verify.py."""
    tau = seeded_tau()
    ctx.params_generate(tau)
    closed = [ctx.params_g2(c) for c in range(GEO.coset_n)]
    bases = [ctx.params_download(c) for c in range(GEO.coset_n)]
    ctx.power_tau_setup(tau)
    for c in range(GEO.coset_n):
        ctx.params_from_power_tau(c)
        van, g2, hg2 = ctx.params_g2(c)
        assert np.array_equal(ctx.params_download(c), bases[c]), c
        assert np.array_equal(van, closed[c][0]), c
        assert np.array_equal(g2, closed[c][1]) and np.array_equal(hg2, closed[c][2]), c
        assert g2_from_canon(g2) == E.G2_GEN
        assert g2_from_canon(hg2) == E.g2_mul(E.G2_GEN, pow(tau, 3 * GEO.n, R))
        vp = V.VerifyParams(tau, GEO.log_col, GEO.log_row, c, [])
        for d in range(GEO.h):
            for k in range(2 << d):
                assert g2_from_canon(van[(2 << d) - 2 + k]) == vp.vanish(d, k), (c, d, k)


def test_bamt_writer_round_trip(ctx):
    """write_amt_params from the device: the G1 sections in the reference's order, and an
    image that the loader reads back into the same bases and verifier data, written out
    again byte for byte."""
    ctx.params_generate(seeded_tau())
    other = Z.Context(GEO)
    try:
        for c in range(GEO.coset_n):
            img = ctx.write_params_file(Z.FILE_AMT_MONT, c)
            bases = ctx.params_download(c)
            ref = bamt_image(bases, GEO.log_n, GEO.h)  # zero G2 sections
            assert len(img) == len(ref)
            assert img[:6] == b"bamt" + bytes([GEO.log_n, GEO.h])
            g1 = 6 + 256
            line = GEO.n * 64
            van = ((2 << GEO.h) - 2) * 128
            assert img[g1:g1 + (GEO.h + 1) * line] == ref[g1:g1 + (GEO.h + 1) * line]
            assert img[-line:] == ref[-line:]
            other.params_load_bamt(c, img)
            assert np.array_equal(other.params_download(c), bases)
            for a, b in zip(other.params_g2(c), ctx.params_g2(c)):
                assert np.array_equal(a, b)
            assert other.write_params_file(Z.FILE_AMT_MONT, c) == img
            # the vanishes section holds the Montgomery form of the device's points
            sec = img[g1 + (GEO.h + 1) * line:g1 + (GEO.h + 1) * line + van]
            v, _, _ = ctx.params_g2(c)
            for k in (0, 1, 5, len(v) - 1):
                for q in range(4):
                    x = int.from_bytes(bytes(v[k][32 * q:32 * q + 32]), "little")
                    assert int.from_bytes(sec[128 * k + 32 * q:128 * k + 32 * q + 32], "little") == x * 2**256 % P
    finally:
        other.close()


class _Reader:
    def __init__(self, b):
        self.b, self.p = b, 0

    def take(self, n):
        v = self.b[self.p:self.p + n]
        assert len(v) == n
        self.p += n
        return v

    def u64(self):
        return int.from_bytes(self.take(8), "little")

    def done(self):
        assert self.p == len(self.b)


def _flag_ok(last_byte, y_neg, inf):
    return (last_byte & 0xC0) == (0x40 if inf else (0x80 if y_neg else 0))


def read_g1(r, compressed=False):
    """ark G1 -> (canonical x, y or None) with the flag checked against y"""
    raw = bytearray(r.take(32 if compressed else 64))
    flags = raw[-1]
    raw[-1] &= 0x3F
    x = int.from_bytes(bytes(raw[:32]), "little")
    if compressed:
        return x, flags
    y = int.from_bytes(bytes(raw[32:]), "little")
    inf = x == 0 and y == 0
    assert _flag_ok(flags, y > (P - 1) // 2, inf)
    return None if inf else (x, y)


def read_g2(r, compressed=False):
    raw = bytearray(r.take(64 if compressed else 128))
    flags = raw[-1]
    raw[-1] &= 0x3F
    v = [int.from_bytes(bytes(raw[32 * k:32 * k + 32]), "little") for k in range(len(raw) // 32)]
    if compressed:
        return (v[0], v[1]), flags
    inf = not any(v)
    yneg = v[3] > (P - 1) // 2 if v[3] else v[2] > (P - 1) // 2
    assert _flag_ok(flags, yneg, inf)
    return None if inf else ((v[0], v[1]), (v[2], v[3]))


def test_ark_files_parse_back(ctx):
    """AMTParams (unmont), AMTVerifyParams and the compressed / mont PowerTau files:
    every section read back by an independent reader equals the device's points."""
    tau = seeded_tau()
    ctx.params_generate(tau)
    n, h = GEO.n, GEO.h
    for c in range(GEO.coset_n):
        lines = O.fr_canonical_array(ctx.params_download(c).reshape(-1, 4), which=1).view(np.uint8).reshape(h + 2, n, 64)
        van, g2, hg2 = ctx.params_g2(c)
        r = _Reader(ctx.write_params_file(Z.FILE_AMT, c))
        assert r.u64() == n and all(read_g1(r) == g1_from_canon(lines[h][i]) for i in range(n))
        assert r.u64() == h
        for d in range(h):
            assert r.u64() == n and all(read_g1(r) == g1_from_canon(lines[d][i]) for i in range(n))
        assert r.u64() == h
        for d in range(h):
            assert r.u64() == 2 << d
            assert all(read_g2(r) == g2_from_canon(van[(2 << d) - 2 + k]) for k in range(2 << d))
        assert read_g2(r) == g2_from_canon(g2)
        assert r.u64() == n and all(read_g1(r) == g1_from_canon(lines[h + 1][i]) for i in range(n))
        assert read_g2(r) == g2_from_canon(hg2)
        r.done()
        for vd in (h, 3):
            r = _Reader(ctx.write_params_file(Z.FILE_AMT_VERIFY, c, vd))
            assert r.u64() == n and all(read_g1(r) == g1_from_canon(lines[h][i]) for i in range(n))
            assert r.u64() == vd
            for d in range(vd):
                assert r.u64() == 2 << d
                assert all(read_g2(r) == g2_from_canon(van[(2 << d) - 2 + k]) for k in range(2 << d))
            assert read_g2(r) == g2_from_canon(g2) and read_g2(r) == g2_from_canon(hg2)
            r.done()
    ctx.power_tau_setup(tau)
    g1, g2pp, hi, hg2 = ctx.power_tau_download()
    r = _Reader(ctx.write_params_file(Z.FILE_PTAU))
    for pts, g in ((g1, 1), (g2pp, 2), (hi, 1), (hg2.reshape(1, 128), 2)):
        if pts.shape[0] > 1:
            assert r.u64() == len(pts)
        for p in pts:
            if g == 1:
                x, flags = read_g1(r, compressed=True)
                xy = g1_from_canon(p)
                assert x == xy[0] and _flag_ok(flags, xy[1] > (P - 1) // 2, False)
            else:
                x, flags = read_g2(r, compressed=True)
                q = g2_from_canon(p)
                y = q[1]
                assert x == q[0] and _flag_ok(flags, (y[1] if y[1] else y[0]) > (P - 1) // 2, False)
    r.done()
    m = ctx.write_params_file(Z.FILE_PTAU_MONT)
    assert m[:5] == b"ptau" + bytes([GEO.log_n])
    r = _Reader(m[5:])
    assert read_g2(r) == g2_from_canon(hg2)
    for pts, w in ((g1, 64), (g2pp, 128), (hi, 64)):
        for p in pts:
            got = r.take(w)
            for q in range(w // 32):
                x = int.from_bytes(bytes(p[32 * q:32 * q + 32]), "little")
                assert int.from_bytes(got[32 * q:32 * q + 32], "little") == x * 2**256 % P
    r.done()


def test_written_verify_params_verify_slices(ctx):
    """A device-encoded blob's slices pass the restated EncodedSlice::verify (pairings
    included) with the verifier params read back from the written verify file."""
    from zg_rocm import wire

    tau = seeded_tau()
    ctx.params_generate(tau)
    out = ctx.encode(Z.seeded_bytes(DATA_SEED + 900, 30000))
    blob = Z.EncodedBlob(GEO, out)
    vps = []
    for c in range(GEO.coset_n):
        r = _Reader(ctx.write_params_file(Z.FILE_AMT_VERIFY, c))
        assert r.u64() == GEO.n
        basis = [read_g1(r) for _ in range(GEO.n)]
        assert r.u64() == GEO.h
        van = []
        for d in range(GEO.h):
            assert r.u64() == 2 << d
            van.append([read_g2(r) for _ in range(2 << d)])
        vp = V.VerifyParams(tau, GEO.log_col, GEO.log_row, c, basis)
        vp.g2, vp.high_g2 = read_g2(r), read_g2(r)
        vp._van = {(d, k): van[d][k] for d in range(GEO.h) for k in range(2 << d)}  # the file's points only
        r.done()
        vps.append(vp)
    com = g1_from_canon(out.commitment)
    root = bytes(out.file_root)
    for i in (5, GEO.row_n + 17, 2 * GEO.row_n + 63):
        s = wire.parse_encoded_slice(blob.encoded_slice(i))
        assert V.verify_slice(s, vps, com, root, GEO) is None, i


def test_build_params_files(ctx, tmp_path):
    """build_params in create mode: the reference's file names, one set per coset, and
    a mont file that loads into a fresh context with the same bases as the builder's."""
    paths = Z.build_params(ctx, str(tmp_path), tau=seeded_tau())
    names = sorted(p.split("/")[-1] for p in paths)
    want = ["power-tau-5DWgDV-11.bin"] + ["amt-%s-5DWgDV-06-11.bin" % k for c in range(3)
                                        for k in ("prove-coset%d" % c, "prove-coset%d-mont" % c, "verify-coset%d" % c)]
    assert names == sorted(want)
    fresh = Z.Context(GEO)
    try:
        for c in range(GEO.coset_n):
            with open(str(tmp_path / ("amt-prove-coset%d-mont-5DWgDV-06-11.bin" % c)), "rb") as f:
                fresh.params_load_bamt(c, f.read())
            assert np.array_equal(fresh.params_download(c), ctx.params_download(c))
    finally:
        fresh.close()


def test_params_io_errors(ctx):
    fresh = Z.Context(GEO)
    try:
        with pytest.raises(Z.ZgrError) as e:
            fresh.write_params_file(Z.FILE_AMT_MONT, 0)
        assert e.value.code == Z.ERR_NOT_LOADED
        with pytest.raises(Z.ZgrError) as e:
            fresh.power_tau_download()
        assert e.value.code == Z.ERR_NOT_LOADED
        with pytest.raises(Z.ZgrError) as e:
            fresh.params_from_power_tau(0)
        assert e.value.code == Z.ERR_NOT_LOADED
        # prove bases from G1 powers alone leave no verifier data to write
        fresh.power_tau_setup(seeded_tau())
        g1, _, hi, _ = fresh.power_tau_download()
        fresh.params_from_pp(0, g1, hi)
        with pytest.raises(Z.ZgrError) as e:
            fresh.write_params_file(Z.FILE_AMT_VERIFY, 0)
        assert e.value.code == Z.ERR_NOT_LOADED
    finally:
        fresh.close()
    with pytest.raises(Z.ZgrError):
        ctx.write_params_file(99, 0)
    with pytest.raises(Z.ZgrError):
        ctx.write_params_file(Z.FILE_AMT_VERIFY, 0, GEO.h + 1)
