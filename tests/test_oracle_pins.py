"""Pins of the CPU oracle against independent mathematics.

The reference holds no golden vectors for this path (every test draws tau from
thread_rng), so the oracle is pinned by:
  * public known-answer tests (Keccak-256 of "" and "abc") and FIPS SHA3-256 from
    hashlib, which shares the Keccak-f[1600] permutation and differs only in padding;
  * textbook restatements in tests/pyref.py (Python ints): affine group law, DFT
    by definition, Lagrange/polynomial evaluation;
  * the reference's own unit tests restated (file:line cited per test);
  * closed forms of the AMT outputs for a known tau.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
import pyref

R = O.R_MOD


# ---------------------------------------------------------------- Keccak
def test_keccak256_known_answers():
    assert O.keccak256(b"").hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"
    assert O.keccak256(b"abc").hex() == "4e03657aea45a94fc7d47ba826c8d667c0d1e6e33a64a036ec44f58fa12d6c45"


@pytest.mark.parametrize("length", [0, 1, 31, 64, 135, 136, 137, 256, 271, 272, 1000])
def test_keccak_permutation_vs_sha3(length):
    msg = os.urandom(length)
    assert O.keccak256(msg, 0x06) == hashlib.sha3_256(msg).digest()
    assert pyref.keccak256_ref(msg, 0x01) == O.keccak256(msg, 0x01)


# ---------------------------------------------------------------- fields
@pytest.mark.parametrize("which,mod", [(0, O.R_MOD), (1, O.P_MOD)])
def test_field_mul_inverse(which, mod):
    rng = np.random.default_rng(which)
    for _ in range(50):
        a = int.from_bytes(rng.bytes(32), "little") % mod
        b = int.from_bytes(rng.bytes(32), "little") % mod
        am, bm = O.to_mont(which, a), O.to_mont(which, b)
        assert am == a * (1 << 256) % mod
        assert O.from_mont(which, O.fmul(which, am, bm)) == a * b % mod
        if a:
            assert O.from_mont(which, O.fmul(which, O.finv(which, am), am)) == 1
    # edge values
    for a in (0, 1, mod - 1):
        assert O.from_mont(which, O.to_mont(which, a)) == a


def test_scalar_to_h256_and_bytes_to_scalar_kats():
    # crates/encoder/src/utils.rs:74-89: zero chunk -> 0; one -> [1, 0, ...]
    blob = O.raw_to_blob(b"", 5)
    assert O.from_mont(0, O.limbs_to_int(blob[0])) == 0
    assert O.from_mont(0, O.to_mont(0, 1)) == 1


# ---------------------------------------------------------------- G1
def test_g1_against_affine_group_law():
    rng = np.random.default_rng(2)
    for _ in range(6):
        k = int.from_bytes(rng.bytes(32), "little") % R
        assert O.point_to_ints(O.g1_mul_gen(k)) == pyref.g1_mul(pyref.G1, k)
    assert O.point_to_ints(O.g1_mul_gen(0)) is None
    assert O.point_to_ints(O.g1_mul_gen(R - 1)) == pyref.g1_neg(pyref.G1)
    p = O.g1_mul_gen(5)
    q = O.g1_mul_gen(7)
    assert O.point_to_ints(O.g1_add(p, q)) == pyref.g1_mul(pyref.G1, 12)
    assert O.point_to_ints(O.g1_add(p, p)) == pyref.g1_mul(pyref.G1, 10)       # doubling branch
    assert O.point_to_ints(O.g1_add(p, O.g1_mul_gen(R - 5))) is None          # P + (-P)
    assert pyref.on_curve(O.point_to_ints(O.g1_mul_gen(123456789)))


def test_msm_against_naive():
    rng = np.random.default_rng(4)
    for n in (1, 3, 17, 40):
        ks = [int.from_bytes(rng.bytes(32), "little") % R for _ in range(n)]
        ks[0] = 0
        pts = [pyref.g1_mul(pyref.G1, i + 2) for i in range(n)]
        bases = np.zeros((n, 8), dtype=np.uint64)
        for i, (x, y) in enumerate(pts):
            bases[i, :4] = O.int_to_limbs(O.to_mont(1, x))
            bases[i, 4:] = O.int_to_limbs(O.to_mont(1, y))
        scal = O.ints_to_limbs(ks)
        exp = pyref.g1_mul(pyref.G1, sum((i + 2) * k for i, k in enumerate(ks)))
        assert O.point_to_ints(O.msm(bases, scal)) == exp


# ---------------------------------------------------------------- domain / NTT
def test_domain_roots():
    # ark-poly get_root_of_unity: the 2^28-th root squared down; order exactly 2^k
    for k in (1, 5, 11, 20, 28):
        w = O.domain_root(k)
        assert w == pyref.fr_root_of_unity(k)
        assert pow(w, 1 << k, R) == 1 and pow(w, 1 << (k - 1), R) != 1


def test_coset_factor_kat():
    # crates/amt/src/blob/encode.rs:349-355: coset_factor(16, 1) == root of unity of order 32
    assert O.coset_factor(4, 1) == pyref.fr_root_of_unity(5)
    # at production size the factors are primitive 2^21 / 2^22 roots (bit-reversed index)
    cf1, cf2 = O.coset_factor(20, 1), O.coset_factor(20, 2)
    assert cf1 == pyref.fr_root_of_unity(21)
    assert cf2 == pyref.fr_root_of_unity(22)
    assert O.coset_factor(20, 0) == 1


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 6])
def test_fft_against_dft(log_n):
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    vals = [int.from_bytes(rng.bytes(32), "little") % R for _ in range(n)]
    a = O.fr_to_mont_array(vals)
    w = pyref.fr_root_of_unity(log_n)
    assert O.fr_from_mont_array(O.fft(a, log_n)) == pyref.dft(vals, w)
    assert O.fr_from_mont_array(O.ifft(a, log_n)) == pyref.idft(vals, w)


def test_erasure_encoding_identity():
    # crates/amt/src/blob/encode.rs:316-347: cosets 0 and 1 are the even / odd entries
    # of the 2N-point FFT of the zero-padded coefficients
    log_n = 6
    n = 1 << log_n
    rng = np.random.default_rng(8)
    data = O.fr_to_mont_array([int.from_bytes(rng.bytes(32), "little") % R for _ in range(n)])
    coeff = O.ifft(data, log_n)
    big = np.zeros((2 * n, 4), dtype=np.uint64)
    big[:n] = coeff
    ans = O.fr_from_mont_array(O.fft(big, log_n + 1))
    assert O.fr_from_mont_array(O.to_coset_blob(data, log_n, 0)) == ans[0::2]
    assert O.fr_from_mont_array(O.to_coset_blob(data, log_n, 1)) == ans[1::2]
    # coset 2 evaluates at cf_2 * w^j, cf_2 a primitive 4N-th root
    c2 = O.fr_from_mont_array(O.to_coset_blob(data, log_n, 2))
    cvals = O.fr_from_mont_array(coeff)
    cf2 = O.coset_factor(log_n, 2)
    w = O.domain_root(log_n)
    for j in (0, 1, 17):
        assert c2[j] == pyref.poly_eval(cvals, cf2 * pow(w, j, R) % R)


# ---------------------------------------------------------------- params
@pytest.mark.parametrize("coset", [0, 1, 2])
def test_params_closed_form_equals_ecfft_restatement(coset):
    # generate.rs:139-232 through the G1 EC-FFT vs the closed form the GPU uses
    tau = 0x1234567890ABCDEF1234567890ABCDEF1234
    for log_n, h in ((3, 2), (4, 3)):
        a = O.params_from_tau(tau, log_n, h, coset, 0)
        b = O.params_from_tau(tau, log_n, h, coset, 1)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("start", [0, 3 * 16, 12345])
def test_power_tau_textbook(start):
    # power_tau (power_tau.rs:28-42): out[i] = [tau^(start + i)] G1, against the textbook
    # affine group law; start 3N is the high powers of setup_inner (power_tau.rs:66-71)
    tau = 0x0BADC0FFEE0DDF00D1234567
    log_n = 4
    out = O.power_tau(tau, start, log_n).reshape(-1, 64)
    for i in range(1 << log_n):
        x = int.from_bytes(bytes(out[i, :32]), "little")
        y = int.from_bytes(bytes(out[i, 32:]), "little")
        assert (x, y) == pyref.g1_mul(pyref.G1, pow(tau, start + i, R)), i


@pytest.mark.parametrize("log_n,h", [(3, 2), (4, 3), (5, 2)])
def test_params_point_equals_params_from_tau(log_n, h):
    # the single-point closed form (used to sample production params) against the whole
    # buffer, at every position of every line of every coset
    tau = 0x77777777777777777777777777777
    n = 1 << log_n
    for coset in range(3):
        full = O.params_from_tau(tau, log_n, h, coset, 0).reshape(h + 2, n, 8)
        for line in range(h + 2):
            for pos in range(n):
                assert np.array_equal(O.params_point(tau, log_n, h, coset, line, pos), full[line, pos]), (coset, line, pos)


def test_gen_basis_brute_force():
    # generate.rs:276-290: basis_t == MSM(g1pp, ifft(one-hot t)) -- here [L_t(tau)] G1
    tau = 987654321987654321
    log_n = 3
    n = 1 << log_n
    bases = O.params_from_tau(tau, log_n, 1, 0, 0).reshape(3, n, 8)
    w = pyref.fr_root_of_unity(log_n)
    for t in range(n):
        onehot = [0] * n
        onehot[t] = 1
        coeff = pyref.idft(onehot, w)
        s = pyref.poly_eval(coeff, tau)
        pos = pyref.bitrev(t, log_n)
        x = O.from_mont(1, O.limbs_to_int(bases[1, pos, :4]))
        y = O.from_mont(1, O.limbs_to_int(bases[1, pos, 4:]))
        assert (x, y) == pyref.g1_mul(pyref.G1, s)


def test_gen_quotients_brute_force():
    # generate.rs:292-312: quot_t = size_inv * w^t * sum_{j<m} tau^j w^{t(m-1-j)}
    tau = 555555555555
    log_n, d = 3, 1
    n, m = 1 << log_n, (1 << log_n) >> d
    bases = O.params_from_tau(tau, log_n, d, 0, 0).reshape(d + 2, n, 8)
    w = pyref.fr_root_of_unity(log_n)
    ninv = pow(n, R - 2, R)
    for t in range(n):
        s = sum(pow(tau, j, R) * pow(w, t * (m - 1 - j), R) for j in range(m)) * ninv * pow(w, t, R) % R
        pos = pyref.bitrev(t, log_n)
        x = O.from_mont(1, O.limbs_to_int(bases[0, pos, :4]))
        assert x == pyref.g1_mul(pyref.G1, s)[0]


# ---------------------------------------------------------------- AMT prover
def _amt_setup(log_n, h, coset, tau, seed):
    n = 1 << log_n
    rng = np.random.default_rng(seed)
    y = [int.from_bytes(rng.bytes(32), "little") % R for _ in range(n)]
    bases = O.params_from_tau(tau, log_n, h, coset, 0)
    ri = [y[pyref.bitrev(i, log_n)] for i in range(n)]
    return y, bases, O.fr_to_mont_array(ri)


@pytest.mark.parametrize("coset", [0, 1])
def test_amt_prove_closed_form(coset):
    """commitment = [f(tau)], node commitments = [sum_{j in S} y_j L_j(tau')],
    proofs = [sum_{j in S} y_j (w^j/N)(tau'^m - w^(jm))/(tau' - w^j)], high = [tau^3N f(tau)]
    with tau' = tau / cf and S = {j : j = bitrev_d(k) mod 2^d}."""
    log_n, h = 5, 3
    n = 1 << log_n
    tau = 31415926535897932384626433
    y, bases, ri = _amt_setup(log_n, h, coset, tau, 3 + coset)
    com, ct, pt, hi = O.amt_prove(bases, log_n, h, ri)
    cf = O.coset_factor(log_n, coset)
    tp = tau * pow(cf, R - 2, R) % R
    w = pyref.fr_root_of_unity(log_n)
    ninv = pow(n, R - 2, R)

    def lag(j):
        wj = pow(w, j, R)
        return wj * ninv * (pow(tp, n, R) - 1) * pow(tp - wj, R - 2, R) % R

    f_tp = sum(y[j] * lag(j) for j in range(n)) % R  # the coset evaluations interpolate f(cf X)
    assert O.point_to_ints(com) == pyref.g1_mul(pyref.G1, f_tp)
    assert O.point_to_ints(hi) == pyref.g1_mul(pyref.G1, f_tp * pow(tau, 3 * n, R))
    for d in range(1, h + 1):
        m = n >> d
        for k in range(1 << d):
            S = [j for j in range(n) if j % (1 << d) == pyref.bitrev(k, d)]
            c_s = sum(y[j] * lag(j) for j in S) % R
            p_s = 0
            for j in S:
                wj = pow(w, j, R)
                p_s += y[j] * wj * ninv * (pow(tp, m, R) - pow(wj, m, R)) * pow(tp - wj, R - 2, R)
            idx = (1 << d) - 2 + k
            assert O.point_to_ints(ct[idx]) == pyref.g1_mul(pyref.G1, c_s), (d, k)
            assert O.point_to_ints(pt[idx]) == pyref.g1_mul(pyref.G1, p_s % R), (d, k)


def test_multiple_multiexp_aggregates_to_amt_prove():
    # prove_gpu.rs:89-123 process_gpu_output == prove.rs:85-103 gen_all_proofs
    log_n, h = 6, 3
    tau = 42424242
    _, bases, ri = _amt_setup(log_n, h, 2, tau, 11)
    com, ct, pt, hi = O.amt_prove(bases, log_n, h, ri)
    lines = O.multiple_multiexp(bases, log_n, h, ri)
    nb = 1 << h

    def psum(points):
        acc = None
        for p in points:
            acc = pyref.g1_add(acc, O.point_to_ints(p))
        return acc

    for d in range(h):
        g = nb >> (d + 1)
        for k in range(1 << (d + 1)):
            assert O.point_to_ints(pt[(1 << (d + 1)) - 2 + k]) == psum(lines[d * nb + k * g:d * nb + (k + 1) * g])
            assert O.point_to_ints(ct[(1 << (d + 1)) - 2 + k]) == psum(lines[h * nb + k * g:h * nb + (k + 1) * g])
    assert O.point_to_ints(hi) == psum(lines[(h + 1) * nb:(h + 2) * nb])
    assert O.point_to_ints(com) == psum(lines[h * nb:(h + 1) * nb])


# ---------------------------------------------------------------- packing / merkle / encode
def test_raw_packing_layout():
    # RawData::try_from (raw_data.rs:55-73): zero pad + u32 LE length in the last 4
    # bytes; RawBlob::from: 31-byte LE chunks (raw_blob.rs:33-40)
    log_n = 5
    n = 1 << log_n
    raw = bytes(range(1, 40))
    blob = O.raw_to_blob(raw, log_n)
    padded = bytearray(31 * n)
    padded[:len(raw)] = raw
    padded[-4:] = len(raw).to_bytes(4, "little")
    for i in range(n):
        assert O.from_mont(0, O.limbs_to_int(blob[i])) == int.from_bytes(padded[31 * i:31 * i + 31], "little")
    with pytest.raises(OverflowError):
        O.raw_to_blob(b"\x01" * (31 * n - 3), log_n)
    O.raw_to_blob(b"\x01" * (31 * n - 4), log_n)


def test_merkle_shape_and_fold():
    # merkle/blob.rs:115-128 (shape) and merkle/slice.rs:88-102 (fold order)
    log_col, log_row, coset_n = 5, 6, 3
    n = 1 << (log_col + log_row)
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, coset_n * n * 32, dtype=np.uint8)
    nl, layers, root = O.merkle(data, coset_n, log_row)
    assert nl == log_row + log_col - 2
    # brute-force the tree for coset 1 with the independent Keccak
    d = bytes(data)
    leaves = [pyref.keccak256_ref(d[256 * i:256 * i + 256]) for i in range(coset_n * n // 8)]
    lay = leaves
    while len(lay) > coset_n:
        lay = [pyref.keccak256_ref(lay[2 * i] + lay[2 * i + 1]) for i in range(len(lay) // 2)]
    assert [bytes(layers[i]) for i in range(3)] == lay
    k = pyref.keccak256_ref
    assert root == k(k(lay[0] + lay[1]) + lay[2])


def test_encode_commitment_closed_form_test_size():
    # full EncodedBlob::build at the test geometry: commitment = [f(tau)] G1 for the
    # interpolant f of the column-major raw blob; every coset agrees (amt/blob.rs:42-47)
    log_col, log_row = 5, 6
    log_n = log_col + log_row
    tau = 0xABCDEF0123456789ABCDEF0123456789ABCDEF
    O.set_threads(8)
    bases = [O.params_from_tau(tau, log_n, log_row, c, 0) for c in range(3)]
    rng = np.random.default_rng(222)
    raw = rng.integers(0, 256, 31 * (1 << log_n) - 4, dtype=np.uint8).tobytes()
    e = O.encode(bases, log_col, log_row, raw)
    rowmaj = O.raw_to_blob(raw, log_n).reshape(1 << log_row, 1 << log_col, 4)
    colmaj = np.ascontiguousarray(rowmaj.transpose(1, 0, 2)).reshape(-1, 4)
    ftau = O.eval_interp(colmaj, log_n, tau)
    assert O.point_to_ints(e.commitment) == pyref.g1_mul(pyref.G1, ftau)
    # and the interpolant evaluation itself against the polynomial form
    coeffs = O.fr_from_mont_array(O.ifft(colmaj, log_n))
    assert ftau == pyref.poly_eval(coeffs, tau)
    # coset-0 data equals the raw blob in row-major canonical form
    canon = O.fr_from_mont_array(rowmaj.reshape(-1, 4)[:64])
    got = [int.from_bytes(bytes(e.data[32 * i:32 * i + 32]), "little") for i in range(64)]
    assert got == canon
