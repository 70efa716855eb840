"""The GLV constants of the device's EC-FFT scalar multiplication (struct Glv in
0g-da-encoder_amd/csrc/kernels.h), no device work.  They are pinned by what the
mathematics fixes: lambda and beta are cube roots of unity mod r and mod p, and the
endomorphism (x, y) -> (beta x, y) is multiplication by lambda on G1; (a1, b1) and
(a2, b2) lie in the lattice {(x, y): x + lambda y = 0 mod r} with determinant r; g1, g2
are the stated roundings; and the split formula gives |k1|, |k2| < 2^127 with
k1 + lambda k2 = k (mod r), at the edges of the scalar range and at random."""
import os
import random
import re

import pairing as E

HDR = os.path.join(os.path.dirname(__file__), "..", "0g-da-encoder_amd", "csrc", "kernels.h")
LAMBDA = 0xB3C4D79D41A917585BFC41088D8DAAA78B17EA66B99C90DD
BETA = 0x59E26BCEA0D48BACD4F263F1ACDB5C4F5763473177FFFFFE


def _glv():
    src = open(HDR).read()
    body = src[src.index("struct Glv {"):]
    body = body[:body.index("\n};")]
    out = {}
    for name, words in re.findall(r"uint32_t (\w+)\[\d+\] = \{([^}]*)\}", body):
        limbs = [int(w.strip().rstrip("u"), 16) for w in words.split(",")]
        radix = 29 if name == "BETA29" else 32
        out[name] = sum(v << (radix * i) for i, v in enumerate(limbs))
    assert "%x" % LAMBDA in body and "%x" % BETA in body  # the documented values
    return out


def test_endomorphism():
    c = _glv()
    p, r = E.P, E.R
    assert (LAMBDA * LAMBDA + LAMBDA + 1) % r == 0 and pow(BETA, 3, p) == 1 and BETA != 1
    assert c["BETA29"] == BETA * pow(2, 261, p) % p  # 29-bit Montgomery form
    for k in (1, 7, 123456789):
        pt = E.g1_mul(E.G1_GEN, k)
        assert E.g1_mul(pt, LAMBDA) == (BETA * pt[0] % p, pt[1])


def test_lattice_and_split():
    c = _glv()
    r = E.R
    a1, a2, b1, b2 = c["A1"], c["A2"], -c["NB1"], c["B2"]
    assert (a1 + b1 * LAMBDA) % r == 0 and (a2 + b2 * LAMBDA) % r == 0
    assert a1 * b2 - a2 * b1 == r
    assert c["G1"] == (b2 * 2 ** 256 + r // 2) // r and c["G2"] == (-b1 * 2 ** 256 + r // 2) // r
    rng = random.Random(5)
    for k in [0, 1, 2, r - 1, r - 2, (r - 1) // 2, LAMBDA, r - LAMBDA] + [rng.randrange(r) for _ in range(3000)]:
        c1 = (k * c["G1"] + 2 ** 255) >> 256
        c2 = (k * c["G2"] + 2 ** 255) >> 256
        assert c1 < 2 ** 64 and c2 < 2 ** 127  # the limb counts glv_split keeps
        k1, k2 = k - c1 * a1 - c2 * a2, -c1 * b1 - c2 * b2
        assert (k1 + LAMBDA * k2 - k) % r == 0
        assert abs(k1) < 2 ** 127 and abs(k2) < 2 ** 127
