"""Independent pure-Python BN254 arithmetic used to pin the C oracle.

Written from the textbook definitions (affine chord-and-tangent addition,
schoolbook DFT, Lagrange interpolation), not from the oracle's code, so that a
mistake in the oracle's limb arithmetic, NTT indexing or MSM shows up as a
mismatch.  Python ints only; intended for small sizes.
"""

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
G1 = (1, 2)
TWO_ADICITY = 28
MONT_R = 1 << 256


def fr_root_of_unity(log_n):
    """Primitive 2^log_n-th root of unity: 5^((r-1)/2^log_n), 5 = Fr generator."""
    return pow(5, (R - 1) >> log_n, R)


def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - 3) % P == 0


def g1_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    x1, y1 = a
    x2, y2 = b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, P - 2, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, P - 2, P) % P
    x3 = (lam * lam - x1 - x2) % P
    y3 = (lam * (x1 - x3) - y1) % P
    return (x3, y3)


def g1_neg(a):
    return None if a is None else (a[0], (-a[1]) % P)


def g1_mul(pt, k):
    k %= R
    acc = None
    add = pt
    while k:
        if k & 1:
            acc = g1_add(acc, add)
        add = g1_add(add, add)
        k >>= 1
    return acc


def g1_msm(points, scalars):
    acc = None
    for p, s in zip(points, scalars):
        acc = g1_add(acc, g1_mul(p, s))
    return acc


def dft(coeffs, omega):
    n = len(coeffs)
    return [sum(c * pow(omega, i * t, R) for i, c in enumerate(coeffs)) % R for t in range(n)]


def idft(evals, omega):
    n = len(evals)
    ninv = pow(n, R - 2, R)
    winv = pow(omega, R - 2, R)
    return [v * ninv % R for v in dft(evals, winv)]


def poly_eval(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % R
    return acc


def bitrev(i, bits):
    return int(format(i, "0%db" % bits)[::-1], 2) if bits else 0


def keccak_f1600_ref(state):
    """Reference permutation on a list of 25 ints, indexed A[x + 5y] (FIPS 202)."""
    RC = []
    lfsr = 1
    for _ in range(24):
        rc = 0
        for j in range(7):
            # rc(t) from the FIPS 202 LFSR x^8 + x^6 + x^5 + x^4 + 1
            bit = lfsr & 1
            if bit:
                rc |= 1 << ((1 << j) - 1)
            lfsr = ((lfsr << 1) ^ (0x171 if lfsr & 0x80 else 0)) & 0xFF
        RC.append(rc)
    M = (1 << 64) - 1

    def rot(v, s):
        s %= 64
        return ((v << s) | (v >> (64 - s))) & M if s else v

    offs = [[0] * 5 for _ in range(5)]
    x, y = 1, 0
    for t in range(24):
        offs[x][y] = (t + 1) * (t + 2) // 2
        x, y = y, (2 * x + 3 * y) % 5
    A = list(state)
    for rnd in range(24):
        C = [A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20] for x in range(5)]
        D = [C[(x - 1) % 5] ^ rot(C[(x + 1) % 5], 1) for x in range(5)]
        A = [A[i] ^ D[i % 5] for i in range(25)]
        B = [0] * 25
        for x in range(5):
            for y in range(5):
                B[y + 5 * ((2 * x + 3 * y) % 5)] = rot(A[x + 5 * y], offs[x][y])
        A = [B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]) for y in range(5) for x in range(5)]
        # the comprehension above iterates y outer, x inner -> index x + 5y
        A[0] ^= RC[rnd]
    return A


def keccak256_ref(msg, pad=0x01):
    rate = 136
    data = bytearray(msg)
    q = rate - (len(data) % rate)
    blk = bytearray(q)
    blk[0] ^= pad
    blk[-1] ^= 0x80
    data += blk
    A = [0] * 25
    for off in range(0, len(data), rate):
        for i in range(rate // 8):
            A[i] ^= int.from_bytes(data[off + 8 * i: off + 8 * i + 8], "little")
        A = keccak_f1600_ref(A)
    return b"".join(A[i].to_bytes(8, "little") for i in range(4))
