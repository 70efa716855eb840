"""BN254 pairing for the verifier-side restatement (TEST INFRASTRUCTURE: only tests/
import this module; it checks outputs, it is never the thing measured or shipped).

The reference verifies AMT proofs with ark-bn254's optimal ate pairing
(crates/amt/src/verify_params.rs:190-225, deferred_verification.rs:94-131).  Those
checks compare products of pairings with each other, so every non-degenerate bilinear
pairing on G1 x G2 accepts and rejects the same inputs.  This module computes the
reduced Tate pairing
    e(P, Q) = f_{r,P}(psi(Q)) ^ ((p^12 - 1) / r)
with plain Python integers:
  * Fq2 = Fq[i] / (i^2 + 1), Fq6 = Fq2[v] / (v^3 - xi), xi = 9 + i, Fq12 = Fq6[w] / (w^2 - v);
  * G2 on the twist y^2 = x^3 + 3 / xi over Fq2, untwisted by psi(x, y) = (x w^2, y w^3)
    (w^6 = xi) onto y^2 = x^3 + 3 over Fq12;
  * the Miller loop runs over the bits of r with affine multiples of P in Fq and
    evaluates each line at psi(Q) as a sparse Fq12 element; vertical lines lie in Fq6
    and vanish under the final exponentiation, which is a plain power.
Pinned in tests/test_pairing_oracle.py by bilinearity, non-degeneracy and order r.
"""
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617

# ---------------------------------------------------------------- Fq2
F2_ZERO = (0, 0)
F2_ONE = (1, 0)


def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_neg(a):
    return ((-a[0]) % P, (-a[1]) % P)


def f2_mul(a, b):
    t0, t1 = a[0] * b[0], a[1] * b[1]
    return ((t0 - t1) % P, ((a[0] + a[1]) * (b[0] + b[1]) - t0 - t1) % P)


def f2_sqr(a):
    return f2_mul(a, a)


def f2_scale(a, k):
    return (a[0] * k % P, a[1] * k % P)


def f2_inv(a):
    n = pow(a[0] * a[0] + a[1] * a[1], P - 2, P)
    return (a[0] * n % P, (-a[1]) * n % P)


XI = (9, 1)


def f2_mul_xi(a):  # (a0 + a1 i)(9 + i)
    return ((9 * a[0] - a[1]) % P, (a[0] + 9 * a[1]) % P)


def f2_pow(a, e):
    r = F2_ONE
    for bit in bin(e)[2:]:
        r = f2_sqr(r)
        if bit == "1":
            r = f2_mul(r, a)
    return r


# ---------------------------------------------------------------- Fq6 (c0 + c1 v + c2 v^2)
F6_ZERO = (F2_ZERO, F2_ZERO, F2_ZERO)
F6_ONE = (F2_ONE, F2_ZERO, F2_ZERO)


def f6_add(a, b):
    return tuple(f2_add(x, y) for x, y in zip(a, b))


def f6_sub(a, b):
    return tuple(f2_sub(x, y) for x, y in zip(a, b))


def f6_neg(a):
    return tuple(f2_neg(x) for x in a)


def f6_mul(a, b):
    a0, a1, a2 = a
    b0, b1, b2 = b
    t0, t1, t2 = f2_mul(a0, b0), f2_mul(a1, b1), f2_mul(a2, b2)
    c0 = f2_add(t0, f2_mul_xi(f2_sub(f2_mul(f2_add(a1, a2), f2_add(b1, b2)), f2_add(t1, t2))))
    c1 = f2_add(f2_sub(f2_mul(f2_add(a0, a1), f2_add(b0, b1)), f2_add(t0, t1)), f2_mul_xi(t2))
    c2 = f2_add(f2_sub(f2_mul(f2_add(a0, a2), f2_add(b0, b2)), f2_add(t0, t2)), t1)
    return (c0, c1, c2)


def f6_mul_v(a):  # a * v
    return (f2_mul_xi(a[2]), a[0], a[1])


# ---------------------------------------------------------------- Fq12 (d0 + d1 w)
F12_ONE = (F6_ONE, F6_ZERO)


def f12_mul(a, b):
    a0, a1 = a
    b0, b1 = b
    t0, t1 = f6_mul(a0, b0), f6_mul(a1, b1)
    c1 = f6_sub(f6_mul(f6_add(a0, a1), f6_add(b0, b1)), f6_add(t0, t1))
    return (f6_add(t0, f6_mul_v(t1)), c1)


def f12_sqr(a):
    return f12_mul(a, a)


def f12_pow(a, e):
    r = F12_ONE
    for bit in bin(e)[2:]:
        r = f12_sqr(r)
        if bit == "1":
            r = f12_mul(r, a)
    return r


FINAL_EXP = (P ** 12 - 1) // R

# ---------------------------------------------------------------- G1 (affine over Fq; None = infinity)
B1 = 3
G1_GEN = (1, 2)


def g1_on_curve(pt):
    return pt is None or (pt[1] * pt[1] - pt[0] ** 3 - B1) % P == 0


def g1_neg(pt):
    return None if pt is None else (pt[0], (-pt[1]) % P)


def g1_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    if a[0] == b[0]:
        if (a[1] + b[1]) % P == 0:
            return None
        lam = 3 * a[0] * a[0] * pow(2 * a[1], P - 2, P) % P
    else:
        lam = (b[1] - a[1]) * pow(b[0] - a[0], P - 2, P) % P
    x = (lam * lam - a[0] - b[0]) % P
    return (x, (lam * (a[0] - x) - a[1]) % P)


def _jdbl(X, Y, Z):  # Jacobian doubling on y^2 = x^3 + b (dbl-2009-l)
    if Z == 0 or Y == 0:
        return 1, 1, 0
    A, B = X * X % P, Y * Y % P
    C = B * B % P
    D = 2 * ((X + B) ** 2 - A - C) % P
    E_ = 3 * A % P
    X3 = (E_ * E_ - 2 * D) % P
    return X3, (E_ * (D - X3) - 8 * C) % P, 2 * Y * Z % P


def _jmadd(X, Y, Z, x, y):  # Jacobian + affine (madd-2007-bl)
    if Z == 0:
        return x, y, 1
    Z2 = Z * Z % P
    U2, S2 = x * Z2 % P, y * Z * Z2 % P
    H, r = (U2 - X) % P, 2 * (S2 - Y) % P
    if H == 0:
        return _jdbl(X, Y, Z) if r == 0 else (1, 1, 0)
    HH = H * H % P
    I = 4 * HH % P
    J, V = H * I % P, X * I % P
    X3 = (r * r - J - 2 * V) % P
    return X3, (r * (V - X3) - 2 * Y * J) % P, ((Z + H) ** 2 - Z2 - HH) % P


def g1_mul(pt, k):
    """k * pt (k reduced mod r) in Jacobian coordinates, affine result."""
    k %= R
    if pt is None or k == 0:
        return None
    X, Y, Z = 1, 1, 0
    for bit in bin(k)[2:]:
        X, Y, Z = _jdbl(X, Y, Z)
        if bit == "1":
            X, Y, Z = _jmadd(X, Y, Z, pt[0], pt[1])
    if Z == 0:
        return None
    zi = pow(Z, P - 2, P)
    zi2 = zi * zi % P
    return X * zi2 % P, Y * zi2 * zi % P


# ---------------------------------------------------------------- G2 (affine over Fq2 on the twist)
B2 = f2_mul((3, 0), f2_inv(XI))  # 3 / xi
G2_GEN = ((10857046999023057135944570762232829481370756359578518086990519993285655852781,
           11559732032986387107991004021392285783925812861821192530917403151452391805634),
          (8495653923123431417604973247489272438418190587263600148770280649306958101930,
           4082367875863433681332203403145435568316851327593401208105741076214120093531))


def g2_on_curve(q):
    if q is None:
        return True
    x, y = q
    return f2_sub(f2_sqr(y), f2_add(f2_mul(f2_sqr(x), x), B2)) == F2_ZERO


def g2_neg(q):
    return None if q is None else (q[0], f2_neg(q[1]))


def g2_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    if a[0] == b[0]:
        if f2_add(a[1], b[1]) == F2_ZERO:
            return None
        lam = f2_mul(f2_scale(f2_sqr(a[0]), 3), f2_inv(f2_scale(a[1], 2)))
    else:
        lam = f2_mul(f2_sub(b[1], a[1]), f2_inv(f2_sub(b[0], a[0])))
    x = f2_sub(f2_sub(f2_sqr(lam), a[0]), b[0])
    return (x, f2_sub(f2_mul(lam, f2_sub(a[0], x)), a[1]))


def g2_mul(q, k):
    """k * q for any integer k >= 0 (not reduced mod r, so r * q can be tested)."""
    acc = None
    for bit in bin(k)[2:] if k else "":
        acc = g2_add(acc, acc)
        if bit == "1":
            acc = g2_add(acc, q)
    return acc


# ---------------------------------------------------------------- Miller loop
def _line(lam, t, q):
    """Line through T (affine, Fq) with slope lam, evaluated at psi(Q):
    y' w^3 - lam x' w^2 + (lam x_T - y_T), as an element of Fq12."""
    x2, y2 = q
    c0 = ((lam * t[0] - t[1]) % P, 0)
    d0 = (c0, f2_scale(f2_neg(x2), lam), F2_ZERO)
    d1 = (F2_ZERO, y2, F2_ZERO)
    return (d0, d1)


def miller(pt, q):
    """f_{r,P}(psi(Q)) up to factors in Fq6 (killed by the final exponentiation)."""
    if pt is None or q is None:
        return F12_ONE
    f = F12_ONE
    t = pt
    for bit in bin(R)[3:]:
        lam = 3 * t[0] * t[0] * pow(2 * t[1], P - 2, P) % P
        f = f12_mul(f12_sqr(f), _line(lam, t, q))
        t = g1_add(t, t)
        if bit == "1":
            if t[0] == pt[0]:  # T = -P: the vertical line of the last step
                t = g1_add(t, pt)
                continue
            lam = (pt[1] - t[1]) * pow(pt[0] - t[0], P - 2, P) % P
            f = f12_mul(f, _line(lam, t, q))
            t = g1_add(t, pt)
    return f


def final_exp(f):
    return f12_pow(f, FINAL_EXP)


def pairing(pt, q):
    return final_exp(miller(pt, q))


def multi_pairing_is_one(pairs):
    """prod_j e(P_j, Q_j) == 1, one final exponentiation for the whole product."""
    f = F12_ONE
    for pt, q in pairs:
        f = f12_mul(f, miller(pt, q))
    return final_exp(f) == F12_ONE
