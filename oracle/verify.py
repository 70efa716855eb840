"""Verifier-side restatement of the reference (TEST INFRASTRUCTURE: only tests/ import
this module).

* Verify params of one coset from a known tau (AMTParams::from_pp, crates/amt/src/
  prove_params/generate.rs:139-195): the G1 basis is the oracle's (line h of its prove
  bases); g2 = G2 generator, high_g2 = [tau^(3N)] G2 (PowerTau::setup_inner,
  crates/amt/src/power_tau.rs:54-80); vanishes[d] (depth d + 1, D = 2^(d+1) points,
  m = N / D) from gen_vanishes (generate.rs:234-263) in closed form
      V[t] = [(tau'^N - 1) / (tau'^m - w_D^t)] G2,  tau' = tau / cf_c,
  stored bit-reversed (enact, generate.rs:122-126).  The closed form is the sum
  tau'^(N-m) + sum_{i=1}^{D-1} tau'^((i-1) m) w_D^(-i t) of gen_vanishes' inverse DFT
  (a geometric series); tests/test_pairing_oracle.py checks the two agree.
* EncodedSlice::verify (crates/encoder/src/encoder/slice.rs:72-95) with
  EncodedSliceAMT::verify (amt/slice.rs:53-105), BlobRow::verify (crates/amt/src/blob/
  encode.rs:224-239), verify_amt_proof (crates/amt/src/verify_params.rs:133-225) and
  EncodedSliceMerkle::verify (merkle/slice.rs:50-104).  Errors are returned as the
  reference's Debug text (VerifierError and friends).
* A deferred check of many slices: every pairing equation of every slice folded with
  random coefficients into one multi-pairing (DeferredVerifier::fast_check_pairing,
  crates/amt/src/deferred_verification.rs:94-131), the MSM equations into one G1
  equality (fast_check_msm, :133-150).
"""
import random

import pairing as E

R = E.R
P = E.P
G28 = 19103219067921713944291392827692070036145651957329286315305642004821462161904


def _bitrev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def _keccak(msg):
    import oracle

    return oracle.keccak256(msg)


def g1_from_canon(b):
    """64 B canonical x || y (zeros = infinity) -> affine tuple / None"""
    if not any(b):
        return None
    return int.from_bytes(b[:32], "little"), int.from_bytes(b[32:64], "little")


def g1_from_wire(b):
    """ark-uncompressed G1 (flags in y's top byte) -> affine tuple / None"""
    raw = bytearray(b)
    flags = raw[63] & 0xC0
    raw[63] &= 0x3F
    if flags & 0x40:
        return None
    return int.from_bytes(bytes(raw[:32]), "little"), int.from_bytes(bytes(raw[32:]), "little")


class VerifyParams:
    """AMTVerifyParams of one coset from a known tau; vanish points on first use."""

    def __init__(self, tau, log_col, log_row, coset, basis_canon):
        self.tau, self.log_col, self.log_row, self.coset = tau % R, log_col, log_row, coset
        self.log_n = log_col + log_row
        self.n = 1 << self.log_n
        self.basis = basis_canon  # list of affine tuples / None, stored (bit-reversed) order
        cf = pow(G28, _bitrev(coset, 28 - self.log_n), R)
        self.tau_c = self.tau * pow(cf, R - 2, R) % R  # tau' = tau / cf (from_pp's w^idx scaling)
        self.g2 = E.G2_GEN
        self.high_g2 = E.g2_mul(E.G2_GEN, pow(self.tau, 3 * self.n, R))
        self._van = {}

    def vanish(self, d, k):
        """vanishes[d][k]: depth d + 1, stored index k"""
        key = (d, k)
        if key not in self._van:
            D = 2 << d
            m = self.n // D
            wD = pow(G28, 1 << (28 - (d + 1)), R)
            t = _bitrev(k, d + 1)
            num = (pow(self.tau_c, self.n, R) - 1) % R
            den = (pow(self.tau_c, m, R) - pow(wD, t, R)) % R
            self._van[key] = E.g2_mul(E.G2_GEN, num * pow(den, R - 2, R) % R)
        return self._van[key]


def _g1_msm(points, scalars):
    acc = None
    for pt, k in zip(points, scalars):
        if pt is not None and k % R:
            acc = E.g1_add(acc, E.g1_mul(pt, k))
    return acc


class Collector:
    """Pairing equations e(a, b) == e(c, d) and G1 equations, either checked at once or
    kept for one batched check."""

    def __init__(self, deferred):
        self.deferred = deferred
        self.pairs = []  # (a, b, c, d)
        self.msm = []  # (points, scalars, answer)

    def pairing(self, a, b, c, d, err):
        if self.deferred:
            self.pairs.append((a, b, c, d))
            return None
        return None if E.multi_pairing_is_one([(a, b), (E.g1_neg(c), d)]) else err

    def g1_equal(self, points, scalars, answer, err):
        if self.deferred:
            self.msm.append((points, scalars, answer))
            return None
        return None if _g1_msm(points, scalars) == answer else err

    def fast_check(self, rng=None):
        """deferred_verification.rs:94-150 with random coefficients: one multi-pairing
        (equal G2 arguments merged) and one G1 equality."""
        rng = rng or random.Random()
        by_g2 = {}
        for a, b, c, d in self.pairs:
            k = rng.randrange(1, R)
            for pt, q, s in ((a, b, k), (c, d, R - k)):
                if pt is None or q is None:
                    continue
                key = (q[0], q[1])
                by_g2[key] = E.g1_add(by_g2.get(key), E.g1_mul(pt, s))
        if not E.multi_pairing_is_one([(pt, q) for q, pt in by_g2.items()]):
            return False
        lhs, rhs = None, None
        for points, scalars, answer in self.msm:
            k = rng.randrange(1, R)
            lhs = E.g1_add(lhs, _g1_msm(points, [s * k % R for s in scalars]))
            rhs = E.g1_add(rhs, E.g1_mul(answer, k) if answer is not None else None)
        return lhs == rhs


def verify_amt_proof(vp, data, batch_index, proof, high, commitment, col):
    """verify_amt_proof (crates/amt/src/verify_params.rs:133-211); returns an
    AmtProofError name or None."""
    depth = len(proof)
    num_batch = 1 << depth
    batch = vp.n // num_batch
    if batch != len(data):
        return "UnexpectedDataLength"
    if batch_index >= num_batch:
        return "IncorrectPosition"
    acc = None
    for d in range(depth - 1, -1, -1):
        com, quot = proof[d]
        vi = batch_index >> (depth - 1 - d)
        err = col.pairing(com, vp.g2, quot, vp.vanish(d, vi ^ 1), "KzgError(%d)" % d)
        if err:
            return err
        acc = E.g1_add(acc, com)
    pts = vp.basis[batch_index * batch:(batch_index + 1) * batch]
    err = col.g1_equal(pts, data, E.g1_add(commitment, E.g1_neg(acc)), "InconsistentCommitment")
    if err:
        return err
    return col.pairing(high, vp.g2, commitment, vp.high_g2, "FailedLowDegreeTest")


def file_root(roots):
    """compute_file_root (crates/encoder/src/encoder/blob.rs:64-72)"""
    return _keccak(_keccak(roots[0] + roots[1]) + roots[2])


def verify_slice(s, params, auth_commitment, auth_root, geo, col=None):
    """EncodedSlice::verify over a parsed slice (zg_rocm.wire.parse_encoded_slice);
    params: one VerifyParams per coset; auth_commitment: affine tuple; auth_root: 32 B.
    Returns None or the reference's error as Debug text."""
    col = col or Collector(False)
    ROW, COL, RE = 1 << geo.log_row, 1 << geo.log_col, geo.row_encoded
    idx = s["index"]
    if idx != s["amt_index"]:
        return "UnmatchedAMTIndex { row_index: %d, amt_index: %d }" % (idx, s["amt_index"])
    if idx != s["leaf_index"]:
        return "UnmatchedMerkleIndex { row_index: %d, merkle_index: %d }" % (idx, s["leaf_index"])
    row = [bytes(x) for x in s["row"]]
    values = [int.from_bytes(x, "little") for x in row]
    # EncodedSliceAMT::verify
    if g1_from_wire(s["commitment"]) != auth_commitment:
        return "AMT(IncorrectCommitment)"
    if len(row) != COL:
        return "AMT(IncorrectRowSize { actual: %d, expected: %d })" % (len(row), COL)
    if s["amt_index"] >= RE:
        return "AMT(RowIndexOverflow { actual: %d, expected_max: %d })" % (s["amt_index"], RE)
    coset, local = divmod(s["amt_index"], ROW)
    if local != s["row_index"]:
        return "AMT(UnmatchedCosetIndex { coset_index: %d, local_index: %d, amt_index: %d })" % (
            coset, local, s["row_index"])
    data = [values[_bitrev(i, geo.log_col)] for i in range(COL)]  # index_reverse
    proof = [(g1_from_wire(c), g1_from_wire(q)) for c, q in s["amt_proof"]]
    err = verify_amt_proof(params[coset], data, _bitrev(s["row_index"], geo.log_row), proof,
                           g1_from_wire(s["high_commitment"]), auth_commitment, col)
    if err:
        return "AMT(IncorrectProof { coset_index: %d, amt_index: %d, error: %s })" % (coset, s["row_index"], err)
    # EncodedSliceMerkle::verify
    roots = [bytes(x) for x in s["merkle_root"]]
    if file_root(roots) != auth_root:
        return "Merkle(IncorrectRoot)"
    li = s["leaf_index"]
    if li >= RE:
        return "Merkle(RowIndexOverflow { actual: %d, expected_max: %d })" % (li, RE)
    layer = [_keccak(b"".join(row[8 * i:8 * i + 8])) for i in range(COL // 8)]
    while len(layer) > 1:
        layer = [_keccak(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
    if layer[0] != bytes(s["leaf"]):
        return "Merkle(IncorrectLocalRoot { row_index: %d })" % li
    pos, acc = li % geo.n, bytes(s["leaf"])
    for sib in s["merkle_proof"]:
        acc = _keccak(bytes(sib) + acc) if pos % 2 == 1 else _keccak(acc + bytes(sib))
        pos >>= 1
    if acc != roots[li // ROW]:
        return "Merkle(IncorrectProof { row_index: %d })" % li
    return None
