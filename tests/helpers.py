"""Shared test helpers: seeded inputs, oracle params, HIP context factory."""
import functools
import os

import numpy as np

import oracle as O
import zg_rocm as Z

TEST_TAU_SEED = 7
DATA_SEED = 222  # crates/encoder/src/encoder/blob.rs:257 uses seed 222 for its blobs
CPU_THREADS = max(1, min(16, os.cpu_count() or 1))


def seeded_tau():
    return Z.tau_from_seed(TEST_TAU_SEED)


@functools.lru_cache(maxsize=None)
def oracle_params(log_col, log_row, coset_n=3, tau=None):
    """Prove bases from the oracle's closed form, one uint64 array per coset."""
    tau = seeded_tau() if tau is None else tau
    O.set_threads(CPU_THREADS)
    log_n = log_col + log_row
    return tuple(O.params_from_tau(tau, log_n, log_row, c, 0) for c in range(coset_n))


def ri_from_y(y_mont, log_n):
    """index_reverse (crates/amt/src/utils.rs:70-88) of a (N, 4) uint64 array."""
    n = 1 << log_n
    idx = np.array([int(format(i, "0%db" % log_n)[::-1], 2) for i in range(n)], dtype=np.int64)
    return np.ascontiguousarray(np.asarray(y_mont).reshape(n, 4)[idx])


def random_fr_mont(rng, n):
    """n uniform Fr elements in Montgomery form (via the oracle's to_mont)."""
    vals = [int.from_bytes(rng.bytes(32), "little") % O.R_MOD for _ in range(n)]
    return O.fr_to_mont_array(vals)


G2_BYTES = 128  # ark-serialize uncompressed G2 affine


def bamt_image(bases_u64, log_n, h):
    """write_amt_params layout (crates/amt/src/prove_params/fast_serde_bn254.rs:20-53):
    header, degree, sub-degree, g2, high_g2, basis, quotients[0..h), vanishes (2^(d+1)
    G2 points at level d), high_basis.  The G2 sections are verifier data and are zero
    placeholders here; `bases_u64` is the device layout quotients || basis || high."""
    line = (1 << log_n) * 64
    b = np.ascontiguousarray(bases_u64).view(np.uint8)
    assert b.size == (h + 2) * line
    parts = [b"bamt" + bytes([log_n, h]), bytes(2 * G2_BYTES), b[h * line:(h + 1) * line].tobytes()]
    parts += [b[d * line:(d + 1) * line].tobytes() for d in range(h)]
    parts += [bytes((2 << d) * G2_BYTES) for d in range(h)]
    parts.append(b[(h + 1) * line:(h + 2) * line].tobytes())
    return b"".join(parts)


class OracleOutputs:
    """zgr_encoded-shaped view of the oracle's Encoded (for zg_rocm.EncodedBlob)."""

    def __init__(self, e):
        self.data, self.commitment, self.comm_tree = e.data, e.commitment, e.comm_tree
        self.proof_tree, self.high, self.merkle, self.file_root = e.proof_tree, e.high, e.merkle, e.root
