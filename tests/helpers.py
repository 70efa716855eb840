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
