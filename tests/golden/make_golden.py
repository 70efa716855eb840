"""Writes tests/golden/encode_test_geometry.json: encode fixtures at the reference's test
geometry (2^6 rows x 2^5 columns, 3 cosets; crates/encoder/src/constants.rs:11-15), params
from the seeded insecure tau (seed 7), blobs of seeded bytes (seed 222, as
crates/encoder/src/encoder/blob.rs:257) of lengths 0, 1, 1234 and MAX_RAW_DATA_SIZE
(encoder/blob.rs:282-286).

The reference is Rust and holds no golden vectors for this path, so the outputs come from
the oracle and are pinned at generation time by facts independent of its encode code:
  * commitment = [f(tau)] G1 and high_commitment = [tau^(3N) f(tau)] G1, f the interpolant
    of the column-major blob (the AMT closed form);
  * coset 0's encoded rows are the raw blob itself (encode.rs:138-140);
  * the file root recomputed from the encoded data with the textbook Keccak-256 in
    tests/pyref.py (leaves of 8 elements, pairs up to the 3 coset roots, then
    k(k(r0 || r1) || r2); merkle/blob.rs:13-34, encoder/blob.rs:64-72).
Large outputs are stored as SHA-256 digests of their bytes.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "0g-da-encoder_amd")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import pyref  # noqa: E402
import zg_rocm as Z  # noqa: E402

GEO = Z.TEST_SETTING
TAU_SEED = 7
DATA_SEED = 222
LENGTHS = [0, 1, 1234, GEO.max_raw_data_size]
OUT = os.path.join(HERE, "encode_test_geometry.json")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def file_root_ref(data):
    """Merkle file root of the encoded data by the textbook Keccak-256."""
    nodes = [pyref.keccak256_ref(bytes(data[i:i + 256])) for i in range(0, len(data), 256)]
    while len(nodes) > 3:
        nodes = [pyref.keccak256_ref(nodes[i] + nodes[i + 1]) for i in range(0, len(nodes), 2)]
    return pyref.keccak256_ref(pyref.keccak256_ref(nodes[0] + nodes[1]) + nodes[2])


def main():
    tau = Z.tau_from_seed(TAU_SEED)
    bases = [O.params_from_tau(tau, GEO.log_n, GEO.log_row, c, 0) for c in range(GEO.coset_n)]
    blobs = []
    for length in LENGTHS:
        raw = Z.seeded_bytes(DATA_SEED, length)
        exp = O.encode(bases, GEO.log_col, GEO.log_row, raw)
        # independent pins
        rowmaj = O.raw_to_blob(raw, GEO.log_n).reshape(GEO.row_n, GEO.col_n, 4)
        colmaj = np.ascontiguousarray(rowmaj.transpose(1, 0, 2)).reshape(-1, 4)
        ftau = O.eval_interp(colmaj, GEO.log_n, tau)
        assert O.point_to_ints(exp.commitment) == O.point_to_ints(O.g1_mul_gen(ftau)), length
        high = O.g1_mul_gen(ftau * pow(tau, 3 * GEO.n, O.R_MOD))
        assert all(np.array_equal(exp.high[64 * c:64 * c + 64], high) for c in range(GEO.coset_n)), length
        canon = O.fr_canonical_array(rowmaj.reshape(-1, 4)).view(np.uint8).reshape(-1)
        assert np.array_equal(exp.data[:GEO.n * 32], canon), length
        assert file_root_ref(exp.data) == bytes(exp.root), length
        blobs.append({
            "length": length,
            "raw_sha256": hashlib.sha256(raw).hexdigest(),
            "commitment": bytes(exp.commitment).hex(),
            "high_commitment": bytes(exp.high[:64]).hex(),
            "file_root": bytes(exp.root).hex(),
            "proof_tree_first_points": [bytes(np.asarray(exp.proof_tree).reshape(-1)[64 * i:64 * i + 64]).hex()
                                        for i in range(4)],
            "proof_tree_sha256": sha(exp.proof_tree),
            "comm_tree_sha256": sha(exp.comm_tree),
            "merkle_sha256": sha(exp.merkle),
            "data_sha256": sha(exp.data),
        })
        print("length %d: commitment %s..., file root %s" % (length, blobs[-1]["commitment"][:16],
                                                             blobs[-1]["file_root"]), flush=True)
    doc = {
        "what": "EncodedBlob outputs at the test geometry (commitment, high commitment, file root as "
                "little-endian canonical bytes; the AMT trees, Merkle layers and encoded data as SHA-256 "
                "of the library's output arrays)",
        "geometry": {"log_col": GEO.log_col, "log_row": GEO.log_row, "cosets": GEO.coset_n},
        "tau": "numpy default_rng(%d).integers(0, 256, 64) little-endian mod r" % TAU_SEED,
        "data": "numpy default_rng(%d).integers(0, 256, length) bytes" % DATA_SEED,
        "blobs": blobs,
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
