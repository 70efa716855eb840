"""C++ host mirror (0g-da-encoder_amd/host/zg_encoder.hpp) driven through its example
program zg_encode_file: params from a known tau and from bamt files
(AMTParams::from_dir_mont, crates/amt/src/prove_params/generate.rs:27-65; layout
fast_serde_bn254.rs:20-53), EncoderService::process_data replies and
EncodedBlob::get_row slices, all checked against the oracle.  The expected slices
are indexed directly from the oracle's arrays (blob.rs:44-49, encode.rs:185-200,
proofs.rs:16-31, merkle/blob.rs:44-57) and the Merkle proofs are verified to the
file root with the textbook Keccak (merkle/slice.rs:51-103)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, OracleOutputs, bamt_image, oracle_params, seeded_tau
from pyref import keccak256_ref
from zg_rocm import wire

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "..", "0g-da-encoder_amd", "lib", "zg_encode_file")
GEO = Z.TEST_SETTING


def _run(args, timeout=120):
    if not os.path.exists(BIN):
        pytest.fail("zg_encode_file is not built (0g-da-encoder_amd/build.py)")
    return subprocess.run([BIN] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          timeout=timeout)


def _geo_args():
    return [GEO.log_col, GEO.log_row, GEO.coset_n]


def write_bamt(path, bases_u64, log_n, h):
    with open(path, "wb") as f:
        f.write(bamt_image(bases_u64, log_n, h))


def expected_row(e, index):
    """EncodedBlob::get_row serialised as zg_encode_file writes it, from oracle arrays."""
    g = GEO
    coset, local = divmod(index, g.row_n)
    data = e.data.reshape(g.coset_n, g.row_n, g.col_n, 32)
    out = [data[coset, local].tobytes()]
    rev = int(format(local, "0%db" % g.log_row)[::-1], 2)
    ct = e.comm_tree.reshape(g.coset_n, -1, 64)
    pt = e.proof_tree.reshape(g.coset_n, -1, 64)
    for d in range(g.log_row):
        at = (2 << d) - 2 + ((rev >> (g.log_row - 1 - d)) ^ 1)
        out += [ct[coset, at].tobytes(), pt[coset, at].tobytes()]
    out += [e.high.reshape(-1, 64)[coset].tobytes(), e.commitment.tobytes()]
    m = e.merkle.reshape(-1, 32)
    layer = lambda d: m[sum(g.coset_n << k for k in range(d)):][:g.coset_n << d]  # noqa: E731
    out += [layer(0).tobytes()]
    out += [layer(d)[(index >> (g.log_row - d)) ^ 1].tobytes() for d in range(g.log_row, 0, -1)]
    out += [layer(g.log_row)[index].tobytes()]
    return b"".join(out)


def verify_merkle_slice(blob_row, index, file_root):
    """EncodedSliceMerkle::verify: row -> row root -> coset root -> file root."""
    g = GEO
    row_len = g.col_n * 32
    row = blob_row[:row_len]
    off = row_len + g.log_row * 128 + 128
    roots = [blob_row[off + 32 * i:off + 32 * i + 32] for i in range(g.coset_n)]
    off += 32 * g.coset_n
    proof = [blob_row[off + 32 * i:off + 32 * i + 32] for i in range(g.log_row)]
    leaf = blob_row[off + 32 * g.log_row:off + 32 * g.log_row + 32]
    layer = [keccak256_ref(row[256 * i:256 * i + 256]) for i in range(g.col_n // 8)]
    while len(layer) > 1:
        layer = [keccak256_ref(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
    assert layer[0] == leaf
    pos, acc = index, leaf
    for sib in proof:
        acc = keccak256_ref(sib + acc if pos % 2 else acc + sib)
        pos >>= 1
    assert acc == roots[index // g.row_n]
    fr = keccak256_ref(keccak256_ref(roots[0] + roots[1]) + roots[2])
    assert fr == file_root


@pytest.fixture(scope="module")
def raws(tmp_path_factory):
    d = tmp_path_factory.mktemp("raw")
    lengths = [0, 1234, GEO.max_raw_data_size]
    paths, data = [], []
    for i, L in enumerate(lengths):
        r = Z.seeded_bytes(DATA_SEED + 40 + i, L)
        p = d / ("%d.bin" % i)
        p.write_bytes(r)
        paths.append(str(p))
        data.append(r)
    return paths, data


ROWS = [0, GEO.row_n - 1, GEO.row_n + 5, GEO.row_encoded - 1]


def _check_outputs(outdir, data, require=None):
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    for k, r in enumerate(data):
        exp = O.encode(params, GEO.log_col, GEO.log_row, r)
        # the protobuf reply against the Python codec over the oracle's outputs
        blob = Z.EncodedBlob(GEO, OracleOutputs(exp))
        req = True if require is None else require[k]
        want_pb = wire.encode_reply({"version": 0, "erasure_commitment": exp.commitment.tobytes(),
                                     "storage_root": exp.root.tobytes(),
                                     "encoded_data": exp.data.tobytes() if req else b"",
                                     "encoded_slice": blob.encoded_slices()})
        assert (outdir / ("%d.pb" % k)).read_bytes() == want_pb
        reply = (outdir / ("%d.reply" % k)).read_bytes()
        assert reply[:64] == exp.commitment.tobytes()
        assert reply[64:96] == exp.root.tobytes()
        assert reply[96:] == exp.data.tobytes()
        # encoded_slice: the C++ serializer against the Python one over the oracle's outputs
        blob = Z.EncodedBlob(GEO, OracleOutputs(exp))
        want = b"".join(blob.encoded_slice(i) for i in range(GEO.row_encoded))
        assert (outdir / ("%d.slices" % k)).read_bytes() == want
        rows = (outdir / ("%d.rows" % k)).read_bytes()
        per = len(rows) // len(ROWS)
        for j, idx in enumerate(ROWS):
            got = rows[j * per:(j + 1) * per]
            assert got == expected_row(exp, idx), "row %d of blob %d" % (idx, k)
            verify_merkle_slice(got, idx, exp.root.tobytes())
        light = (outdir / ("%d.light" % k)).read_bytes()
        assert light == b"".join(blob.light_slice(idx) for idx in ROWS)


@pytest.mark.gpu
def test_host_mirror_from_tau(tmp_path, raws):
    paths, data = raws
    tau_hex = (seeded_tau() % O.R_MOD).to_bytes(32, "little").hex()
    res = _run(_geo_args() + ["tau:" + tau_hex, tmp_path] + paths + ["--rows", ",".join(map(str, ROWS))])
    assert res.returncode == 0, res.stderr.decode()
    _check_outputs(tmp_path, data)


@pytest.mark.gpu
def test_host_mirror_from_bamt_dir(tmp_path, raws):
    paths, data = raws
    pdir = tmp_path / "params"
    pdir.mkdir()
    out = tmp_path / "out"
    out.mkdir()
    params = oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n)
    for c in range(GEO.coset_n):
        name = "amt-prove-coset%d-mont-5DWgDV-%02d-%02d.bin" % (c, GEO.log_row, GEO.log_n)
        write_bamt(str(pdir / name), params[c], GEO.log_n, GEO.log_row)
    res = _run(_geo_args() + ["dir:%s" % pdir, out] + paths + ["--rows", ",".join(map(str, ROWS))])
    assert res.returncode == 0, res.stderr.decode()
    _check_outputs(out, data)


@pytest.mark.gpu
def test_host_mirror_protobuf_requests(tmp_path, raws):
    """EncodeBlobRequest files in, EncodeBlobReply protobuf out (require_data both ways)."""
    _, data = raws
    reqs = []
    require = [k % 2 == 0 for k in range(len(data))]
    for k, r in enumerate(data):
        p = tmp_path / ("%d.req" % k)
        p.write_bytes(wire.encode_request(r, require[k]))
        reqs.append(str(p))
    out = tmp_path / "out"
    out.mkdir()
    tau_hex = (seeded_tau() % O.R_MOD).to_bytes(32, "little").hex()
    res = _run(_geo_args() + ["tau:" + tau_hex, out] + reqs + ["--rows", ",".join(map(str, ROWS)), "--pb"])
    assert res.returncode == 0, res.stderr.decode()
    _check_outputs(out, data, require)


@pytest.mark.gpu
def test_host_mirror_too_large(tmp_path):
    p = tmp_path / "big.bin"
    p.write_bytes(bytes(GEO.max_raw_data_size + 1))
    tau_hex = (seeded_tau() % O.R_MOD).to_bytes(32, "little").hex()
    res = _run(_geo_args() + ["tau:" + tau_hex, tmp_path, p])
    assert res.returncode == 3
    assert b"TooLargeBlob" in res.stderr


def test_oracle_slices_verify():
    """The slice layout above, read from the oracle's own outputs, passes the
    reference's Merkle verification (no device work)."""
    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    exp = O.encode(params, GEO.log_col, GEO.log_row, Z.seeded_bytes(DATA_SEED + 41, 1234))
    for idx in ROWS:
        verify_merkle_slice(expected_row(exp, idx), idx, exp.root.tobytes())
    with pytest.raises(AssertionError):  # a sibling from the wrong side fails
        verify_merkle_slice(expected_row(exp, 1), 0, exp.root.tobytes())


def test_host_mirror_usage_and_bad_params(tmp_path):
    """Host-side argument and file errors (no device work)."""
    assert _run([]).returncode == 2
    res = _run(_geo_args() + ["tau:zz", tmp_path, tmp_path / "x"])
    assert res.returncode == 4


@pytest.mark.gpu
def test_host_mirror_recovery(tmp_path, raws):
    """zg::recover_from_da_slice (C++ mirror of crates/recovery/src/lib.rs:15-43) over
    rows of the blob's own encoding: the raw bytes back; too few rows fail with the
    reference's text."""
    paths, data = raws
    tau_hex = (seeded_tau() % O.R_MOD).to_bytes(32, "little").hex()
    ids = list(range(GEO.row_n + 7, 2 * GEO.row_n + 7)) + [GEO.row_encoded - 1]
    out = tmp_path / "ok"
    out.mkdir()
    res = _run(_geo_args() + ["tau:" + tau_hex, out] + paths + ["--recover", ",".join(map(str, ids))])
    assert res.returncode == 0, res.stderr.decode()
    for k, r in enumerate(data):
        assert (out / ("%d.recovered" % k)).read_bytes() == r
    few = tmp_path / "few"
    few.mkdir()
    res = _run(_geo_args() + ["tau:" + tau_hex, few, paths[1], "--recover", ",".join(map(str, ids[:GEO.row_n - 1]))])
    assert res.returncode == 5 and b"Not enough valid lines" in res.stderr
