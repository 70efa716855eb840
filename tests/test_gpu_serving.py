"""Serving layer (zgr_server): requests submitted while others are on the GPU are
admitted by staging-arena footprint, overlap on the device, and each comes out
bit-exact with the same blob encoded alone."""
import time

import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import DATA_SEED, oracle_params, seeded_tau

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING
FIELDS = ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data")


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    c.params_generate(seeded_tau())
    yield c
    c.close()


def test_served_mix_matches_oracle(ctx):
    params = oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n)
    rng = np.random.default_rng(5)
    lengths = [int(x) for x in rng.integers(0, GEO.max_raw_data_size + 1, 12)] + [0, GEO.max_raw_data_size]
    raws = [Z.seeded_bytes(DATA_SEED + 10 + i, L) for i, L in enumerate(lengths)]
    # a tight arena (two full blobs) and few in flight force admission to wait
    srv = Z.Server(ctx, hbm_budget=2 * GEO.max_blob_size, max_inflight=3)
    try:
        tickets = [srv.submit(r) for r in raws]
        outs = {}
        deadline = time.time() + 60
        while len(outs) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t not in outs:
                    o = srv.poll(t, 10)
                    if o is not None:
                        outs[t] = o
        assert len(outs) == len(tickets)
        st = srv.stats()
        assert st["completed"] == len(tickets) and st["queued"] == 0 and st["admitted"] == 0
    finally:
        srv.close()
    for r, t in zip(raws, tickets):
        exp = O.encode(list(params), GEO.log_col, GEO.log_row, r)
        got = outs[t]
        for a, b in ((got.commitment, exp.commitment), (got.file_root, exp.root), (got.proof_tree, exp.proof_tree),
                     (got.comm_tree, exp.comm_tree), (got.high, exp.high), (got.merkle, exp.merkle),
                     (got.data, exp.data)):
            assert np.array_equal(a, b)


def test_server_admission_errors(ctx):
    srv = Z.Server(ctx, hbm_budget=1000, max_inflight=2)
    try:
        with pytest.raises(Z.ZgrError) as e:
            srv.submit(b"\x01" * (GEO.max_raw_data_size + 1))
        assert e.value.code == Z.ERR_TOO_LARGE
        with pytest.raises(Z.ZgrError) as e:
            srv.submit(b"\x01" * 5000)  # can never fit the 1000-byte arena
        assert e.value.code == Z.ERR_OOM
        t = srv.submit(b"\x02" * 100)
        assert srv.poll(t) is not None
        # a collected ticket is unknown afterwards
        assert Z.lib().zgr_server_poll(srv.h, t, 0) == Z.ERR_INVALID_ARG
    finally:
        srv.close()
    # the context is usable again once the server is gone
    out = ctx.encode(b"abc")
    assert out.commitment.any()


def test_grpc_front_end_on_device(ctx):
    """encoder.Encoder/EncodeBlob over loopback, served from the device through
    zgr_server by concurrent gRPC workers; every reply equals the oracle's."""
    from concurrent.futures import ThreadPoolExecutor

    import grpc

    from helpers import OracleOutputs
    from zg_rocm import grpc_front, wire

    params = list(oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n))
    lengths = [0, 1, 1234, 40000, GEO.max_raw_data_size, 777]
    raws = [Z.seeded_bytes(DATA_SEED + 90 + i, L) for i, L in enumerate(lengths)]
    srv = Z.Server(ctx, hbm_budget=3 * GEO.max_blob_size, max_inflight=4)
    server, port = grpc_front.serve(Z.EncoderService(ctx, srv), "127.0.0.1:0", max_workers=4)
    try:
        with grpc_front.EncoderClient("127.0.0.1:%d" % port) as client:
            with ThreadPoolExecutor(4) as pool:
                replies = list(pool.map(lambda kr: client.encode_blob_raw(kr[1], kr[0] % 2 == 0, timeout=60),
                                        enumerate(raws)))
            with pytest.raises(grpc.RpcError) as ei:
                client.encode_blob(bytes(GEO.max_raw_data_size + 1), timeout=60)
            assert ei.value.code() == grpc.StatusCode.UNKNOWN
            assert ei.value.details().startswith("TooLargeBlob { actual: %d" % (GEO.max_raw_data_size + 1))
    finally:
        server.stop(0)
        srv.close()
    for k, (r, got) in enumerate(zip(raws, replies)):
        exp = O.encode(params, GEO.log_col, GEO.log_row, r)
        want = wire.encode_reply({"version": 0, "erasure_commitment": exp.commitment.tobytes(),
                                  "storage_root": exp.root.tobytes(),
                                  "encoded_data": exp.data.tobytes() if k % 2 == 0 else b"",
                                  "encoded_slice": Z.EncodedBlob(GEO, OracleOutputs(exp)).encoded_slices()})
        assert got == want, "request %d (length %d)" % (k, len(r))
