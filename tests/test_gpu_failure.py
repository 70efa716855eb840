"""Error paths forced from the host (no kernel is made to fail): request cancellation in
the server, and failures injected at the issue of a blob (zgr_test_inject_fault) in a
batch and in the server.  After each, the context keeps encoding bit-exact outputs.

The reference has neither: its requests run to completion and its GPU errors panic
(grpc/src/service.rs:48-65, crates/amt/src/prove_params/prove_gpu.rs:48,84)."""
import time

import numpy as np
import pytest

import zg_rocm as Z
from helpers import DATA_SEED, seeded_tau

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING
FIELDS = ("commitment", "file_root", "proof_tree", "comm_tree", "high", "merkle", "data")


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    c.params_generate(seeded_tau())
    yield c
    c.close()


@pytest.fixture(scope="module")
def blobs(ctx):
    """eight raw blobs and their plain single-call encodings (pinned elsewhere)"""
    raws = [Z.seeded_bytes(DATA_SEED + 40 + i, GEO.max_raw_data_size - 977 * i) for i in range(8)]
    return raws, [ctx.encode(r) for r in raws]


def same(a, b):
    return all(np.array_equal(getattr(a, f), getattr(b, f)) for f in FIELDS)


def drain(srv, tickets, timeout=60):
    """ticket -> outputs or the ZgrError it completed with"""
    res, deadline = {}, time.time() + timeout
    while len(res) < len(tickets) and time.time() < deadline:
        for t in tickets:
            if t in res:
                continue
            try:
                o = srv.poll(t, 10)
            except Z.ZgrError as e:
                res[t] = e
                continue
            if o is not None:
                res[t] = o
    assert len(res) == len(tickets)
    return res


def test_cancel_before_the_gpu(ctx, blobs):
    raws, exp = blobs
    # one request admitted at a time: the later ones wait in the queue
    srv = Z.Server(ctx, hbm_budget=2 * GEO.max_blob_size, max_inflight=1)
    try:
        tickets = [srv.submit(r) for r in raws]
        cancelled = {t: srv.cancel(t) for t in tickets[4:]}
        res = drain(srv, tickets)
        st = srv.stats()
        assert st["completed"] == len(tickets) and st["queued"] == 0 and st["admitted"] == 0
        with pytest.raises(Z.ZgrError) as e:  # a collected ticket is no longer pending
            srv.cancel(tickets[0])
        assert e.value.code == Z.ERR_INVALID_ARG
    finally:
        srv.close()
    assert any(cancelled.values())
    for i, t in enumerate(tickets):
        if cancelled.get(t):
            assert isinstance(res[t], Z.ZgrError) and res[t].code == Z.ERR_CANCELLED, i
        else:
            assert not isinstance(res[t], Exception) and same(res[t], exp[i]), i


def test_injected_fault_in_a_batch(ctx, blobs):
    raws, exp = blobs
    dev = [Z.DeviceBuffer(r) for r in raws[:4]]
    ptrs, lens = [d.address for d in dev], [len(r) for r in raws[:4]]
    outs = [Z.EncodedBlobOutputs(GEO) for _ in range(4)]
    ctx.inject_fault(Z.FAULT_ISSUE, 3)  # the third blob's issue fails, the second is in flight
    with pytest.raises(Z.ZgrError) as e:
        ctx.encode_batch_device(ptrs, lens, outs)
    assert e.value.code == Z.ERR_DEVICE and "injected" in e.value.msg
    # the fault fired once; nothing of the failed call leaks into the next ones
    for i in range(4):
        assert same(ctx.encode(raws[i]), exp[i]), i
    outs = [Z.EncodedBlobOutputs(GEO) for _ in range(4)]
    ctx.encode_batch_device(ptrs, lens, outs)
    for i in range(4):
        assert same(outs[i], exp[i]), i


def test_injected_fault_in_the_server(ctx, blobs):
    raws, exp = blobs
    ctx.inject_fault(Z.FAULT_ISSUE, 2)
    srv = Z.Server(ctx, hbm_budget=4 * GEO.max_blob_size, max_inflight=1)  # issue order = submit order
    try:
        tickets = [srv.submit(r) for r in raws[:5]]
        res = drain(srv, tickets)
    finally:
        srv.close()
    for i, t in enumerate(tickets):
        if i == 1:
            assert isinstance(res[t], Z.ZgrError) and res[t].code == Z.ERR_DEVICE and "injected" in res[t].msg
        else:
            assert not isinstance(res[t], Exception) and same(res[t], exp[i]), i


def test_inject_fault_arguments(ctx, blobs):
    raws, exp = blobs
    for site, countdown in ((-1, 1), (Z.FAULT_ISSUE + 1, 1), (Z.FAULT_ISSUE, -1)):
        with pytest.raises(Z.ZgrError) as e:
            ctx.inject_fault(site, countdown)
        assert e.value.code == Z.ERR_INVALID_ARG
    ctx.inject_fault(Z.FAULT_ISSUE, 1)
    ctx.inject_fault(Z.FAULT_ISSUE, 0)  # disarmed again
    assert same(ctx.encode(raws[0]), exp[0])
