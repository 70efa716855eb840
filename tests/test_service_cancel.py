"""EncoderService's withdrawal of a served request whose caller has gone (the gRPC
handler passes its context's is_active), on the host with a stand-in server: the
request is cancelled while it waits, and runs to completion once it is on the GPU."""
import pytest

import zg_rocm as Z

GEO = Z.TEST_SETTING


class StandInServer:
    """Plays zgr_server: poll() pends `pending_polls` times, cancel() succeeds while
    `queued`."""

    def __init__(self, queued, pending_polls=3):
        self.queued, self.left = queued, pending_polls
        self.cancelled, self.timeouts = [], []

    def submit(self, data, want_data):
        return 41

    def poll(self, ticket, timeout_ms=-1):
        self.timeouts.append(timeout_ms)
        if self.cancelled:
            raise Z.ZgrError(Z.ERR_CANCELLED, "cancelled")
        if self.left:
            self.left -= 1
            return None
        return "outputs"

    def cancel(self, ticket):
        if self.queued:
            self.cancelled.append(ticket)
            return True
        return False


class StandInCtx:
    geo = GEO


def test_withdrawn_while_queued():
    srv = StandInServer(queued=True)
    svc = Z.EncoderService(StandInCtx(), srv)
    with pytest.raises(Z.RequestCancelled):
        svc._build(b"abc", False, alive=lambda: False)
    assert srv.cancelled == [41]


def test_runs_on_once_issued():
    srv = StandInServer(queued=False)
    svc = Z.EncoderService(StandInCtx(), srv)
    blob = svc._build(b"abc", False, alive=lambda: False)
    assert blob.out == "outputs" and srv.cancelled == []
    assert all(t == Z.EncoderService.POLL_MS for t in srv.timeouts)


def test_no_alive_callback_waits_without_timeouts():
    srv = StandInServer(queued=True, pending_polls=0)
    svc = Z.EncoderService(StandInCtx(), srv)
    assert svc._build(b"abc", False).out == "outputs"
    assert srv.timeouts == [-1]
