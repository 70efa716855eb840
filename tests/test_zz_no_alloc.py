"""zgr_merkle_noalloc (the Merkle entry point that allocates nothing: its layers live in
a fixed device scratch array) against zgr_merkle on the same data, through the C ABI:
two calls in a row on one context, the second with a smaller input, so that the
scratch array's contents from the first call cannot leak into the second's result."""
import ctypes

import numpy as np
import pytest

import zg_rocm as Z

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING  # coset_n = 3: layers halve down to 3 roots


def _noalloc(ctx, data, levels):
    n = data.size // 32
    top = sum(GEO.coset_n << k for k in range(levels + 1))
    src = Z.DeviceBuffer(data.tobytes())
    layers = Z.DeviceBuffer(bytes(top * 32))
    root = Z.DeviceBuffer(bytes(32))
    Z._check(Z.lib().zgr_merkle_noalloc(ctx.h, ctypes.c_void_p(src.address), n, levels,
                                        ctypes.c_void_p(layers.address), ctypes.c_void_p(root.address)))
    return np.frombuffer(layers.to_host(), dtype=np.uint8).reshape(-1, 32), root.to_host()


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    yield c
    c.close()


def test_noalloc_equals_merkle_twice(ctx):
    rng = np.random.default_rng(707)
    # a production blob's encoded data (3 x 2^20 elements), then a much smaller input
    for n_elems, levels in ((3 << 20, 10), (3 << 12, 6)):
        data = rng.integers(0, 256, n_elems * 32, dtype=np.uint8)
        want_layers, want_root = ctx.merkle(data, levels)
        got_layers, got_root = _noalloc(ctx, data, levels)
        assert got_root == want_root, n_elems
        assert np.array_equal(got_layers, want_layers), n_elems


def test_noalloc_rejects_oversize(ctx):
    # about 8 bytes of layers per element: 4 Mi elements fit the 32 MiB scratch, these do not
    too_many = (32 << 20) // 8 + 3 * 8 * 1024
    src = Z.DeviceBuffer(bytes(32))  # never read: the size check comes first
    out = Z.DeviceBuffer(bytes(32))
    rc = Z.lib().zgr_merkle_noalloc(ctx.h, ctypes.c_void_p(src.address), too_many, 0, ctypes.c_void_p(out.address),
                                    ctypes.c_void_p(out.address))
    assert rc == Z.ERR_INVALID_ARG
    assert b"scratch" in Z.lib().zgr_last_error()
