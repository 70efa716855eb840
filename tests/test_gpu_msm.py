"""zgr_msm (one MSM of arbitrary bases, the isolated-primitive MSM of 2^20 points at
production size) on both of its paths -- GLV halves through the wide sort and bucket
kernels with one bucket group per 16-bit window (the default) and the per-chunk windowed
kernels (ZGR_MSM=windowed) -- against the oracle's Pippenger (oracle/zg_oracle.c), on
scalars chosen for the signed-digit edge cases."""
import os

import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import seeded_tau
from test_gpu_parity import _canon_to_mont_bases, _random_points

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING
R = O.R_MOD


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(GEO)
    c.params_generate(seeded_tau())
    yield c
    c.close()


def scalars_mont(vals):
    return O.fr_to_mont_array([v % R for v in vals])


def edge_scalars(rng, n):
    """uniform values, then digits at the window edges: 2^15, 2^15 + 1 and 2^16 - 1 in a
    window (carries into the next), r - 1, zero, small, and repeats (shared buckets)"""
    vals = [int.from_bytes(rng.bytes(32), "little") % R for _ in range(n)]
    special = [0, 1, 2, (1 << 15), (1 << 15) + 1, (1 << 16) - 1, (1 << 16), R - 1, R - 2,
               sum((1 << 15) << (16 * w) for w in range(15)), sum(((1 << 16) - 1) << (16 * w) for w in range(15))]
    for i, v in enumerate(special[:n]):
        vals[(i * 7) % n] = v % R
    if n > 20:
        vals[20] = vals[21]
    return vals


def run_both(ctx, bases, scal):
    got = ctx.msm(bases, scal)
    os.environ["ZGR_MSM"] = "windowed"
    try:
        win = ctx.msm(bases, scal)
    finally:
        del os.environ["ZGR_MSM"]
    return got, win


@pytest.mark.parametrize("n", [1, 2, 11, 257, 1500, 2048])
def test_msm_paths_match_oracle(ctx, n):
    rng = np.random.default_rng(700 + n)
    bases = _canon_to_mont_bases(_random_points(rng, n, with_inf=False))
    vals = edge_scalars(rng, n)
    exp = O.msm(bases, np.array([O.int_to_limbs(v) for v in vals]))
    got, win = run_both(ctx, bases, scalars_mont(vals))
    assert np.array_equal(got, exp)
    assert np.array_equal(win, exp)


def test_msm_all_zero_and_single_point(ctx):
    rng = np.random.default_rng(5)
    bases = _canon_to_mont_bases(_random_points(rng, 64, with_inf=False))
    got, win = run_both(ctx, bases, scalars_mont([0] * 64))
    assert not got.any() and not win.any()  # the point at infinity
    # one base times k is [k] P
    for k in (1, 3, (1 << 16) - 1, R - 1):
        s = [0] * 64
        s[17] = k
        got, _ = run_both(ctx, bases, scalars_mont(s))
        exp = O.msm(bases[17:18], np.array([O.int_to_limbs(k)]))
        assert np.array_equal(got, exp), k


@pytest.mark.parametrize("n", [5, 300])
def test_msm_bases_at_infinity(ctx, n):
    """An all-zero base -- which the param files never hold (write_g1 refuses the point at
    infinity, crates/amt/src/prove_params/fast_serde_bn254.rs:139-142) -- is read as the
    point at infinity on either path: the result is the MSM over the other bases."""
    rng = np.random.default_rng(1300 + n)
    bases = _canon_to_mont_bases(_random_points(rng, n, with_inf=False))
    bases[0::3] = 0
    vals = edge_scalars(rng, n)
    keep = [i for i in range(n) if i % 3]
    exp = O.msm(bases[keep], np.array([O.int_to_limbs(vals[i]) for i in keep]))
    got, win = run_both(ctx, bases, scalars_mont(vals))
    assert np.array_equal(got, exp)
    assert np.array_equal(win, exp)
    # every base at infinity
    got, win = run_both(ctx, np.zeros_like(bases), scalars_mont(vals))
    assert not got.any() and not win.any()


@pytest.mark.parametrize("segs", ["1", "3", "8192"])
def test_msm_segment_counts(ctx, segs):
    """ZGR_MSM_SEGS: one sums wave for everything, an odd count, and more waves than
    (bucket, window) runs (most segments empty) give the same point"""
    n = 1500
    rng = np.random.default_rng(1700)
    bases = _canon_to_mont_bases(_random_points(rng, n, with_inf=False))
    vals = edge_scalars(rng, n)
    exp = O.msm(bases, np.array([O.int_to_limbs(v) for v in vals]))
    os.environ["ZGR_MSM_SEGS"] = segs
    try:
        got = ctx.msm(bases, scalars_mont(vals))
    finally:
        del os.environ["ZGR_MSM_SEGS"]
    assert np.array_equal(got, exp)


def test_msm_repeated_calls_and_sizes(ctx):
    """the path's buffers are reused across calls of different sizes"""
    for n in (2048, 3, 2048, 700):
        rng = np.random.default_rng(900 + n)
        bases = _canon_to_mont_bases(_random_points(rng, n, with_inf=False))
        vals = [int.from_bytes(rng.bytes(32), "little") % R for _ in range(n)]
        exp = O.msm(bases, np.array([O.int_to_limbs(v) for v in vals]))
        assert np.array_equal(ctx.msm(bases, scalars_mont(vals)), exp), n
