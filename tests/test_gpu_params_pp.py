"""AMTParams::from_pp on the device (zgr_params_from_pp: coset scaling and the EC-FFTs
of gen_basis / gen_quotients, crates/amt/src/prove_params/generate.rs:139-232) from
powers of tau, at the test geometry: bit-exact with the oracle's closed form for every
coset, and an encode over those params equals the oracle's."""
import numpy as np
import pytest

import oracle as O
import zg_rocm as Z
from helpers import CPU_THREADS, DATA_SEED, OracleOutputs, oracle_params, seeded_tau

pytestmark = pytest.mark.gpu

GEO = Z.TEST_SETTING


@pytest.fixture(scope="module")
def powers():
    O.set_threads(CPU_THREADS)
    tau = seeded_tau()
    return O.power_tau(tau, 0, GEO.log_n), O.power_tau(tau, 3 * GEO.n, GEO.log_n)


def test_from_pp_equals_closed_form(powers):
    g1pp, high = powers
    params = oracle_params(GEO.log_col, GEO.log_row, GEO.coset_n)
    ctx = Z.Context(GEO)
    try:
        for c in range(GEO.coset_n):
            ctx.params_from_pp(c, g1pp, high)
            got = ctx.params_download(c).reshape(GEO.h + 2, GEO.n, 8)
            want = np.ascontiguousarray(params[c], dtype=np.uint64).reshape(GEO.h + 2, GEO.n, 8)
            bad = [L for L in range(GEO.h + 2) if not np.array_equal(got[L], want[L])]
            assert not bad, "coset %d, lines %s differ" % (c, bad)
        raw = Z.seeded_bytes(DATA_SEED + 700, 5000)
        out = ctx.encode(raw)
        exp = O.encode(list(params), GEO.log_col, GEO.log_row, raw)
        assert bytes(out.commitment) == exp.commitment.tobytes()
        assert np.array_equal(out.proof_tree, exp.proof_tree) and np.array_equal(out.data, exp.data)
    finally:
        ctx.close()


def test_from_pp_argument_errors(powers):
    g1pp, high = powers
    ctx = Z.Context(GEO)
    try:
        with pytest.raises(Z.ZgrError):
            ctx.params_from_pp(0, g1pp[:64 * (GEO.n // 2)], high[:64 * (GEO.n // 2)])
        with pytest.raises(Z.ZgrError):
            ctx.params_from_pp(GEO.coset_n, g1pp, high)
    finally:
        ctx.close()
