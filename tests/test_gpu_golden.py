"""The HIP encode path against the committed encode fixtures at the test geometry
(tests/golden/encode_test_geometry.json): the commitment, high commitment, file root, AMT
trees, Merkle layers and encoded data of four blobs, through the C ABI, on the default
plan and on 12-bit windows everywhere."""
import hashlib

import pytest

import zg_rocm as Z
from test_golden import check_against_golden, load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["default", "narrow_only"])
def ctx(request):
    c = Z.make_context(Z.TEST_SETTING, wide=request.param == "default")
    c.params_generate(Z.tau_from_seed(7))
    yield c
    c.close()


@pytest.mark.parametrize("i", range(4))
def test_encode_matches_golden(ctx, i):
    g = load_golden()["blobs"][i]
    raw = Z.seeded_bytes(222, g["length"])
    assert hashlib.sha256(raw).hexdigest() == g["raw_sha256"]
    out = ctx.encode(raw)
    check_against_golden(g, out.commitment, out.high, out.file_root, out.proof_tree, out.comm_tree, out.merkle,
                         out.data)
