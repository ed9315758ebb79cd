"""GPU parity on the committed golden fixtures: the HIP path (both kernel
families where the shape allows the row x column one) and the fp8
cast-transpose reproduce the fixtures bit for bit."""
import numpy as np
import pytest

from common import GOLDEN_CODES, first_diff, load_golden, same_bits
from test_gpu_parity import families, run_hip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("code", GOLDEN_CODES)
def test_hip_matches_fixture(gpu, code):
    meta, arr = load_golden(code)
    tA, tB = meta["type_a"], meta["type_b"]
    for n, c in enumerate(meta["cases"]):
        want = arr["out%d" % n]
        for fam in families(tA, tB, c["perm"], c["size"], c["lda"], c["ldb"]):
            got, info = run_hip(gpu, tA, tB, arr["A%d" % n], arr["B%d" % n], c["alpha"], c["beta"], c["perm"],
                                c["size"], c["lda"], c["ldb"], config=dict(family=fam))
            assert same_bits(got, want), (code, n, c, fam, first_diff(got, want))


def test_fp8_matches_fixture(gpu):
    import torch
    from common import to_device
    from ttc_amd import runtime as rt
    meta, arr = load_golden("fp8")
    amax = torch.zeros(1, dtype=torch.float32, device=gpu)
    for n, c in enumerate(meta["cases"]):
        want = arr["out%d" % n]
        dA = to_device(arr["A%d" % n], gpu)
        dB = torch.zeros(want.size, dtype=torch.uint8, device=gpu)
        plan = rt.Plan(c["type_a"], c["type_b"], c["perm"], c["size"], beta_is_zero=True)
        plan.execute_fp8(dA, dB, c["alpha"], scale=c["scale"], amax_out=amax)
        torch.cuda.synchronize()
        assert same_bits(dB.cpu().numpy(), want), (n, c)
        assert float(amax.item()) == c["amax"], (n, c)
