"""The committed golden fixtures (tests/golden, written by make_golden.py
from the oracle): the oracle must keep reproducing them bit for bit, and
their cases must cover every type pair and beta class."""
import numpy as np
import pytest

from common import GOLDEN_CODES, load_golden, oracle_run, same_bits
from oracle import oracle as orc


@pytest.mark.parametrize("code", GOLDEN_CODES)
def test_oracle_reproduces_fixture(oracle_lib, code):
    meta, arr = load_golden(code)
    tA, tB = meta["type_a"], meta["type_b"]
    betas = set()
    for n, c in enumerate(meta["cases"]):
        out = oracle_run(tA, tB, arr["A%d" % n], arr["B%d" % n], c["alpha"], c["beta"], c["perm"], c["size"],
                         c["lda"], c["ldb"])
        assert same_bits(out, arr["out%d" % n]), (code, n, c)
        betas.add(c["beta"])
    assert {0.0, 1.2, -1.0} <= betas and len(meta["cases"]) >= 12


def test_oracle_reproduces_fp8_fixture(oracle_lib):
    meta, arr = load_golden("fp8")
    for n, c in enumerate(meta["cases"]):
        out = np.zeros_like(arr["out%d" % n])
        amax = orc.cast_fp8(c["type_a"], c["type_b"], arr["A%d" % n], out, c["alpha"], c["scale"], c["perm"],
                            c["size"])
        assert same_bits(out, arr["out%d" % n]) and amax == c["amax"], (n, c)
    assert len(meta["cases"]) == 36


def test_fixture_files_are_small():
    import os
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for f in os.listdir(d):
        if f.endswith(".npz"):
            assert os.path.getsize(os.path.join(d, f)) <= 64 * 1024, f
