"""GPU: TTC's search moved on device.  The autotuner's candidates all
compute the same bits, its fastest configuration is the one emitted, and
generateTransposition (ttc/ttc.py:614-1130) measures, caches in the SQLite
database and answers a repeated request from the cache."""
import os

import pytest

from common import S, D
from test_gpu_parity import check

pytestmark = pytest.mark.gpu


def test_autotune_candidates_agree_and_are_ranked(gpu):
    from ttc_amd import autotune
    perm, size = [2, 0, 1], [64, 48, 80]
    results = autotune.tune(S, S, perm, size, alpha=1.1, beta=0.5, max_impl=14, reps=2)
    assert len(results) >= 5
    assert [r["ms"] for r in results] == sorted(r["ms"] for r in results)
    assert all(r["GBs"] > 0 for r in results)
    fams = {r["info"]["family"] for r in results}
    assert fams == {1, 2}
    for r in results:
        check(gpu, S, S, perm, size, 1.1, 0.5, config=r["config"], seed=3)


def test_generate_transposition_measures_caches_and_emits(gpu, tmp_path, monkeypatch):
    from ttc_amd import ttc, ttc_util
    monkeypatch.setenv("TTC_ROOT", str(tmp_path))
    monkeypatch.chdir(tmp_path)

    def args():
        a = ttc_util.TTCargs([1, 0, 2], [200, 150, 8])
        a.floatTypeA = a.floatTypeB = "double"
        a.alpha, a.beta = 1.0, 1.0
        a.maxNumImplementations = 6
        a.silent = 1
        return a
    name, gibs = ttc.generateTransposition(args())
    assert name == "dCuTranspose_102_200x150x8" and gibs > 0
    assert os.path.exists(os.path.join(str(tmp_path), "ttc", "ttc.db"))
    assert os.path.exists(os.path.join(str(tmp_path), "ttc_transpositions", name + ".hip"))
    name2, gibs2 = ttc.generateTransposition(args())  # found in the database: no new measurement
    assert name2 == name and gibs2 == pytest.approx(gibs)
