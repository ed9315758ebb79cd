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


def test_autotune_fp8_candidates_agree(gpu):
    """The search also covers fp8 cast-transposes (timed with delayed
    scaling); every candidate reproduces the oracle's bits and amax."""
    import numpy as np
    import torch
    from common import make_case, same_bits, to_device
    from oracle import oracle as orc
    from ttc_amd import autotune
    from ttc_amd import runtime as rt
    perm, size = [1, 0], [512, 448]
    results = autotune.tune(S, rt.E4M3, perm, size, alpha=1.0, beta=0.0, max_impl=8, reps=2)
    assert len(results) >= 3 and {r["info"]["family"] for r in results} >= {1, 2}
    A, _ = make_case(S, S, perm, size, seed=5, dist="heavy")
    want = np.zeros(A.size, np.uint8)
    want_amax = orc.cast_fp8(S, rt.E4M3, A, want, 0.75, 2.0, perm, size)
    dA = to_device(A, gpu)
    for r in results:
        plan = rt.Plan(S, rt.E4M3, perm, size, beta_is_zero=True, config=r["config"] or None)
        dB = torch.zeros(A.size, dtype=torch.uint8, device=gpu)
        amax = torch.zeros(1, dtype=torch.float32, device=gpu)
        plan.execute_fp8(dA, dB, 0.75, scale=2.0, amax_out=amax)
        torch.cuda.synchronize()
        assert same_bits(dB.cpu().numpy(), want), r["config"]
        assert float(amax.item()) == want_amax


def test_hot_cache_modes_time_and_key_the_database(gpu, tmp_path, monkeypatch):
    """--hotA/--hotB: the search times with A (B) reused between launches and
    stores the result under its own database key."""
    from ttc_amd import autotune, sql_util, ttc, ttc_util
    import sqlite3
    for hot_a, hot_b in ((True, False), (False, True), (True, True)):
        res = autotune.tune(S, S, [1, 0], [300, 200], alpha=1.0, beta=1.0, max_impl=3, reps=2, hot_a=hot_a,
                            hot_b=hot_b)
        assert res and all(r["GBs"] > 0 for r in res)
    monkeypatch.setenv("TTC_ROOT", str(tmp_path))
    monkeypatch.chdir(tmp_path)
    for hot in (0, 1):
        a = ttc_util.TTCargs([1, 0], [300, 200])
        a.floatTypeA = a.floatTypeB = "float"
        a.alpha, a.beta = 1.0, 1.0
        a.maxNumImplementations = 3
        a.silent = 1
        a.hotA = hot
        ttc.generateTransposition(a)
    con = sqlite3.connect(os.path.join(str(tmp_path), "ttc", "ttc.db"))
    assert sorted(r[0] for r in con.execute("SELECT hotA FROM measurements")) == [0, 1]
    con.close()
