"""GPU checks of the throughput mode (concurrent independent
transpositions on one GPU, the question of the reference's --mpi harness,
ttc/transposeGenerator.py:459-460, 653-654): every stream's and every
process's output is checked on sampled positions, and the aggregate figures
are consistent with the work done."""
import pytest

pytestmark = pytest.mark.gpu


def test_streams_mode(gpu):
    from ttc_amd import runtime as rt
    from ttc_amd import throughput
    out = throughput.run_streams(rt.S, rt.S, [1, 0, 2], [1000, 768, 16], alpha=1.5, beta=0.5, streams=(1, 2, 4),
                                 reps=5)
    assert [r["streams"] for r in out["results"]] == [1, 2, 4]
    for r in out["results"]:
        assert r["verified"], r
        assert r["transposes"] == 5 * r["streams"] and r["GBps"] > 0
        assert abs(r["transposes_per_s"] * out["bytes_per_transpose"] / 1e9 - r["GBps"]) < 0.01 * r["GBps"] + 1
    out = throughput.run_streams(rt.D, rt.D, [2, 0, 1], [64, 300, 200], beta=0.0, streams=(3,), reps=3)
    assert out["results"][0]["verified"]


def test_processes_mode(gpu):
    from ttc_amd import runtime as rt
    from ttc_amd import throughput
    out = throughput.run_processes(rt.S, rt.S, [1, 0], [2048, 1536], alpha=1.0, beta=1.0, processes=2, reps=10)
    assert out["verified"] and out["transposes"] == 20 and out["GBps"] > 0
