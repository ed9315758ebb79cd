"""GPU parity at the full sizes of the benchmark's two throughput workloads.

Plain transpositions at bandwidth scale: the 16 seeded random permutations of
rank 2..6 at 200 MiB of A, in fp32 and in fp64, with beta = 0 (B not read)
and beta = 1, alpha = 1; case 15 extracts a padded sub-tensor.  Every op is
checked over its WHOLE B buffer, bit for bit, against the oracle run on the
same inputs on the host — that covers every output element and, for case
15, that the padding of B is left untouched (the oracle never writes it).
Each configuration the autotuner may pick for the op (the candidates it
times) is then run at full size and compared with that verified result on
the device.

Batched grouped transpositions: the 512 mixed ops (ranks 2..5, 4 KiB to
64 MiB of A, fp32 / bf16 / fp64, random perms, alpha in [0.5, 2], beta in
{0, 1}) through the grouped launch, its hipGraph replay and the per-op
loop, with the planner's and with the batch tuner's per-op configurations;
every op's whole B buffer is compared with the oracle.  As in the
reference's test driver (test-framework/ttc_test.py:124-156) one failing
op fails the test, and the message names it.
"""
import numpy as np
import pytest

from common import BF16, D, S, extents, oracle_run, same_bits

pytestmark = pytest.mark.gpu

C2_TYPES = {4: S, 8: D}
NP_TYPES = {S: np.float32, D: np.float64, BF16: np.uint16}


def _c2_case(k, es):
    from ttc_amd import workloads as wl
    case = wl.plain_c2()[k]
    size = case["sizes"][es]
    lda = case["lda"][es] if case["lda"] else None
    ldb = case["ldb"][es] if case["ldb"] else None
    return case["perm"], size, lda, ldb


def _diff_report(got, want):
    g = got.view(np.uint8).reshape(-1, got.itemsize)
    w = want.view(np.uint8).reshape(-1, want.itemsize)
    bad = np.nonzero((g != w).any(axis=1))[0]
    return "%d of %d elements differ, first at %s" % (len(bad), len(got), bad[:8].tolist())


@pytest.mark.parametrize("es", [4, 8])
@pytest.mark.parametrize("k", range(16))
def test_c2_full_size_vs_oracle(gpu, k, es):
    import torch
    from ttc_amd import autotune
    from ttc_amd import runtime as rt
    perm, size, lda, ldb = _c2_case(k, es)
    t = C2_TYPES[es]
    na, nb = extents(perm, size, lda, ldb)
    rng = np.random.default_rng(1000 + k)
    hA = rng.uniform(-1.0, 1.0, na).astype(NP_TYPES[t])
    hB0 = rng.uniform(-1.0, 1.0, nb).astype(NP_TYPES[t])
    dA = torch.from_numpy(hA).to(gpu)
    dB0 = torch.from_numpy(hB0).to(gpu)
    dB = torch.empty_like(dB0)
    print("c2 case %d es=%d perm=%s size=%s lda=%s ldb=%s" % (k, es, perm, size, lda, ldb), flush=True)
    for bz in (True, False):
        beta = 0.0 if bz else 1.0
        want = oracle_run(t, t, hA, hB0, 1.0, beta, perm, size, lda, ldb)
        dB.copy_(dB0)
        plan = rt.Plan(t, t, perm, size, lda, ldb, beta_is_zero=bz)
        plan.execute(dA, dB, 1.0, beta)
        torch.cuda.synchronize()
        got = dB.cpu().numpy()
        assert same_bits(got, want), ("case %d es %d beta %g planner %s: %s"
                                      % (k, es, beta, plan.info()["tile"], _diff_report(got, want)))
        # every candidate the autotuner times for this op, against the verified result
        dwant = torch.from_numpy(want).to(gpu)
        cands = autotune.candidates(t, t, perm, size, lda, ldb, bz, max_impl=12, nt_stores=bz)
        for cfg in cands[1:]:
            cp = rt.Plan(t, t, perm, size, lda, ldb, beta_is_zero=bz, config=cfg)
            dB.copy_(dB0)
            cp.execute(dA, dB, 1.0, beta)
            torch.cuda.synchronize()
            assert torch.equal(dB.view(torch.uint8), dwant.view(torch.uint8)), (
                "case %d es %d beta %g candidate %s (%s)" % (k, es, beta, cfg, cp.info()["tile"]))
        del dwant


def _host(torch, t):
    """Flat host copy in the oracle's storage type (bf16 as uint16 bits)."""
    if t.dtype == torch.bfloat16:
        return t.cpu().view(torch.int16).numpy().view(np.uint16)
    return t.cpu().numpy()


def _c5_ops(torch, gpu):
    """The 512 ops with device buffers drawn as the benchmark draws them,
    host copies of A and of the original B, and the oracle's result."""
    from ttc_amd import runtime as rt
    from ttc_amd import workloads as wl
    gen = torch.Generator(device=gpu)
    gen.manual_seed(5)
    tmap = {0: (S, torch.float32), 1: (D, torch.float64), 4: (BF16, torch.bfloat16)}
    ops = []
    for o in wl.batch_c5():
        t, tdt = tmap[o["type"]]
        n = int(np.prod(o["size"]))
        A = (torch.rand(n, generator=gen, device=gpu, dtype=torch.float32) * 2 - 1).to(tdt)
        B = (torch.rand(n, generator=gen, device=gpu, dtype=torch.float32) * 2 - 1).to(tdt)
        ops.append({"type_a": t, "type_b": t, "perm": o["perm"], "size": o["size"], "A": A, "B": B,
                    "B0": B.clone(), "alpha": o["alpha"], "beta": o["beta"]})
    torch.cuda.synchronize()
    for o in ops:
        hA = _host(torch, o["A"])
        hB = _host(torch, o["B0"])
        want = oracle_run(o["type_a"], o["type_b"], hA, hB, o["alpha"], o["beta"], o["perm"], o["size"])
        o["want"] = torch.from_numpy(want).to(gpu)
    assert len(ops) == 512 and rt.DTYPE_BYTES[S] == 4
    return ops


def _check_c5(torch, ops, label):
    torch.cuda.synchronize()
    bad = []
    for i, o in enumerate(ops):
        if not torch.equal(o["B"].view(torch.uint8), o["want"].view(torch.uint8)):
            bad.append((i, o["type_a"], o["perm"], o["size"], o["alpha"], o["beta"]))
    assert not bad, "%s: %d of %d ops differ from the oracle, first %s" % (label, len(bad), len(ops), bad[:4])


def _restore(ops):
    for o in ops:
        o["B"].copy_(o["B0"])


def test_c5_batch_full_size_vs_oracle(gpu):
    import torch
    from ttc_amd import autotune
    from ttc_amd import runtime as rt
    ops = _c5_ops(torch, gpu)
    stream = torch.cuda.Stream()
    print("c5: 512 ops, planner configurations", flush=True)
    batch = rt.Batch(ops)
    assert batch.info()["bytes"] > 0
    _restore(ops)
    batch.execute(stream)
    stream.synchronize()
    _check_c5(torch, ops, "grouped launch")
    for replay in range(3):  # each replay from the original B, eager work in between
        _restore(ops)
        torch.cuda.synchronize()
        batch.graph_launch(stream)
        stream.synchronize()
        _check_c5(torch, ops, "graph replay %d" % replay)
        torch.randn(1 << 20, device=gpu).mul_(2.0).sum().item()
    _restore(ops)
    torch.cuda.synchronize()
    batch.execute_loop(stream)
    stream.synchronize()
    _check_c5(torch, ops, "per-op loop")
    batch.close()
    # the batch tuner's per-op configurations (every problem of >= 1 MiB)
    print("c5: batch tuner configurations", flush=True)
    cfgs, tuned = autotune.tune_batch(ops, max_impl=3, reps=1)
    assert tuned > 100
    tb = rt.Batch([dict(o, config=c) for o, c in zip(ops, cfgs)])
    for replay in range(2):
        _restore(ops)
        torch.cuda.synchronize()
        tb.graph_launch(stream)
        stream.synchronize()
        _check_c5(torch, ops, "tuned graph replay %d" % replay)
    _restore(ops)
    torch.cuda.synchronize()
    tb.execute(stream)
    stream.synchronize()
    _check_c5(torch, ops, "tuned grouped launch")
    tb.close()
