"""Throughput mode: many independent transpositions at once on one GPU.

The reference's --mpi harness runs the generated transposition in several
processes at the same time and reports their combined bandwidth
(ttc/transposeGenerator.py:459-460, 653-654; ttc/ttc.py:927-929); its flag
is accepted but has no effect on the GPU path there.  Here the same question
— what do concurrent, independent transpositions achieve together on one
device — is asked in two ways:

  streams:    N HIP streams in one process, each with its own buffers,
              launches interleaved round robin (`run_streams`);
  processes:  N processes sharing the GPU, each with its own buffers and
              stream, started together behind a barrier (`run_processes`).

Both report the aggregate GB/s (algorithmic bytes, ttc/transposeGenerator.py:
665-668) and transposes/s over the whole region, beside the same plan on one
stream, and check every output against the operation's definition on a
sample of positions.

  python -m ttc_amd.throughput --perm=1,0,2 --size=1000,768,16 --dataType=s --beta=1 --streams=1,2,4,8
  python -m ttc_amd.throughput --perm=1,0 --size=8192,8192 --processes=4
"""
import argparse
import json
import math
import multiprocessing as mp
import time

from . import runtime as rt
from . import ttc_util


def _buffers(torch, type_a, type_b, perm, size, dev, gen):
    na = int(math.prod(size))
    A = (torch.rand(na, generator=gen, device=dev, dtype=torch.float64) * 2 - 1).to(rt.torch_dtype(type_a))
    B = (torch.rand(na, generator=gen, device=dev, dtype=torch.float64) * 2 - 1).to(rt.torch_dtype(type_b))
    return A, B


def _sample_check(torch, plan, A, B_after, B_before_sample, idx_a, idx_b, alpha, beta):
    """B_after at sampled output positions against alpha*A + beta*B_before,
    evaluated in the types of the operation (float32 / float64, real types)."""
    a = A[idx_a].double()
    b0 = B_before_sample.double()
    if plan.type_a == rt.S:
        x = (a.float() * torch.tensor(alpha, dtype=torch.float32, device=A.device))
        if beta != 0:
            x = x + b0.float() * torch.tensor(beta, dtype=torch.float32, device=A.device)
        want = x.to(B_after.dtype)
    else:
        x = a * alpha
        if beta != 0:
            x = x + b0 * beta
        want = x.to(B_after.dtype)
    return bool(torch.equal(B_after[idx_b], want))


def _sample_positions(torch, perm, size, n, dev, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.integers(0, s, n) for s in size])
    oa = np.zeros(n, np.int64)
    m = 1
    for k, s in enumerate(size):
        oa += idx[k] * m
        m *= s
    ob = np.zeros(n, np.int64)
    m = 1
    for q in range(len(size)):
        ob += idx[perm[q]] * m
        m *= size[perm[q]]
    return torch.from_numpy(oa).to(dev), torch.from_numpy(ob).to(dev)


def run_streams(type_a, type_b, perm, size, alpha=1.0, beta=0.0, streams=(1, 2, 4, 8), reps=20, config=None):
    """Aggregate throughput of `n` concurrent streams for each n in
    `streams`; every stream transposes its own A into its own B, `reps`
    times, launches interleaved across streams.  Real types (s, d) only:
    the sample check evaluates the operation in torch."""
    import torch
    if type_a not in (rt.S, rt.D) or type_a != type_b:
        raise ValueError("throughput mode: same-type s or d transpositions")
    dev = torch.device("cuda:0")
    plan = rt.Plan(type_a, type_b, perm, size, beta_is_zero=beta == 0, config=config)
    nbytes = plan.info()["bytes_algorithmic"]
    gen = torch.Generator(device=dev)
    gen.manual_seed(17)
    nmax = max(streams)
    bufs = [_buffers(torch, type_a, type_b, perm, size, dev, gen) for _ in range(nmax)]
    ss = [torch.cuda.Stream(device=dev) for _ in range(nmax)]
    ia, ib = _sample_positions(torch, perm, size, 4096, dev, 3)
    results = []
    for n in streams:
        # correctness: one launch per stream from a known B
        before = [B[ib].clone() for _, B in bufs[:n]]
        torch.cuda.synchronize()
        for (A, B), s in zip(bufs[:n], ss):
            plan.execute(A, B, alpha, beta, stream=s)
        torch.cuda.synchronize()
        ok = all(_sample_check(torch, plan, A, B, b0, ia, ib, alpha, beta)
                 for (A, B), b0 in zip(bufs[:n], before))
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(n):
            starts[i].record(ss[i])
        for _ in range(reps):
            for (A, B), s in zip(bufs[:n], ss):
                plan.execute(A, B, alpha, beta, stream=s)
        for i in range(n):
            ends[i].record(ss[i])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # the region spans the earliest start to the latest end over all
        # streams (event times relative to stream 0's start)
        t_start = min(0.0, min(starts[0].elapsed_time(starts[i]) for i in range(n)))
        t_end = max(starts[0].elapsed_time(ends[i]) for i in range(n))
        span_ms = t_end - t_start
        total = n * reps
        results.append({"streams": n, "transposes": total, "ms": round(span_ms, 4),
                        "GBps": round(total * nbytes / (span_ms * 1e-3) / 1e9, 1),
                        "transposes_per_s": round(total / (span_ms * 1e-3), 1), "verified": ok,
                        "wall_s": round(wall, 3)})
    return {"perm": list(perm), "size": list(size), "type": rt.DTYPE_NAMES[type_a], "alpha": alpha, "beta": beta,
            "bytes_per_transpose": nbytes, "tile": plan.info()["tile"], "results": results}


def _worker(args):
    """One process of the processes mode: its own buffers and stream; waits
    at the barrier, then times `reps` launches with wall-clock bounds that
    the parent combines."""
    (type_a, type_b, perm, size, alpha, beta, reps, seed, barrier, q) = args
    import torch
    dev = torch.device("cuda:0")
    plan = rt.Plan(type_a, type_b, perm, size, beta_is_zero=beta == 0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    A, B = _buffers(torch, type_a, type_b, perm, size, dev, gen)
    ia, ib = _sample_positions(torch, perm, size, 4096, dev, seed)
    s = torch.cuda.Stream(device=dev)
    before = B[ib].clone()
    plan.execute(A, B, alpha, beta, stream=s)
    torch.cuda.synchronize()
    ok = _sample_check(torch, plan, A, B, before, ia, ib, alpha, beta)
    barrier.wait()
    t0 = time.time()
    for _ in range(reps):
        plan.execute(A, B, alpha, beta, stream=s)
    s.synchronize()
    t1 = time.time()
    q.put((t0, t1, ok, plan.info()["bytes_algorithmic"]))


def run_processes(type_a, type_b, perm, size, alpha=1.0, beta=0.0, processes=4, reps=50):
    """Aggregate throughput of `processes` processes sharing the GPU
    (spawned; each initialises its own HIP context)."""
    if processes < 1 or processes > 16:
        raise ValueError("processes must be 1..16 (one GPU)")
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(processes)
    q = ctx.Queue()
    procs = [ctx.Process(target=_worker, args=((type_a, type_b, list(perm), list(size), alpha, beta, reps, 100 + i,
                                                barrier, q),)) for i in range(processes)]
    for p in procs:
        p.start()
    got = [q.get(timeout=600) for _ in procs]
    for p in procs:
        p.join(timeout=60)
    if any(p.exitcode != 0 for p in procs):
        raise RuntimeError("a throughput worker failed (exit codes %s)" % [p.exitcode for p in procs])
    t0 = min(g[0] for g in got)
    t1 = max(g[1] for g in got)
    nbytes = got[0][3]
    total = processes * reps
    return {"perm": list(perm), "size": list(size), "type": rt.DTYPE_NAMES[type_a], "processes": processes,
            "transposes": total, "wall_s": round(t1 - t0, 4),
            "GBps": round(total * nbytes / (t1 - t0) / 1e9, 1), "transposes_per_s": round(total / (t1 - t0), 1),
            "verified": all(g[2] for g in got)}


def main(argv=None):
    ap = argparse.ArgumentParser(description="concurrent transpositions on one GPU (the --mpi harness's question)")
    ap.add_argument("--perm", required=True)
    ap.add_argument("--size", required=True)
    ap.add_argument("--dataType", default="s", help="s or d")
    ap.add_argument("--alpha", type=float, default=1.0)
    ap.add_argument("--beta", type=float, default=0.0)
    ap.add_argument("--streams", default="1,2,4,8", help="comma-separated stream counts")
    ap.add_argument("--processes", type=int, default=0, help="processes sharing the GPU (0: streams mode)")
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    perm = [int(x) for x in a.perm.split(",")]
    size = [int(x) for x in a.size.split(",")]
    ta, tb = (ttc_util.dtypeCode(t) for t in ttc_util.dataTypeFromString(a.dataType))
    if a.processes:
        out = run_processes(ta, tb, perm, size, a.alpha, a.beta, a.processes, a.reps)
    else:
        out = run_streams(ta, tb, perm, size, a.alpha, a.beta, [int(x) for x in a.streams.split(",")], a.reps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
