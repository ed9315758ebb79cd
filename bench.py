#!/usr/bin/env python3
"""Benchmark of the MI355X transposition path.

Workload (one "step"): the 16 seeded random-permutation cases of rank 2..6,
each at 200 MiB of A, in fp32 and in fp64, with beta = 0 (B not read) and
beta = 1 (B read) — 64 out-of-place transpositions
B[pi(i)] = alpha*A[i] + beta*B[pi(i)], inputs resident in HBM.  All beta = 0
ops run first, then all beta = 1 ops, so consecutive ops never touch the same
buffers and no op finds its data in the 256 MB last-level cache.

Plans: by default each op is autotuned first (at most 200 candidates, TTC's
default number of implementations, timed on its own buffers outside the
timed region), as TTC generates and measures variants before use
(ttc/ttc.py:871-1022); the planner's untuned choice is timed in the same run
and reported beside it.

Metric: GB/s (10^9 B/s) of algorithmic traffic, N*sA + N*sB (+ N*sB when
beta != 0) (ttc/transposeGenerator.py:665-668), over the whole timed region,
plus a roofline line for the dominant kernel and the oracle's CPU rate on a
bounded sample.  `--batch` adds the grouped-launch comparison (512 mixed ops,
one persistent launch, hipGraph replay) in transposes/s.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
  python bench.py --calibrate-copy     (a plain 200 MiB copy, for counter calibration)

roofline.traffic comes from the committed counter summary
profiles/c2_hbm_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
benchmark, scaled by the copy calibration; see ttc_amd/profiling.py) when it
lists the dominant kernel, else null.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c2_hbm_traffic.json")
CALIB_BYTES = 200 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_name(info):
    """The kernel a plan launches, spelled as rocprof shows its template
    (ttc::<family kernel><Op, VA, VB, BETA, NT[, FAST]>)."""
    ops = {(0, 0): "ttc::OpF32<float, float>", (1, 1): "ttc::OpDD"}
    tf = lambda x: "true" if x else "false"
    op = ops.get((info["type_a"], info["type_b"]), "?")
    if info["family"] == 2:
        return "ttc::transpose_rc_kernel<%s, %d, %d, %s, %s, %s>" % (
            op, info["vec_a"], info["vec_b"], tf(not info["beta_is_zero"]), tf(info["nt_store"]),
            tf(info["lane_fixed"]))
    return "ttc::transpose_kernel<%s, %d, %d, %s, %s>" % (op, info["vec_a"], info["vec_b"],
                                                        tf(not info["beta_is_zero"]), tf(info["nt_store"]))


def build_c2(torch, rt, wl, dev, tune=0):
    """The 64 ops of one step.  tune > 0: each op's configuration comes from
    the autotuner (at most `tune` candidates timed on its own buffers before
    the timed region, TTC's generate-then-run flow, ttc/ttc.py:871-1022)."""
    ops = []
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    bufs = {}
    for case in wl.plain_c2():
        for es, t, tdt in ((4, rt.S, torch.float32), (8, rt.D, torch.float64)):
            size = case["sizes"][es]
            lda = case["lda"][es] if case["lda"] else None
            ldb = case["ldb"][es] if case["ldb"] else None
            r = len(size)
            LA = lda or size
            LB = ldb or [size[case["perm"][q]] for q in range(r)]
            na = 1
            for x in LA[:-1]:
                na *= x
            na *= size[-1]
            nb = 1
            for x in LB[:-1]:
                nb *= x
            nb *= size[case["perm"][-1]]
            A = torch.rand(na, generator=gen, device=dev, dtype=tdt) * 2 - 1
            B = torch.rand(nb, generator=gen, device=dev, dtype=tdt) * 2 - 1
            bufs[(case["k"], es)] = (A, B)
            for bz in (True, False):
                cfg = None
                if tune > 0:
                    from ttc_amd import autotune
                    res = autotune.tune(t, t, case["perm"], size, lda, ldb, 1.0, 0.0 if bz else 1.0, tune, reps=5)
                    cfg = res[0]["config"] if res else None
                plan = rt.Plan(t, t, case["perm"], size, lda, ldb, beta_is_zero=bz, config=cfg)
                plan0 = rt.Plan(t, t, case["perm"], size, lda, ldb, beta_is_zero=bz)
                info = plan.info()
                ops.append({"k": case["k"], "es": es, "bz": bz, "plan": plan, "plan0": plan0, "A": A, "B": B,
                            "alpha": 1.0, "beta": 0.0 if bz else 1.0, "perm": case["perm"], "size": size,
                            "lda": lda, "ldb": ldb, "bytes": info["bytes_algorithmic"], "info": info,
                            "kernel": kernel_name(info)})
    # beta = 0 ops first, then beta = 1 ops
    ops.sort(key=lambda o: (not o["bz"], o["k"], o["es"]))
    return ops, bufs


def sample_offsets(torch, op, nsample=4096):
    """Sampled element positions of an op: (A offsets, B offsets) as device
    index tensors (each output depends on exactly one input)."""
    import numpy as np
    rng = np.random.default_rng(op["k"] * 10 + op["es"])
    size, perm = op["size"], op["perm"]
    r = len(size)
    LA = op["lda"] or size
    LB = op["ldb"] or [size[perm[q]] for q in range(r)]
    idx = np.stack([rng.integers(0, s, nsample) for s in size])
    offA = np.zeros(nsample, np.int64)
    m = 1
    for k in range(r):
        offA += idx[k] * m
        m *= LA[k]
    offB = np.zeros(nsample, np.int64)
    m = 1
    for q in range(r):
        offB += idx[perm[q]] * m
        m *= LB[q]
    dev = op["A"].device
    return torch.from_numpy(offA).to(dev), torch.from_numpy(offB).to(dev)


def verify_pass(torch, ops, stream, key):
    """Runs every op once and checks sampled outputs against the element
    formula (Appendix-A semantics with alpha = 1: beta = 0 gives B = A,
    beta = 1 gives B = fl(A + B_old), one rounding in B's type).  The beta = 0
    ops run first; the beta = 1 ops then read the B they wrote.  Returns the
    number of ops whose sample differs."""
    bad = 0
    for bz in (True, False):
        sel = [o for o in ops if o["bz"] == bz]
        samples = [sample_offsets(torch, o) for o in sel]
        old = [o["B"][ib].clone() for o, (_, ib) in zip(sel, samples)]
        # the snapshots run on torch's current stream, the ops on `stream`
        torch.cuda.synchronize()
        for o in sel:
            o[key].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
        torch.cuda.synchronize()
        for o, (ia, ib), b0 in zip(sel, samples, old):
            a = o["A"][ia]
            want = a if bz else a + b0  # alpha = 1: fl(1*a) = a, fl(1*b) = b
            if not torch.equal(o["B"][ib], want):
                bad += 1
                log("  verification FAILED: k=%d %s beta=%g" % (o["k"], "f32" if o["es"] == 4 else "f64",
                                                                 o["beta"]))
    return bad


def run_c2(torch, ops, stream, steps, warmup, key="plan"):
    """Times `steps` passes over the ops with one pair of events around the
    whole region; nothing else is enqueued between the launches (an event
    recorded between two kernels adds a marker packet and costs several
    microseconds per op).  Before that, every op is checked once on
    sampled outputs (verify_pass)."""
    ok = verify_pass(torch, ops, stream, key) == 0
    for _ in range(warmup):
        for o in ops:
            o[key].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(steps):
        for o in ops:
            o[key].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    return t0.elapsed_time(t1), wall, ok


def copy_ceiling(torch, rt, dev, stream, rounds=10, nbuf=8):
    """The practical ceiling of this launch pattern: this library's identity
    transposition (fused to one row, the row-copy kernel, beta = 0) of
    200 MiB of fp32, launched back to back over `nbuf` distinct buffer pairs
    as the C2 step launches its ops, so that no launch finds its data in the
    256 MB last-level cache.  GB/s of read + write bytes."""
    n = (200 << 20) // 4
    bufs = [(torch.rand(n, device=dev), torch.empty(n, device=dev)) for _ in range(nbuf)]
    plan = rt.Plan(rt.S, rt.S, [0], [n], beta_is_zero=True)
    for A, B in bufs:
        plan.execute(A, B, 1.0, 0.0, stream=stream)
    torch.cuda.synchronize()
    ok = all(torch.equal(A, B) for A, B in bufs)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(rounds):
        for A, B in bufs:
            plan.execute(A, B, 1.0, 0.0, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / (rounds * nbuf)
    info = plan.info()
    del bufs
    torch.cuda.empty_cache()
    return {"GBps": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1), "avg_launch_ms": round(ms, 5), "bytes": 2 * n * 4,
            "launches": rounds * nbuf, "verified": ok, "kernel": "row copy (family %d), vector %d" % (
                info["family"], info["vec_a"])}


def per_op_ms(torch, ops, stream, reps):
    """[[ms of op i in rep r]]: each launch between its own pair of events
    (diagnostic: the events add a few microseconds to each launch)."""
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in ops]
          for _ in range(reps)]
    torch.cuda.synchronize()
    for r in range(reps):
        for i, o in enumerate(ops):
            ev[r][i][0].record(stream)
            o["plan"].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
            ev[r][i][1].record(stream)
    torch.cuda.synchronize()
    return [[ev[r][i][0].elapsed_time(ev[r][i][1]) for i in range(len(ops))] for r in range(reps)]


def kernel_avg_ms(torch, ops, stream, reps):
    """Average launch duration of one kernel: its ops launched in turn, `reps`
    rounds, with one pair of events around the whole region.  Consecutive
    launches touch different tensors, so no launch finds its data in cache."""
    for o in ops:
        o["plan"].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(reps):
        for o in ops:
            o["plan"].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
    t1.record(stream)
    torch.cuda.synchronize()
    n = reps * len(ops)
    return t0.elapsed_time(t1) / n, sum(o["bytes"] for o in ops) * reps // n


def host_threads():
    """CPU threads this process may use (its affinity mask; OMP_NUM_THREADS
    when set lower)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and 0 < int(env) < n:
        n = int(env)
    return n


def l3_bytes():
    """Total last-level cache of the host (distinct L3 instances in sysfs)."""
    seen, total = set(), 0
    base = "/sys/devices/system/cpu"
    try:
        for cpu in os.listdir(base):
            d = os.path.join(base, cpu, "cache", "index3")
            if not cpu.startswith("cpu") or not os.path.isdir(d):
                continue
            with open(os.path.join(d, "shared_cpu_list")) as f:
                key = f.read().strip()
            if key in seen:
                continue
            seen.add(key)
            with open(os.path.join(d, "size")) as f:
                v = f.read().strip()
            total += int(v[:-1]) * 1024 if v.endswith("K") else int(v)
    except (OSError, ValueError):
        return 0
    return total


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_oracle_rate(oracle, np, ops, nthreads, trash, reps=4, blocked=False):
    """Bytes and seconds over (tA, A, B, perm, size, beta) ops: min of
    `reps` runs of each, each after a cache-trash sweep
    (ttc/transposeGenerator.py:1493-1502, min of nRepeat as 452-837).
    blocked: the blocked, vectorised baseline instead of the oracle's
    scalar loop nest."""
    tot_bytes, tot_s = 0, 0.0
    for t, A, B, perm, size, beta in ops:
        best = None
        for _ in range(reps):
            oracle.trash_cache(trash, nthreads)
            t0 = time.perf_counter()
            if blocked:
                oracle.blocked_transpose(t, A, B, 1.0, beta, perm, size, nthreads=nthreads)
            else:
                oracle.transpose(t, t, A, B, 1.0, beta, perm, size, nthreads=nthreads)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        n = int(np.prod(size))
        tot_bytes += n * A.itemsize * (2 if beta == 0 else 3)
        tot_s += best
    return tot_bytes, tot_s


def cpu_baseline(budget_s):
    """The CPU path on the host threads this process may use, on a bounded
    sample of the same workload (C2 cases in fp32 and fp64, beta 0 and 1,
    min of 4 reps each after a cache-trash sweep of twice the host's L3, at
    least 512 MiB, until `budget_s` seconds are spent): TTC's blocked,
    vectorised CPU method (oracle/ttc_cpu_blocked.cpp, AVX-512 or AVX2
    micro-kernels; `value`) and, beside it, the oracle's scalar loop nest
    (the reference code the generator emits for checking)."""
    import numpy as np

    from oracle import oracle
    from ttc_amd import workloads as wl
    nthreads = host_threads()
    cases = wl.plain_c2()
    done, tot = [], {"blocked": [0, 0.0], "oracle": [0, 0.0]}
    trash_bytes = max(512 << 20, 2 * l3_bytes())
    trash = np.ones(trash_bytes // 8)
    t_start = time.perf_counter()
    order = [(k, es) for k in (2, 6, 11, 3, 0, 9, 14, 7, 12, 5, 1, 8, 4, 10, 13, 15) for es in (4, 8)]
    for k, es in order:
        if time.perf_counter() - t_start > budget_s:
            break
        case = cases[k]
        size, perm = case["sizes"][es], case["perm"]
        if case["lda"] or case["ldb"]:
            continue  # the padded case: dense extents only here
        n = int(np.prod(size))
        rng = np.random.default_rng(k)
        dt = np.float32 if es == 4 else np.float64
        A = rng.random(n, dtype=dt)
        B = rng.random(n, dtype=dt)
        t = oracle.S if es == 4 else oracle.D
        ops = [(t, A, B, perm, size, 0.0), (t, A, B, perm, size, 1.0)]
        for name in ("blocked", "oracle"):
            b, s_ = cpu_oracle_rate(oracle, np, ops, nthreads, trash, blocked=name == "blocked")
            tot[name][0] += b
            tot[name][1] += s_
        done.append("k%d/%s" % (k, "f32" if es == 4 else "f64"))
        del A, B
    isa = {2: "AVX-512", 1: "AVX2"}.get(oracle.blocked_isa(), "none")
    rate = {k: v[0] / v[1] / 1e9 for k, v in tot.items()}
    return {"value": round(rate["blocked"], 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "oracle_loop_nest_GBps": round(rate["oracle"], 3),
            "sample": "TTC's blocked CPU method (%s micro-kernels, OpenMP %d threads, %s) on C2 cases %s, beta 0 "
                      "and 1 each, min of 4 reps after a %d MiB cache-trash sweep (2x L3); the oracle's scalar loop "
                      "nest on the same ops: oracle_loop_nest_GBps" % (isa, nthreads, cpu_model(), ",".join(done),
                                                                       trash_bytes >> 20)}


def run_c1(torch, rt, wl, dev, stream, reps=20, nbuf=4):
    """The plain CPU-runnable case at full size (fp32, perm (1,0,2), size
    (1000,768,16), alpha = beta = 1): GPU rate over launches alternating
    between `nbuf` buffer pairs, and the oracle's rate on the host."""
    import numpy as np

    from oracle import oracle
    c1 = wl.plain_c1()
    perm, size = c1["perm"], c1["size"]
    n = int(np.prod(size))
    plan = rt.Plan(rt.S, rt.S, perm, size, beta_is_zero=False)
    nb = plan.info()["bytes_algorithmic"]
    bufs = [(torch.rand(n, device=dev) * 2 - 1, torch.rand(n, device=dev) * 2 - 1) for _ in range(nbuf)]
    for A, B in bufs:
        plan.execute(A, B, c1["alpha"], c1["beta"], stream=stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for r in range(reps):
        A, B = bufs[r % nbuf]
        plan.execute(A, B, c1["alpha"], c1["beta"], stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del bufs
    nthreads = host_threads()
    rng = np.random.default_rng(0)
    hA = rng.uniform(-1, 1, n).astype(np.float32)
    hB = rng.uniform(-1, 1, n).astype(np.float32)
    trash = np.ones(max(512 << 20, 2 * l3_bytes()) // 8)
    b, s_ = cpu_oracle_rate(oracle, np, [(oracle.S, hA, hB, perm, size, 1.0)], nthreads, trash)
    gpu = nb / (ms * 1e-3) / 1e9
    cpu = b / s_ / 1e9
    return {"ms": round(ms, 4), "GBps": round(gpu, 1), "frac": round(gpu / HBM_PEAK_GBS, 4),
            "tile": plan.info()["tile"], "cpu_oracle_GBps": round(cpu, 2), "cpu_threads": nthreads,
            "gpu_over_cpu": round(gpu / cpu, 1)}


def run_batch(torch, rt, wl, dev, stream, replays, tune=0):
    """512 mixed ops: the grouped persistent launch (graph replay) against
    plain per-op launch loops over the same ops and buffers.

    The plain loops are what a user writes with ttc_plan_create /
    ttc_execute: one plan per op with its own best single-op configuration
    (the planner's, or the autotuner's when tuning: every kernel family,
    row copy and two LDS buffers included), launched one after another on
    one stream — issued from Python, and captured once into a graph and
    replayed (no host time per launch).  The speed-up is quoted against the
    fastest of them.  Diagnostics: the loop over the batch's own plans
    (grouped kinds only) and one synchronous legacy call per op.
    tune > 0: every op of at least 1 MiB takes the autotuner's configuration
    (at most `tune` candidates) in the batch and in the plain loops."""
    ops = wl.batch_c5()
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    tmap = {0: (rt.S, torch.float32), 1: (rt.D, torch.float64), 4: (rt.BF16, torch.bfloat16)}
    desc = []
    for o in ops:
        t, tdt = tmap[o["type"]]
        n = 1
        for x in o["size"]:
            n *= x
        A = (torch.rand(n, generator=gen, device=dev, dtype=torch.float32) * 2 - 1).to(tdt)
        B = (torch.rand(n, generator=gen, device=dev, dtype=torch.float32) * 2 - 1).to(tdt)
        desc.append({"type_a": t, "type_b": t, "perm": o["perm"], "size": o["size"], "A": A, "B": B,
                     "alpha": o["alpha"], "beta": o["beta"]})
    batch = rt.Batch(desc)
    info = batch.info()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def plain_loops(cfgs):
        """(eager ms, graph-replay ms) of the per-op plan loop."""
        plans = [rt.Plan(d["type_a"], d["type_b"], d["perm"], d["size"], beta_is_zero=d["beta"] == 0, config=c)
                 for d, c in zip(desc, cfgs)]

        def loop():
            for p, d in zip(plans, desc):
                p.execute(d["A"], d["B"], d["alpha"], d["beta"], stream=stream)
        t_eager = timed(loop, max(5, replays // 100))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            loop()
            stream.synchronize()
            with torch.cuda.graph(g, stream=stream):
                loop()

        def replay():
            with torch.cuda.stream(stream):
                g.replay()
        t_graph = timed(replay, max(10, replays // 10))
        del g
        return t_eager, t_graph

    cfgs = [None] * len(desc)
    t_group = timed(lambda: batch.graph_launch(stream, per_op_loop=False), replays)
    t_group_planner, tuned = None, 0
    if tune > 0:
        from ttc_amd import autotune
        cfgs, tuned = autotune.tune_batch(desc, max_impl=tune)
        tb = rt.Batch([dict(d, config=c) for d, c in zip(desc, cfgs)])
        t_group_planner = t_group
        if timed(lambda: tb.graph_launch(stream, per_op_loop=False), replays // 4 + 1) < t_group:
            batch.close()
            batch, info = tb, tb.info()  # the search's configurations, timed afresh below
            t_group = timed(lambda: batch.graph_launch(stream, per_op_loop=False), replays)
        else:
            tb.close()
    t_group_eager = timed(lambda: batch.execute(stream), max(10, replays // 10))
    loops = {"planner": plain_loops([None] * len(desc))}
    if tune > 0:
        loops["tuned"] = plain_loops(cfgs)
    best_name, best_ms = min(((k + "_" + m, v[i]) for k, v in loops.items() for i, m in enumerate(("eager", "graph"))),
                             key=lambda x: x[1])
    t_loop_batch = timed(lambda: batch.execute_loop(stream), max(10, replays // 10))
    t_loop_batch_graph = timed(lambda: batch.graph_launch(stream, per_op_loop=True), max(10, replays // 10))

    # the reference wrapper's semantics: one synchronous call per op on the
    # null stream (ttc/CUDAtranspose.py:395-429), here without its per-call
    # metadata upload (plans are cached by shape)
    def legacy_loop():
        for d in desc:
            rt.legacy_transpose(d["type_a"], d["type_b"], d["A"], d["B"], d["alpha"], d["beta"], d["perm"],
                                d["size"], beta_is_zero=d["beta"] == 0)
    legacy_loop()
    torch.cuda.synchronize()
    reps = 3
    w0 = time.perf_counter()
    for _ in range(reps):
        legacy_loop()
    t_sync = (time.perf_counter() - w0) * 1e3 / reps
    n = len(ops)
    out = {"ops": n, "replays": replays, "tiles": info["tiles"], "workgroups": info["workgroups"],
           "bytes_per_launch": info["bytes"],
           "configs": ("autotuned (<= %d candidates) for the %d distinct problems of >= 1 MiB, planner for the "
                       "rest" % (tune, tuned)) if tune else "planner",
           "grouped_graph_ms": round(t_group, 4),
           "grouped_graph_ms_planner": round(t_group_planner, 4) if t_group_planner else None,
           "grouped_eager_ms": round(t_group_eager, 4),
           "plain_loop_ms": {k + "_" + m: round(v[i], 4) for k, v in loops.items()
                             for i, m in enumerate(("eager", "graph"))},
           "best_plain_loop": best_name,
           "transposes_per_s_grouped": round(n / (t_group * 1e-3), 1),
           "transposes_per_s_best_plain_loop": round(n / (best_ms * 1e-3), 1),
           "speedup_vs_best_plain_loop": round(best_ms / t_group, 3),
           "grouped_GBps": round(info["bytes"] / (t_group * 1e-3) / 1e9, 1),
           "diagnostic": {"loop_batch_plans_ms": round(t_loop_batch, 4),
                          "loop_batch_plans_graph_ms": round(t_loop_batch_graph, 4),
                          "loop_sync_legacy_ms": round(t_sync, 3),
                          "speedup_vs_sync_legacy_loop": round(t_sync / t_group, 3)}}
    return out


def run_workloads(torch, rt, wl, dev, stream, reps=5, tune=0):
    """Throughput of the mixed-precision / complex workload (perm (3,1,0,2),
    size (256,192,128,96), alpha 1.3, beta 0.7) and of the fp8
    cast-transposes (2-D 8192x14336 perm (1,0); 4-D (128,4096,32,8) perm
    (0,2,1,3)) with the planner's configurations.  `reps` launches between
    one pair of events; the fp8 launches alternate between two buffer pairs
    so that the working set exceeds the last-level cache.  tune > 0: each
    fp8 cast takes the autotuner's configuration (at most `tune`
    candidates) when it times faster than the planner's, whose rate is
    reported beside it.  GB/s counts
    algorithmic bytes (fp8: N*s_A + N; the current-scaling amax pass re-reads
    A, which only current_GBps_both_passes counts)."""
    import math

    def timed(fn, n=None):
        n = n or reps
        fn(0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for r in range(n):
            fn(r)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    out = {}
    c3 = wl.mixed_c3()
    n = math.prod(c3["size"])
    types = {"sb": (rt.S, rt.BF16, torch.float32, torch.bfloat16), "ds": (rt.D, rt.S, torch.float64, torch.float32),
             "c": (rt.C, rt.C, torch.complex64, torch.complex64), "z": (rt.Z, rt.Z, torch.complex128, torch.complex128)}
    for code in c3["dtypes"]:
        ta, tb, da, db = types[code]
        A = torch.randn(n, device=dev, dtype=da)
        B = torch.randn(n, device=dev, dtype=db)
        plan = rt.Plan(ta, tb, c3["perm"], c3["size"], beta_is_zero=False)
        ms = timed(lambda r: plan.execute(A, B, c3["alpha"], c3["beta"], stream=stream))
        nb = plan.info()["bytes_algorithmic"]
        out["mixed_" + code] = {"ms": round(ms, 4), "GBps": round(nb / (ms * 1e-3) / 1e9, 1),
                                "frac": round(nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        del A, B, plan
        torch.cuda.empty_cache()
    for shape in wl.fp8_c4():
        n = math.prod(shape["size"])
        for ta, da, an in ((rt.S, torch.float32, "f32"), (rt.BF16, torch.bfloat16, "bf16")):
            As = [torch.randn(n, device=dev, dtype=da) for _ in range(2)]
            for tb, bn in ((rt.E4M3, "e4m3"), (rt.E5M2, "e5m2")):
                Bs = [torch.empty(n, dtype=torch.uint8, device=dev) for _ in range(2)]
                amax = torch.zeros(1, dtype=torch.float32, device=dev)
                sinv = torch.zeros(1, dtype=torch.float32, device=dev)
                plan = rt.Plan(ta, tb, shape["perm"], shape["size"], beta_is_zero=True)
                nb = n * (rt.DTYPE_BYTES[ta] + 1)
                res = {}
                if tune > 0:  # the planner's choice, then the autotuner's (delayed scaling)
                    from ttc_amd import autotune
                    ms0 = timed(lambda r: plan.execute_fp8(As[r % 2], Bs[r % 2], 1.0, scale=1.0, stream=stream), 20)
                    res["planner_delayed_GBps"] = round(nb / (ms0 * 1e-3) / 1e9, 1)
                    best = autotune.tune(ta, tb, shape["perm"], shape["size"], alpha=1.0, beta=0.0, max_impl=tune,
                                         reps=3)
                    if best and best[0]["config"]:
                        cand = rt.Plan(ta, tb, shape["perm"], shape["size"], beta_is_zero=True,
                                       config=best[0]["config"])
                        # keep what is faster here; the figures below time it afresh
                        if timed(lambda r: cand.execute_fp8(As[r % 2], Bs[r % 2], 1.0, scale=1.0,
                                                            stream=stream), 20) < ms0:
                            plan = cand
                    res["plan"] = {k: plan.info()[k] for k in ("family", "tile", "vec_a", "vec_b")}
                for mode, mname in ((rt.SCALE_DELAYED, "delayed"), (rt.SCALE_CURRENT, "current")):
                    ms = timed(lambda r: plan.execute_fp8(As[r % 2], Bs[r % 2], 1.0, scale=1.0, amax_out=amax,
                                                          scale_inv_out=sinv, mode=mode, stream=stream), 20)
                    res[mname + "_ms"] = round(ms, 4)
                    res[mname + "_GBps"] = round(nb / (ms * 1e-3) / 1e9, 1)
                # both passes of current scaling: the amax pass reads A once more
                res["current_GBps_both_passes"] = round((nb + n * rt.DTYPE_BYTES[ta]) / (res["current_ms"] * 1e-3) / 1e9, 1)
                res["frac_delayed"] = round(res["delayed_GBps"] / HBM_PEAK_GBS, 4)
                out["fp8_%s_%s_%s" % (an, bn, "x".join(str(x) for x in shape["size"]))] = res
                del Bs, plan
            del As
            torch.cuda.empty_cache()
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    k = data.get("kernels", {}).get(kernel)
    return k["traffic_bytes"] if k else None


def calibrate_copy(torch):
    """A plain device copy of CALIB_BYTES, three times (the reference stream
    for the counter scale factors)."""
    dev = torch.device("cuda:0")
    x = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": CALIB_BYTES, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--batch", dest="batch", action="store_true", default=True)
    ap.add_argument("--no-batch", dest="batch", action="store_false")
    ap.add_argument("--batch-replays", type=int, default=1000)
    ap.add_argument("--no-workloads", dest="workloads", action="store_false", default=True,
                    help="skip the mixed-precision / complex and fp8 cast-transpose throughput leg")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only the plain copy used to scale counters")
    ap.add_argument("--tune", type=int, default=200,
                    help="autotune each op over at most N candidates before timing, as TTC generates the fastest "
                         "variant before use; 200 is TTC's own default (ttc/ttc.py:1160); 0: the planner's "
                         "choice only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus %d is not supported (a transposition shards as "
              "independent sub-transpositions, one slab of B's outermost dimension per GPU, with no "
              "exchange)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from ttc_amd import runtime as rt
    from ttc_amd import workloads as wl
    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        calibrate_copy(torch)
        return
    dev = torch.device("cuda:0")
    rt.lib()
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        ops, bufs = build_c2(torch, rt, wl, dev, args.tune)
    torch.cuda.synchronize()
    log("c2: %d ops per step" % len(ops))
    value_planner, ok0 = None, True
    if args.tune > 0:  # the planner's own choice, for comparison
        t0_ms, _, ok0 = run_c2(torch, ops, stream, args.steps, args.warmup, key="plan0")
        value_planner = sum(o["bytes"] for o in ops) * args.steps / (t0_ms * 1e-3) / 1e9
    total_ms, wall, ok = run_c2(torch, ops, stream, args.steps, args.warmup)
    ok = ok and ok0
    step_bytes = sum(o["bytes"] for o in ops)
    value = step_bytes * args.steps / (total_ms * 1e-3) / 1e9
    ceiling = copy_ceiling(torch, rt, dev, stream)
    log("  copy ceiling (row copy, 200 MiB fp32, %d launches over 8 buffer pairs): %.1f GB/s" % (
        ceiling["launches"], ceiling["GBps"]))
    # per-op and per-kernel times (diagnostic pass, each launch between events)
    per_op = per_op_ms(torch, ops, stream, 5)
    groups = {}
    for r in per_op:
        for i, o in enumerate(ops):
            g = groups.setdefault(o["kernel"], {"ms": 0.0, "bytes": 0, "launches": 0, "ops": []})
            g["ms"] += r[i]
            g["bytes"] += o["bytes"]
            g["launches"] += 1
            if r is per_op[0]:
                g["ops"].append(o)
    for i, o in enumerate(ops):
        t = sorted(r[i] for r in per_op)
        log("  k=%2d %s %s rank=%d tile=%s vec=%d/%d %8.1f GB/s  min %.3f ms  median %.3f ms  %s" % (
            o["k"], "f32" if o["es"] == 4 else "f64", "b0" if o["bz"] else "b1", o["info"]["rank"],
            o["info"]["tile"], o["info"]["vec_a"], o["info"]["vec_b"], o["bytes"] / (t[0] * 1e-3) / 1e9, t[0],
            t[len(t) // 2], o["kernel"]))
    for k, g in sorted(groups.items(), key=lambda kv: -kv[1]["ms"]):
        log("  kernel %-48s launches=%4d avg=%.4f ms  %.1f GB/s" % (
            k, g["launches"], g["ms"] / g["launches"], g["bytes"] / (g["ms"] * 1e-3) / 1e9))
    # the roofline line is for the kernel with the most time, timed in a region of its own
    dom = max(groups, key=lambda k: groups[k]["ms"])
    avg_ms, bytes_per_launch = kernel_avg_ms(torch, groups[dom]["ops"], stream, max(2, args.steps))
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    log("  roofline kernel %s: %d ops x %d rounds, avg %.4f ms, %.1f GB/s" % (
        dom, len(groups[dom]["ops"]), max(2, args.steps), avg_ms, achieved))
    out = {
        "metric": "GB/s (r+w) & % of HBM peak per transpose; batched transposes/sec vs plain",
        # a step whose sampled outputs differ from the element formula has no value
        "value": round(value, 2) if ok else None, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32,f64", "data": "synthetic",
        "config": {"workload": "16 seeded random perms rank 2-6, 200 MiB A each, fp32+fp64, beta 0 and 1 "
                               "(64 transposes per step)", "bytes_per_step": step_bytes,
                   "plans": ("autotuned, <= %d candidates per op" % args.tune) if args.tune else "planner heuristic",
                   "value_planner_heuristic": round(value_planner, 2) if value_planner else None,
                   "value_GiBps": round(value * 1e9 / 2 ** 30, 2),
                   "pct_of_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2), "verified_sample": ok,
                   "timed_wall_s": round(wall, 3)},
        "copy_ceiling_GBps": ceiling["GBps"],
        "frac_of_copy": round(value / ceiling["GBps"], 4),
        "copy_ceiling": ceiling,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                     "frac_of_copy": round(achieved / ceiling["GBps"], 4),
                     "avg_launch_ms": round(avg_ms, 5), "bytes_per_launch": bytes_per_launch,
                     "launches": len(groups[dom]["ops"]) * max(2, args.steps)},
    }
    if args.batch:
        del bufs, ops
        torch.cuda.empty_cache()
        try:
            out["batched"] = run_batch(torch, rt, wl, dev, stream, args.batch_replays, min(args.tune, 4))
        except Exception as e:  # reported, not hidden
            out["batched"] = {"error": str(e)}
    if args.workloads:
        try:
            out["workloads"] = {"plain_c1": run_c1(torch, rt, wl, dev, stream)}
            out["workloads"].update(run_workloads(torch, rt, wl, dev, stream, tune=min(args.tune, 32)))
        except Exception as e:  # reported, not hidden
            out.setdefault("workloads", {})["error"] = str(e)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)
    if not ok:
        log("bench.py: sampled outputs differ from the element formula; no value reported")
        sys.exit(1)


if __name__ == "__main__":
    main()
