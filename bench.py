#!/usr/bin/env python3
"""Benchmark of the MI355X transposition path.

Workload (one "step"): the 16 seeded random-permutation cases of rank 2..6,
each at 200 MiB of A, in fp32 and in fp64, with beta = 0 (B not read) and
beta = 1 (B read) — 64 out-of-place transpositions
B[pi(i)] = alpha*A[i] + beta*B[pi(i)], inputs resident in HBM.  All beta = 0
ops run first, then all beta = 1 ops, so consecutive ops never touch the same
buffers and no op finds its data in the 256 MB last-level cache.

Plans: by default each op is autotuned first (at most 12 candidates timed
on its own buffers, outside the timed region), as TTC generates and measures
variants before use (ttc/ttc.py:871-1022); the planner's untuned choice is
timed in the same run and reported beside it.

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
    (ttc::<family kernel><Op, VA, VB, BETA, NT>)."""
    ops = {(0, 0): "ttc::OpF32<float, float>", (1, 1): "ttc::OpDD"}
    fam = "transpose_rc_kernel" if info["family"] == 2 else "transpose_kernel"
    return "ttc::%s<%s, %d, %d, %s, %s>" % (fam, ops.get((info["type_a"], info["type_b"]), "?"), info["vec_a"],
                                            info["vec_b"], "false" if info["beta_is_zero"] else "true",
                                            "true" if info["nt_store"] else "false")


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


def verify_sample(torch, op, nsample=4096):
    """For a beta = 0, alpha = 1 op: B at sampled output positions equals A
    at the matching input positions (exact)."""
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
    a = op["A"][torch.from_numpy(offA).to(dev)]
    b = op["B"][torch.from_numpy(offB).to(dev)]
    return bool(torch.equal(a, b))


def run_c2(torch, ops, stream, steps, warmup, key="plan"):
    """Times `steps` passes over the ops with one pair of events around the
    whole region; nothing else is enqueued between the launches (an event
    recorded between two kernels adds a marker packet and costs several
    microseconds per op)."""
    # sample check of every beta = 0 op before the beta = 1 ops reuse B
    for o in ops:
        if o["bz"]:
            o[key].execute(o["A"], o["B"], o["alpha"], o["beta"], stream=stream)
    torch.cuda.synchronize()
    ok = all(verify_sample(torch, o) for o in ops if o["bz"])
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


def cpu_baseline(budget_s):
    """The oracle (C++ restatement of TTC's reference loop nest, OpenMP on all
    host cores) on a bounded sample of the same workload."""
    import numpy as np

    from oracle import oracle
    from ttc_amd import workloads as wl
    nthreads = os.cpu_count() or 1
    cases = wl.plain_c2()
    done, tot_bytes, tot_s = [], 0, 0.0
    trash = np.ones((512 << 20) // 8)
    t_start = time.perf_counter()
    for case in [cases[i] for i in (2, 6, 11, 3, 0, 9)]:
        if time.perf_counter() - t_start > budget_s:
            break
        size, perm = case["sizes"][4], case["perm"]
        n = int(np.prod(size))
        rng = np.random.default_rng(case["k"])
        A = rng.uniform(-1, 1, n).astype(np.float32)
        B = rng.uniform(-1, 1, n).astype(np.float32)
        for bz in (True, False):
            best = None
            for _ in range(2):
                oracle.trash_cache(trash, nthreads)
                t = time.perf_counter()
                oracle.transpose(oracle.S, oracle.S, A, B, 1.0, 0.0 if bz else 1.0, perm, size, nthreads=nthreads)
                dt = time.perf_counter() - t
                best = dt if best is None else min(best, dt)
            tot_bytes += n * 4 * (2 if bz else 3)
            tot_s += best
            done.append("k%d/%s" % (case["k"], "b0" if bz else "b1"))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(tot_bytes / tot_s / 1e9, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": "oracle loop nest (OpenMP, %d threads, %s) on fp32 C2 cases %s, min of 2 reps each after a "
                      "512 MiB cache-trash sweep" % (nthreads, model, ",".join(done))}


def run_batch(torch, rt, wl, dev, stream, replays, tune=0):
    """512 mixed ops: grouped persistent launch (graph replay) vs the per-op
    launch loop with the same kernels.  tune > 0: every op of at least 1 MiB
    takes the autotuner's configuration (at most `tune` candidates); the
    planner's choices are timed beside it."""
    ops = wl.batch_c5()
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    tmap = {0: (rt.S, torch.float32), 1: (rt.D, torch.float64), 4: (rt.BF16, torch.bfloat16)}
    desc = []
    keep = []
    for o in ops:
        t, tdt = tmap[o["type"]]
        n = 1
        for x in o["size"]:
            n *= x
        A = (torch.rand(n, generator=gen, device=dev, dtype=torch.float32) * 2 - 1).to(tdt)
        B = (torch.rand(n, generator=gen, device=dev, dtype=torch.float32) * 2 - 1).to(tdt)
        keep.append((A, B))
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
    t_loop = timed(lambda: batch.execute_loop(stream), max(10, replays // 10))
    t_loop_graph = timed(lambda: batch.graph_launch(stream, per_op_loop=True), max(10, replays // 10))

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
    return {"ops": n, "replays": replays, "tiles": info["tiles"], "workgroups": info["workgroups"],
            "bytes_per_launch": info["bytes"],
            "configs": ("autotuned (<= %d candidates) for the %d distinct problems of >= 1 MiB, planner for the "
                        "rest" % (tune, tuned)) if tune else "planner",
            "grouped_graph_ms": round(t_group, 4),
            "grouped_graph_ms_planner": round(t_group_planner, 4) if t_group_planner else None,
            "grouped_eager_ms": round(t_group_eager, 4),
            "loop_async_ms": round(t_loop, 4), "loop_graph_ms": round(t_loop_graph, 4),
            "loop_sync_legacy_ms": round(t_sync, 3),
            "transposes_per_s_grouped": round(n / (t_group * 1e-3), 1),
            "transposes_per_s_loop": round(n / (t_loop * 1e-3), 1),
            "speedup_vs_async_loop": round(t_loop / t_group, 3),
            "speedup_vs_graph_loop": round(t_loop_graph / t_group, 3),
            "speedup_vs_sync_legacy_loop": round(t_sync / t_group, 3),
            "grouped_GBps": round(info["bytes"] / (t_group * 1e-3) / 1e9, 1)}


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
    A and is not counted)."""
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
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--batch", dest="batch", action="store_true", default=True)
    ap.add_argument("--no-batch", dest="batch", action="store_false")
    ap.add_argument("--batch-replays", type=int, default=1000)
    ap.add_argument("--no-workloads", dest="workloads", action="store_false", default=True,
                    help="skip the mixed-precision / complex and fp8 cast-transpose throughput leg")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only the plain copy used to scale counters")
    ap.add_argument("--tune", type=int, default=12,
                    help="autotune each op over at most N candidates before timing, as TTC generates the fastest "
                         "variant before use (0: the planner's choice only)")
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
    value_planner = None
    if args.tune > 0:  # the planner's own choice, for comparison
        t0_ms, _, _ = run_c2(torch, ops, stream, args.steps, args.warmup, key="plan0")
        value_planner = sum(o["bytes"] for o in ops) * args.steps / (t0_ms * 1e-3) / 1e9
    total_ms, wall, ok = run_c2(torch, ops, stream, args.steps, args.warmup)
    step_bytes = sum(o["bytes"] for o in ops)
    value = step_bytes * args.steps / (total_ms * 1e-3) / 1e9
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
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32,f64", "data": "synthetic",
        "config": {"workload": "16 seeded random perms rank 2-6, 200 MiB A each, fp32+fp64, beta 0 and 1 "
                               "(64 transposes per step)", "bytes_per_step": step_bytes,
                   "plans": ("autotuned, <= %d candidates per op" % args.tune) if args.tune else "planner heuristic",
                   "value_planner_heuristic": round(value_planner, 2) if value_planner else None,
                   "value_GiBps": round(value * 1e9 / 2 ** 30, 2),
                   "pct_of_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2), "verified_sample": ok,
                   "timed_wall_s": round(wall, 3)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
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
            out["workloads"] = run_workloads(torch, rt, wl, dev, stream, tune=min(args.tune, 8))
        except Exception as e:  # reported, not hidden
            out["workloads"] = {"error": str(e)}
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
