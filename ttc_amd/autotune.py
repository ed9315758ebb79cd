"""On-device search over kernel configurations.

TTC's method — enumerate blockings x loop orders, prune to about
sqrt(maxImplementations) of each by cost models, compile and time every
survivor, keep the fastest (ttc/GPUtransposeGenerator.py:176-230, 328-349;
ttc/ttc.py:871-1022) — is kept, with the search space of these kernels:
kernel family (row x column tiles, box tiles, or row copy when dimension 0
stays in place), tile shape (through run-length
targets on A's and B's side), vector width per side, tile numbering,
tile-walk order, work-groups per CU and non-temporal stores.  Nothing is compiled per candidate: every candidate is a launch of a
prebuilt kernel with different arguments, timed with HIP events.
"""
import math

from . import runtime as rt
from . import ttc_util


def tile_score(info, size_a, size_b):
    """Same model as the planner (csrc/ttc_plan.cpp tile_score and
    choose_rc_tile): fraction of useful tile slots x b/(b+64) per side for
    contiguous runs of b bytes, each side weighted by its share of the bytes
    (exponent 2 s/(s_A + s_B)), x V/(V+256), times u/(0.3+0.7u) for the
    row x column kernel, whose tiles fill a fraction u of its slots."""
    n = info["elements"]
    if n == 0:
        return 0.0
    eff = n / float(info["tiles"] * info["tile_volume"])
    ba, bb = info["run_a"] * size_a, info["run_b"] * size_b
    v = info["tile_volume"]
    wa, wb = 2.0 * size_a / (size_a + size_b), 2.0 * size_b / (size_a + size_b)
    sc = eff * (ba / (ba + 64.0)) ** wa * (bb / (bb + 64.0)) ** wb * (v / (v + 256.0))
    if info["family"] == 2:
        cap = info["threads"] * min(32, 128 // max(size_a, size_b))
        u = v / float(cap)
        sc *= u / (0.3 + 0.7 * u)
    return sc


def _row_group(info):
    """A's leading dimensions that form a row x column tile's rows: from
    dimension 0 up to B's stride-1 dimension, each covered by the tile
    (csrc/ttc_plan.cpp rc_groups)."""
    r, perm, size, tile = info["rank"], info["perm"], info["size"], info["tile"]
    rows = []
    for k in range(r):
        if k == perm[0] or (k > 0 and (tile[k] == 1 or size[k - 1] % tile[k - 1])):
            break
        rows.append(k)
    return rows


def tile_walk_cost(info, tile_order):
    """TTC's loop-order cost (ttc_util.walkCost, ttc/ttc_util.py:167-209)
    of the order in which a row x column plan numbers its tiles: the tile
    grid is a loop nest whose loop over dimension k has ceil(n_k / t_k)
    iterations and steps t_k elements in A and in B.  tile_order 1 walks A
    order (dimension 0 innermost), 2 B order, 3 / 4 B order after an inner
    group of 8 / 16 tiles along the rows' outermost tiled dimension."""
    r, perm, size, tile = info["rank"], info["perm"], info["size"], info["tile"]
    sa, sb = info["stride_a"], info["stride_b"]
    g = [-(-size[k] // tile[k]) if size[k] else 1 for k in range(r)]
    order = list(range(r)) if tile_order == 1 else list(perm)  # innermost first
    inner_first = [(g[k], tile[k] * sa[k], tile[k] * sb[k]) for k in order if g[k] > 1]
    if tile_order in (3, 4):
        G = 8 if tile_order == 3 else 16
        kg = next((k for k in reversed(_row_group(info)) if g[k] > 1), None)
        if kg is not None:
            gi = min(G, g[kg])
            outer = -(-g[kg] // gi)
            inner_first = [(gi, tile[kg] * sa[kg], tile[kg] * sb[kg])]
            for k in order:
                if k == kg:
                    if outer > 1:
                        inner_first.append((outer, gi * tile[k] * sa[k], gi * tile[k] * sb[k]))
                elif g[k] > 1:
                    inner_first.append((g[k], tile[k] * sa[k], tile[k] * sb[k]))
    return ttc_util.walkCost(list(reversed(inner_first)))


def candidates(type_a, type_b, perm, size, lda=None, ldb=None, beta_is_zero=False, max_impl=200,
               nt_stores=False, blockings=None):
    """List of config dicts, best-first by the model, at most max_impl.  The
    planner's own choice ({}) comes first unless `blockings` is given.

    nt_stores: also try non-temporal B stores (beta == 0 only; the command
    line's --streamingStores).  They measure well in a loop over a few
    buffers and lose in longer streams of different transpositions, so they
    are not tried by default.
    blockings: [(a, b), ...] restricts the tiles to those derived from these
    run-length targets on A's and B's side, as the reference restricts its
    blockings (--blockings, ttc/ttc.py:1296-1297, ttc/GPUtransposeGenerator.py:
    144-150).
    Tile walks of row x column tiles are tried in the order of TTC's
    loop-order cost of their tile grid (tile_walk_cost), as the reference
    sorts its loop orders by getCostLoop before pruning."""
    base = rt.Plan(type_a, type_b, perm, size, lda, ldb, beta_is_zero).info()
    sa, sb = rt.DTYPE_BYTES[type_a], rt.DTYPE_BYTES[type_b]
    vecs = []
    for va in (base["vec_a"], 1):
        for vb in (base["vec_b"], 1):
            if (va, vb) not in vecs:
                vecs.append((va, vb))
    tiles = {}
    runs = [1 << k for k in range(0, 14)]
    # (0, 0): the planner's own search for each family and widths, which
    # also weighs tiles covering short leading extents whole
    targets = list(blockings) if blockings else [(0, 0)] + [(ra, rb) for ra in runs for rb in runs
                                                           if ra * rb <= 4 * 16384]
    for fam in (2, 1):
        for va, vb in vecs:
            for ra, rb in targets:
                cfg = {"vec_a": va, "vec_b": vb, "run_a": int(ra), "run_b": int(rb), "family": fam}
                try:
                    info = rt.Plan(type_a, type_b, perm, size, lda, ldb, beta_is_zero, cfg).info()
                except rt.TTCError:
                    continue
                key = (fam, tuple(info["tile"]), va, vb)
                if key not in tiles:
                    tiles[key] = (tile_score(info, sa, sb),
                                  {"tile": info["tile"], "vec_a": va, "vec_b": vb, "family": fam}, info)
    ranked = [(c, i) for _, c, i in sorted(tiles.values(), key=lambda x: -x[0])]
    # distinct tile shapes first (each with its best-scoring widths), then the
    # other widths of shapes already listed, so that pruning keeps shapes
    seen, first, rest = set(), [], []
    for c, i in ranked:
        key = (c["family"], tuple(c["tile"]))
        (rest if key in seen else first).append((c, i))
        seen.add(key)
    ranked = first + rest
    # the best tile of each family first, so that pruning never drops a family
    heads = []
    for fam in (2, 1):
        best = next((x for x in ranked if x[0]["family"] == fam), None)
        if best is not None:
            heads.append(best)
    ranked = heads + [x for x in ranked if all(x is not h for h in heads)]
    # launch variants: tile walk (grid order), tile numbering (row x column
    # tiles, per tile in loop-order-cost order), store mode
    grid_orders = (4, 3, 1)
    nts = (2, 1) if beta_is_zero and nt_stores else (0,)
    n_launch = len(grid_orders) * 3 * len(nts)
    # sqrt pruning of both axes, as the reference combines its best blockings
    # with its best loop orders
    budget = max(1, max_impl - (0 if blockings else 1))
    k = max(1, int(math.ceil(math.sqrt(budget))))
    while len(ranked) * n_launch > budget and len(ranked) > k:
        ranked.pop()
    while len(ranked) * n_launch > budget and n_launch > 1:
        n_launch -= 1
    out = [] if blockings else [{}]
    if base["family"] == 3 and not blockings:  # row copy (perm[0] == 0): resident work-groups per CU
        out += [{"family": 3, "wg_per_cu": w} for w in (4, 2)]
    for t, info in ranked:
        if t["family"] == 2:
            tords = sorted((2, 3, 1), key=lambda o: (tile_walk_cost(info, o), o))
        else:
            tords = (2,)  # box tiles have one numbering
        launch = [{"grid_order": go, "tile_order": to, "nt_store": nt}
                  for go in grid_orders for to in tords for nt in nts][:n_launch]
        for l in launch:
            c = dict(t)
            c.update(l)
            out.append(c)
    # waves per CU: the best tile with fewer resident work-groups than the
    # occupancy limit (the default), where the budget leaves room
    if ranked:
        t0, i0 = ranked[0]
        l0 = {"grid_order": grid_orders[0], "nt_store": nts[0],
              "tile_order": min((2, 3, 1), key=lambda o: (tile_walk_cost(i0, o), o)) if t0["family"] == 2 else 2}
        for w in (2, 3):
            if len(out) < max_impl:
                c = dict(t0)
                c.update(l0)
                c["wg_per_cu"] = w
                out.append(c)
        # one LDS buffer (two barriers per tile, more resident work-groups)
        if len(out) < max_impl and t0["family"] == 2:
            c = dict(t0)
            c.update(l0)
            c["lds_buffers"] = 1
            out.append(c)
    return out[:max(1, max_impl)]


def measure(plan, pairs, alpha, beta, reps=3, stream=None):
    """Steady-state time per launch in ms: `reps` rounds of back-to-back
    launches over the buffer pairs in `pairs`, timed with HIP events; the
    median round.  Alternating between pairs whose total footprint exceeds
    the last-level cache means every launch reads from and writes back to
    HBM, dirty-line write-back of the previous launch included — as in a
    stream of transpositions.  (A cache flush by writing a scratch buffer
    right before a single timed launch would charge that flush's write-back
    to the kernel.)"""
    import torch
    s = stream or torch.cuda.current_stream()  # launches and events on the same stream
    if plan.info()["type_b"] in (rt.E4M3, rt.E5M2):  # fp8 cast-transpose, delayed scaling
        def run(A, B):
            plan.execute_fp8(A, B, alpha, scale=1.0, stream=s)
    else:
        def run(A, B):
            plan.execute(A, B, alpha, beta, stream=s)
    for A, B in pairs:
        run(A, B)
    rounds = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for A, B in pairs:
            run(A, B)
        e1.record(s)
        e1.synchronize()
        rounds.append(e0.elapsed_time(e1) / len(pairs))
    rounds.sort()
    return rounds[len(rounds) // 2]


def _device_inputs(torch, type_a, type_b, perm, size, lda, ldb, dev):
    r = len(size)
    LA = list(lda) if lda else list(size)
    LB = list(ldb) if ldb else [size[perm[q]] for q in range(r)]
    na = int(math.prod(LA[:-1])) * size[-1] if r > 1 else size[0]
    nb = int(math.prod(LB[:-1])) * size[perm[-1]] if r > 1 else size[0]
    A = torch.empty(na * rt.DTYPE_BYTES[type_a], dtype=torch.uint8, device=dev).random_(0, 60)
    B = torch.empty(nb * rt.DTYPE_BYTES[type_b], dtype=torch.uint8, device=dev).random_(0, 60)
    return A, B


def tune(type_a, type_b, perm, size, lda=None, ldb=None, alpha=1.0, beta=0.0, max_impl=200, reps=5,
         progress=None, nt_stores=False, hot_a=False, hot_b=False, blockings=None):
    """Times every candidate on device buffers of the problem's shape.
    Returns results sorted fastest first: dicts with config, info, ms, GBs
    (10^9 B/s) and GiBs (2^30 B/s, the reference's unit).  hot_a / hot_b:
    the reference's cache modes (ttc/transposeGenerator.py:648-651,
    744-747): every launch of a round reuses the same A (resp. B), which
    therefore stays in cache when it fits, instead of a fresh buffer."""
    import torch
    dev = torch.device("cuda:0")
    bz = beta == 0
    A, B = _device_inputs(torch, type_a, type_b, perm, size, lda, ldb, dev)
    pairs = [(A, B)]
    # enough distinct buffers that a round streams well past the 256 MB
    # last-level cache, and at least four launches per round
    footprint = A.numel() + B.numel()
    while (footprint * len(pairs) < (1536 << 20) or len(pairs) < 4) and len(pairs) < 16:
        if hot_a and hot_b:
            pairs.append((A, B))
        else:
            A2, B2 = _device_inputs(torch, type_a, type_b, perm, size, lda, ldb, dev)
            pairs.append((A if hot_a else A2, B if hot_b else B2))
    results = []
    cands = candidates(type_a, type_b, perm, size, lda, ldb, bz, max_impl, nt_stores, blockings)
    for i, cfg in enumerate(cands):
        try:
            plan = rt.Plan(type_a, type_b, perm, size, lda, ldb, bz, cfg)
        except rt.TTCError:
            continue
        info = plan.info()
        ms = measure(plan, pairs, alpha, beta, reps)
        nbytes = info["bytes_algorithmic"]
        results.append({"config": cfg, "info": info, "ms": ms, "GBs": nbytes / (ms * 1e-3) / 1e9,
                        "GiBs": nbytes / (ms * 1e-3) / float(1 << 30)})
        if progress:
            progress(i + 1, len(cands))
    results.sort(key=lambda r: r["ms"])
    return results


def op_bytes(op):
    """Algorithmic bytes of one batch op (dict as for runtime.Batch)."""
    n = int(math.prod(op["size"]))
    ta, tb = op["type_a"], op.get("type_b", op["type_a"])
    bz = op.get("beta_is_zero", op.get("beta", 0.0) == 0)
    return n * (rt.DTYPE_BYTES[ta] + rt.DTYPE_BYTES[tb] * (1 if bz else 2))


def tune_batch(ops, max_impl=4, min_bytes=1 << 20, reps=3):
    """Per-op configurations for a grouped batch (the `config` of each op
    dict for runtime.Batch).  Every op of at least `min_bytes` algorithmic
    bytes is tuned on its own shape, each distinct problem once; smaller
    ops, whose time is launch and tail latency rather than tile shape, keep
    the planner's choice (None).  Returns (configs, number of problems
    tuned)."""
    memo = {}
    out = []
    for op in ops:
        if op_bytes(op) < min_bytes:
            out.append(None)
            continue
        ta, tb = op["type_a"], op.get("type_b", op["type_a"])
        bz = op.get("beta_is_zero", op.get("beta", 0.0) == 0)
        key = (ta, tb, tuple(op["perm"]), tuple(op["size"]), tuple(op.get("lda") or ()),
               tuple(op.get("ldb") or ()), bz)
        if key not in memo:
            res = tune(ta, tb, op["perm"], op["size"], op.get("lda"), op.get("ldb"), op.get("alpha", 1.0),
                       0.0 if bz else 1.0, max_impl, reps)
            memo[key] = res[0]["config"] if res and res[0]["config"] else None
        out.append(memo[key])
    return out, len(memo)
