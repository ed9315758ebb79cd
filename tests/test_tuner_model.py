"""CPU tests of the search's cost models and restrictions: TTC's loop-order
cost (ttc/ttc_util.py:167-209) pinned by brute force, its use to order the
tile walks the tuner tries, and the --blockings / --streamingStores
restrictions of the candidate list (ttc/ttc.py:1296-1297, ttc/ttc_util.py:
301-314)."""
import itertools
import math
import os
import random

import pytest

from ttc_amd import autotune, runtime as rt, ttc, ttc_util


def _brute_walk(loops):
    """Sum of squared A and B address jumps between consecutive visits of a
    loop nest given outermost first as (extent, strideA, strideB)."""
    prev, cost = None, 0.0
    for idx in itertools.product(*[range(n) for n, _, _ in loops]):
        a = sum(i * sa for i, (_, sa, _) in zip(idx, loops))
        b = sum(i * sb for i, (_, _, sb) in zip(idx, loops))
        if prev is not None:
            cost += (a - prev[0]) ** 2 + (b - prev[1]) ** 2
        prev = (a, b)
    return cost


def test_walk_cost_is_the_sum_of_squared_jumps():
    rng = random.Random(5)
    for _ in range(200):
        r = rng.randint(1, 4)
        loops = [(rng.randint(1, 5), rng.randint(1, 50), rng.randint(1, 50)) for _ in range(r)]
        assert ttc_util.walkCost(loops) == _brute_walk(loops), loops


def test_cost_loop_over_element_strides():
    assert ttc_util.getCostLoop([0, 1], [1, 0], [2, 3]) == 30.0
    rng = random.Random(6)
    for _ in range(60):
        r = rng.randint(1, 4)
        size = [rng.randint(1, 4) for _ in range(r)]
        perm = list(range(r))
        rng.shuffle(perm)
        loop = list(range(r))
        rng.shuffle(loop)
        pad = rng.random() < 0.5
        lda = [s + rng.randint(0, 2) for s in size] if pad else None
        ldb = [size[perm[q]] + rng.randint(0, 2) for q in range(r)] if pad else None
        sA, sB = [1] * r, [0] * r
        LA = lda or size
        LB = ldb or [size[perm[q]] for q in range(r)]
        for k in range(1, r):
            sA[k] = sA[k - 1] * LA[k - 1]
        pos = [1] * r
        for q in range(1, r):
            pos[q] = pos[q - 1] * LB[q - 1]
        for q in range(r):
            sB[perm[q]] = pos[q]
        want = _brute_walk([(size[k], sA[k], sB[k]) for k in loop])
        assert ttc_util.getCostLoop(loop, perm, size, lda, ldb) == want


def _grid_walk(info, tile_order):
    """Tile origins (A and B element offsets) in the order a row x column
    plan numbers them: A order (1), B order (2), or B order after an inner
    group of G = 8 / 16 tiles along the rows' outermost tiled dimension."""
    r, perm, size, tile = info["rank"], info["perm"], info["size"], info["tile"]
    sa, sb = info["stride_a"], info["stride_b"]
    g = [-(-size[k] // tile[k]) for k in range(r)]
    order = list(range(r)) if tile_order == 1 else list(perm)
    coords = []
    if tile_order in (3, 4):
        G = 8 if tile_order == 3 else 16
        kg = next((k for k in reversed(autotune._row_group(info)) if g[k] > 1), None)
    else:
        kg = None
    if kg is None:
        dims = [k for k in order if g[k] > 1]
        for idx in itertools.product(*[range(g[k]) for k in reversed(dims)]):
            x = dict(zip(reversed(dims), idx))
            coords.append(x)
    else:
        gi = min(G, g[kg])
        outer = -(-g[kg] // gi)
        rest = [k for k in order if g[k] > 1 and k != kg]
        # loops outermost first: B order of the rest (with kg's outer part in kg's place), inner group last
        outer_loops = []
        for k in reversed(order):
            if k == kg and outer > 1:
                outer_loops.append(("o", k))
            elif k != kg and g[k] > 1:
                outer_loops.append(("d", k))
        ranges = [range(outer) if t == "o" else range(g[k]) for t, k in outer_loops] + [range(gi)]
        for idx in itertools.product(*ranges):
            x = {}
            for (t, k), v in zip(outer_loops, idx[:-1]):
                x[k] = x.get(k, 0) + (v * gi if t == "o" else v)
            x[kg] = x.get(kg, 0) + idx[-1]
            if x[kg] < g[kg]:
                coords.append(x)
            else:
                coords.append(None)  # a padding id past the last real tile: no tile, no jump
    pts = []
    for x in coords:
        if x is None:
            continue
        pts.append((sum(v * tile[k] * sa[k] for k, v in x.items()), sum(v * tile[k] * sb[k] for k, v in x.items())))
    return pts


@pytest.mark.parametrize("perm,size", [([1, 0], [300, 200]), ([2, 0, 1], [64, 40, 30]), ([1, 2, 0], [60, 50, 40]),
                                       ([3, 1, 0, 2], [40, 12, 30, 9]), ([2, 1, 0], [100, 3, 90])])
def test_tile_walk_cost_matches_the_walk(perm, size):
    """For A and B order the model equals the squared jumps of the tile
    origins in the kernel's numbering; grouped orders with a full last
    group likewise."""
    rt.lib()
    for cfg in autotune.candidates(rt.S, rt.S, perm, size, max_impl=40):
        if cfg.get("family") != 2:
            continue
        info = rt.Plan(rt.S, rt.S, perm, size, config=dict(cfg, tile_order=2)).info()
        for order in (1, 2, 3):
            pts = _grid_walk(info, order)
            brute = sum((a1 - a0) ** 2 + (b1 - b0) ** 2 for (a0, b0), (a1, b1) in zip(pts, pts[1:]))
            g = [-(-info["size"][k] // info["tile"][k]) for k in range(info["rank"])]
            kg = next((k for k in reversed(autotune._row_group(info)) if g[k] > 1), None)
            if order == 3 and kg is not None and g[kg] % min(8, g[kg]):
                continue  # a partial last group: the walk skips padding ids the model counts
            assert autotune.tile_walk_cost(info, order) == brute, (cfg, order)


def test_blockings_restrict_the_candidates():
    rt.lib()
    perm, size = [1, 0], [1000, 768]
    cands = autotune.candidates(rt.S, rt.S, perm, size, max_impl=40, blockings=[(64, 64)])
    assert cands and {} not in cands
    allowed = set()
    for fam in (1, 2):
        for va, vb in ((4, 4), (4, 1), (1, 4), (1, 1)):
            try:
                allowed.add(tuple(rt.Plan(rt.S, rt.S, perm, size, config={"vec_a": va, "vec_b": vb, "run_a": 64,
                                                                            "run_b": 64, "family": fam}).info()["tile"]))
            except rt.TTCError:
                pass
    assert {tuple(c["tile"]) for c in cands} <= allowed


def test_streaming_stores_join_the_search_only_when_beta_is_zero():
    rt.lib()
    perm, size = [2, 0, 1], [64, 40, 30]
    nt = autotune.candidates(rt.S, rt.S, perm, size, beta_is_zero=True, max_impl=40, nt_stores=True)
    assert any(c.get("nt_store") == 1 for c in nt)
    plain = autotune.candidates(rt.S, rt.S, perm, size, beta_is_zero=True, max_impl=40)
    assert not any(c.get("nt_store") == 1 for c in plain)
    b1 = autotune.candidates(rt.S, rt.S, perm, size, beta_is_zero=False, max_impl=40, nt_stores=True)
    assert not any(c.get("nt_store") == 1 for c in b1)
    assert ttc_util.streamingStoresApplicable([], size, perm, 0.0, 1, 4, 1) == 1
    assert ttc_util.streamingStoresApplicable([], size, perm, 1.0, 1, 4, 1) == 0
    assert ttc_util.streamingStoresApplicable([], size, perm, 0.0, 1, 4, 0) == 0
    assert ttc_util.streamingStoresApplicable([41, 64, 30], size, perm, 0.0, 16, 4, 1) == 0  # 164 B rows


@pytest.fixture()
def workdir(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("TTC_ROOT", str(tmp_path / "root"))
    return tmp_path


def test_command_line_flags_reach_the_emitted_configuration(workdir):
    a = ttc.parse_args(["--perm=2,0,1", "--size=64,40,30", "--streamingStores", "--generateOnly"])
    name, _ = ttc.generateTransposition(a)
    src = (workdir / "ttc_transpositions" / (name + ".hip")).read_text()
    cfg = src.split("static const ttc_config_t cfg = ")[1].split(";")[0]
    assert cfg.split("}, ")[1].split(", ")[3] == "1"  # nt_store
    a = ttc.parse_args(["--perm=2,0,1", "--size=64,40,30", "--beta=1", "--streamingStores", "--generateOnly"])
    name, _ = ttc.generateTransposition(a)
    src = (workdir / "ttc_transpositions" / (name + ".hip")).read_text()
    cfg = src.split("static const ttc_config_t cfg = ")[1].split(";")[0]
    assert cfg.split("}, ")[1].split(", ")[3] == "0"  # B is read: no streaming stores
    a = ttc.parse_args(["--perm=1,0", "--size=1000,768", "--blockings=16x256", "--generateOnly"])
    name, _ = ttc.generateTransposition(a)
    src = (workdir / "ttc_transpositions" / (name + ".hip")).read_text()
    want = rt.Plan(rt.S, rt.S, [1, 0], [1000, 768], beta_is_zero=True, config={"run_a": 16, "run_b": 256}).info()
    assert "{{%s}" % ", ".join(str(t) for t in want["tile"] + [0] * 6) in src
