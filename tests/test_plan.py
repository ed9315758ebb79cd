"""CPU tests of the host side of libttc_hip.so: the library loads and
exports every symbol of include/ttc_hip.h, index fusion agrees with the
Python restatement of ttc/ttc_util.py:214-276 and leaves the result of the
operation unchanged, and plans satisfy the kernels' constraints.  No compute
is launched (there is no GPU here)."""
import os
import re

import numpy as np
import pytest

from common import CODE, D, S, make_case, oracle_run, same_bits
from ttc_amd import runtime as rt
from ttc_amd import ttc_util

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ttc_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+[\s\*]+\**(ttc_[a-z0-9_]+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_header_symbol():
    import ctypes
    names = header_functions()
    assert len(names) >= 16
    L = ctypes.CDLL(rt.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(rt.EXPORTS)


def test_version_and_sizes():
    assert "gfx950" in rt.version()
    L = rt.lib()
    assert [L.ttc_dtype_size(t) for t in range(8)] == [4, 8, 8, 16, 2, 2, 1, 1]


def _py_fuse(perm, size, lda, ldb):
    s, p = list(size), list(perm)
    la, lb = list(lda or []), list(ldb or [])
    ttc_util.fuseIndices(s, p, [], la, lb)
    return p, s, la, lb


def random_problem(rng, max_rank=7):
    r = int(rng.integers(1, max_rank + 1))
    size = [int(x) for x in rng.integers(1, 6, r)]
    perm = [int(x) for x in rng.permutation(r)]
    if rng.random() < 0.5:  # make fusable runs likely
        k = int(rng.integers(0, r))
        perm = sorted(perm[:k]) + perm[k:]
    lda = [s + int(rng.integers(0, 2)) for s in size] if rng.random() < 0.3 else None
    ldb = [size[perm[q]] + int(rng.integers(0, 2)) for q in range(r)] if rng.random() < 0.3 else None
    return perm, size, lda, ldb


def test_fusion_matches_python_restatement():
    rng = np.random.default_rng(1)
    for _ in range(400):
        perm, size, lda, ldb = random_problem(rng)
        cp, cs, cla, clb = rt.fuse_indices(perm, size, lda, ldb)
        pp, ps, pla, plb = _py_fuse(perm, size, lda, ldb)
        assert cp == pp and cs == ps, (perm, size, lda, ldb)
        if lda:
            assert cla == pla
        if ldb:
            assert clb == plb


def test_fusion_known_cases():
    # contiguous in both -> one index (the identity is a copy)
    assert rt.fuse_indices([0, 1, 2], [4, 5, 6])[:2] == ([0], [120])
    # (0,1) move together
    assert rt.fuse_indices([2, 0, 1], [4, 5, 6])[:2] == ([1, 0], [20, 6])
    # padding in A breaks the run
    assert rt.fuse_indices([2, 0, 1], [4, 5, 6], [5, 5, 6], None)[:2] == ([2, 0, 1], [4, 5, 6])
    # padding on the LAST index of a run does not
    assert rt.fuse_indices([2, 0, 1], [4, 5, 6], [4, 7, 6], None)[:2] == ([1, 0], [20, 6])
    # configs of the benchmark are left alone
    assert rt.fuse_indices([1, 0, 2], [1000, 768, 16])[:2] == ([1, 0, 2], [1000, 768, 16])
    assert rt.fuse_indices([3, 1, 0, 2], [256, 192, 128, 96])[:2] == ([3, 1, 0, 2], [256, 192, 128, 96])


def test_fusion_is_idempotent_and_preserves_the_result(oracle_lib):
    rng = np.random.default_rng(2)
    for i in range(120):
        perm, size, lda, ldb = random_problem(rng, 6)
        fp, fs, fla, flb = rt.fuse_indices(perm, size, lda, ldb)
        assert rt.fuse_indices(fp, fs, fla, flb)[:2] == (fp, fs)
        A, B = make_case(S, S, perm, size, lda, ldb, seed=i, dist="normal")
        want = oracle_run(S, S, A, B, 1.5, -0.5, perm, size, lda, ldb)
        got = oracle_run(S, S, A, B, 1.5, -0.5, fp, fs, fla if lda else None, flb if ldb else None)
        assert same_bits(got, want), (perm, size, lda, ldb, fp, fs)


def _check_plan(i, perm, size):
    r = i["rank"]
    assert r == len(i["perm"]) and sorted(i["perm"]) == list(range(r))
    V = 1
    tiles = 1
    for k in range(r):
        assert 1 <= i["tile"][k] <= max(1, i["size"][k])
        V *= i["tile"][k]
        tiles *= -(-i["size"][k] // i["tile"][k]) if i["size"][k] else 0
    assert V == i["tile_volume"]
    assert tiles == i["tiles"] or i["elements"] == 0 or i["family"] == 3
    assert i["tile"][0] % i["vec_a"] == 0 and i["tile"][i["perm"][0]] % i["vec_b"] == 0
    assert i["size"][0] % i["vec_a"] == 0 and i["size"][i["perm"][0]] % i["vec_b"] == 0
    def slots(vec, nbytes):  # ttc_params.h slots_for
        return min(16, max(4, 64 // (vec * nbytes)))
    def rc_elems(nbytes):  # ttc_params.h rc_elems
        return min(32, 128 // nbytes)
    if i["family"] == 3:
        # row copy: dimension 0 stays in place; one vector width for both
        # sides, at most 16 bytes on each; no LDS
        assert i["perm"][0] == 0 and i["vec_a"] == i["vec_b"]
        assert i["tile"][0] == i["size"][0] and V == i["size"][0]
        assert i["vec_a"] * max(rt.DTYPE_BYTES[i["type_a"]], rt.DTYPE_BYTES[i["type_b"]]) <= 16
        assert i["lds_bytes"] == 0 and i["elements"] == int(np.prod(size))
        return
    if i["family"] == 1:
        assert i["slots_a"] <= slots(i["vec_a"], rt.DTYPE_BYTES[i["type_a"]])
        assert i["slots_b"] <= slots(i["vec_b"], rt.DTYPE_BYTES[i["type_b"]])
        assert i["rows"] == 0 and i["cols"] == 0
    else:
        assert i["family"] == 2
        ea, eb = rc_elems(rt.DTYPE_BYTES[i["type_a"]]), rc_elems(rt.DTYPE_BYTES[i["type_b"]])
        assert i["slots_a"] * i["vec_a"] == ea and i["slots_b"] * i["vec_b"] == eb
        assert V <= 256 * min(ea, eb)
        _check_rc_groups(i)
    assert 0 < i["lds_bytes"] <= (81920 if i["family"] == 2 else 65536)
    assert i["lds_buffers"] == (2 if i["family"] == 2 else 1)
    partial = sum(1 for k in range(r) if i["size"][k] % i["tile"][k])
    assert partial <= 2
    assert i["elements"] == int(np.prod(size))


def _check_rc_groups(i):
    """A row x column tile: some split into rows = A's leading dimensions
    0..na-1 (never B's stride-1 dimension) and columns = B's leading
    dimensions perm[0..nb-1] followed by the other tiled dimensions,
    disjoint, with a partial last tile only on a group's outermost
    dimension, whose row and column counts are the plan's."""
    r, t, n, perm = i["rank"], i["tile"], i["size"], i["perm"]
    assert perm[0] != 0
    big = {k for k in range(r) if t[k] > 1}
    found = False
    for na in range(1, r + 1):
        ga = list(range(na))
        for nb in range(1, r + 1):
            gb = list(perm[:nb])
            if set(ga) & set(gb):
                continue
            gb = gb + sorted(big - set(ga) - set(gb))
            if any(n[k] % t[k] for k in ga[:-1] + gb[:-1]):
                continue
            rows, cols = int(np.prod([t[k] for k in ga])), int(np.prod([t[k] for k in gb]))
            if rows == i["rows"] and cols == i["cols"] and rows * cols == i["tile_volume"]:
                found = True
    assert found, (t, n, perm, i["rows"], i["cols"])


@pytest.mark.parametrize("code", ["s", "d", "c", "z", "sd", "ds", "cz", "zc", "b", "h", "sb", "sh", "bs", "hs"])
def test_plans_satisfy_kernel_constraints(code):
    tA, tB = CODE[code]
    rng = np.random.default_rng(3)
    for trial in range(150):
        r = int(rng.integers(1, 7))
        size = [int(x) for x in np.exp(rng.uniform(0, np.log(3000), r)).astype(int) + 1]
        while np.prod(size, dtype=np.float64) > 2e8:
            size[int(np.argmax(size))] //= 2
        perm = [int(x) for x in rng.permutation(r)]
        bz = bool(trial % 2)
        plan = rt.Plan(tA, tB, perm, size, beta_is_zero=bz)
        i = plan.info()
        _check_plan(i, perm, size)
        sa, sb = rt.DTYPE_BYTES[tA], rt.DTYPE_BYTES[tB]
        n = i["elements"]
        assert i["bytes_algorithmic"] == n * sa + n * sb * (1 if bz else 2)


def test_explicit_config_is_honoured():
    plan = rt.Plan(S, S, [1, 0, 2], [1000, 768, 16], config=dict(tile=[32, 64, 1], vec_a=2, vec_b=1))
    i = plan.info()
    assert i["tile"] == [32, 64, 1] and i["vec_a"] == 2 and i["vec_b"] == 1


def test_invalid_arguments_are_rejected():
    with pytest.raises(rt.TTCError) as e:
        rt.Plan(S, S, [0, 0], [3, 4])
    assert e.value.status == rt.ERROR_INVALID_VALUE
    with pytest.raises(rt.TTCError):
        rt.Plan(S, S, [1, 0], [3, 4], lda=[2, 4])       # lda < size (ttc/ttc.py:743-746)
    with pytest.raises(rt.TTCError):
        rt.Plan(S, S, [1, 0], [3, 4], ldb=[3, 4])       # ldb[0] < size[perm[0]] (ttc/ttc.py:739-742)
    with pytest.raises(rt.TTCError) as e:
        rt.Plan(D, rt.BF16, [1, 0], [3, 4])
    assert e.value.status == rt.ERROR_UNSUPPORTED
    with pytest.raises(rt.TTCError):
        rt.Plan(S, rt.E4M3, [1, 0], [3, 4], beta_is_zero=False)  # fp8 output only with beta == 0
    with pytest.raises(rt.TTCError):
        rt.Plan(S, S, [1, 0], [6, 4], config=dict(vec_a=4))   # 6 % 4 != 0
    with pytest.raises(rt.TTCError):
        rt.Plan(S, S, [1, 0], [100, 100], config=dict(tile=[100, 100]))  # volume above the kernel limit


def test_empty_and_extent_one_problems_plan():
    i = rt.Plan(S, S, [1, 0], [0, 5]).info()
    assert i["elements"] == 0 and i["tiles"] == 0
    perm = [15 - k for k in range(16)]
    size = [2, 1, 2, 1, 2, 1, 2, 1, 1, 2, 1, 1, 1, 2, 1, 1]
    i = rt.Plan(S, S, perm, size).info()  # extent-1 dimensions are dropped after fusion
    assert i["rank"] <= 8 and i["elements"] == 64


def test_too_many_partial_dims_is_unsupported():
    with pytest.raises(rt.TTCError) as e:
        rt.Plan(S, S, [2, 1, 0], [10, 10, 10], config=dict(tile=[3, 3, 3], vec_a=1, vec_b=1))
    assert e.value.status == rt.ERROR_UNSUPPORTED


def test_autotune_candidates_cover_families_orders_and_waves():
    """The search space: the planner's choice first, both kernel families,
    tile numberings, and fewer work-groups per CU; every candidate plans."""
    from ttc_amd import autotune
    perm, size = [2, 0, 1], [64, 48, 80]
    cands = autotune.candidates(S, S, perm, size, beta_is_zero=False, max_impl=12)
    assert cands[0] == {} and 2 <= len(cands) <= 12
    assert {c["family"] for c in cands[1:]} == {1, 2}
    assert any(c.get("tile_order") == 3 for c in cands) and any(c.get("wg_per_cu") for c in cands)
    assert any(c.get("tile_order") == 1 for c in autotune.candidates(S, S, perm, size, max_impl=40))
    for c in cands:
        rt.Plan(S, S, perm, size, beta_is_zero=False, config=c or None).info()
    assert len(autotune.candidates(S, S, perm, size, max_impl=1)) == 1


def test_lds_buffers():
    """Row x column plans stage through two LDS tile buffers by default and
    one on request; the second buffer costs exactly one more data region."""
    one = rt.Plan(S, S, [1, 0], [3000, 2000], config={"family": 2, "lds_buffers": 1}).info()
    two = rt.Plan(S, S, [1, 0], [3000, 2000], config={"family": 2}).info()
    assert one["lds_buffers"] == 1 and two["lds_buffers"] == 2 and one["tile"] == two["tile"]
    data = two["lds_bytes"] - one["lds_bytes"]  # one tile data region
    tables = one["lds_bytes"] - data
    assert data % 16 == 0 and data >= one["rows"] * one["cols"] * 4
    # the four offset tables, plus the lane-fixed walk's slot tables and padding
    assert tables % 4 == 0 and tables >= (2 * one["rows"] + 2 * one["cols"]) * 4
    with pytest.raises(rt.TTCError):
        rt.Plan(S, S, [1, 0], [3000, 2000], config={"lds_buffers": 3})


def test_batch_op_bytes():
    from ttc_amd import autotune
    assert autotune.op_bytes(dict(type_a=S, perm=[1, 0], size=[10, 20], beta=0.0)) == 200 * 8
    assert autotune.op_bytes(dict(type_a=D, type_b=S, perm=[1, 0], size=[10, 20], beta=1.0)) == 200 * 16


def test_row_copy_family_for_perm0():
    """Dimension 0 in place after fusion -> row copy, with the widest vector
    both sides allow; box tiles on request; never for other permutations."""
    for ta, tb, want_v in ((S, S, 4), (D, D, 2), (S, rt.E4M3, 4), (rt.BF16, rt.E5M2, 8), (S, rt.BF16, 4),
                           (rt.Z, rt.Z, 1)):
        i = rt.Plan(ta, tb, [0, 2, 1, 3], [128, 4096, 32, 8], beta_is_zero=True).info()
        assert i["family"] == 3 and i["vec_a"] == want_v == i["vec_b"], (ta, tb, i)
        assert i["tiles"] == 4096 * 32 * 8
    assert rt.Plan(S, S, [0, 2, 1], [6, 5, 7], beta_is_zero=False).info()["vec_a"] == 2
    assert rt.Plan(S, S, [0, 2, 1], [7, 5, 3], beta_is_zero=False).info()["vec_a"] == 1
    assert rt.Plan(S, S, [0, 1, 2], [7, 5, 3]).info()["family"] == 3  # identity: one row
    assert rt.Plan(S, S, [0, 2, 1], [8, 5, 7], config={"family": 1}).info()["family"] == 1
    # B's stride-1 dimension has extent 1 and a padded leading dimension:
    # once it is dropped, dimension 0 is in place but strided on B (3), so
    # rows are not contiguous and the row copy must not take it
    for ta in (S, rt.BF16):
        i = rt.Plan(ta, ta, [1, 0, 2], [4, 1, 16], [5, 1, 18], [3, 6, 18]).info()
        assert i["perm"] == [0, 1] and i["stride_b"] == [3, 18], i
        assert i["family"] == 1 and i["vec_a"] == i["vec_b"] == 1, i
        with pytest.raises(rt.TTCError):
            rt.Plan(ta, ta, [1, 0, 2], [4, 1, 16], [5, 1, 18], [3, 6, 18], config={"family": 3})
    # the same on A's side: A's stride-1 dimension has extent 1
    i = rt.Plan(S, S, [1, 0, 2], [1, 4, 16], [3, 4, 16], None).info()
    assert i["perm"] == [0, 1] and i["stride_a"] == [3, 12] and i["stride_b"] == [1, 4], i
    assert i["family"] == 1 and i["vec_a"] == 1, i
    with pytest.raises(rt.TTCError):
        rt.Plan(S, S, [1, 0, 2], [8, 5, 7], config={"family": 3})
    # explicit widths of the row copy: one width, from its own list
    i = rt.Plan(S, rt.BF16, [0, 2, 1], [64, 48, 80], config={"family": 3, "vec_a": 4, "vec_b": 4}).info()
    assert i["family"] == 3 and i["vec_a"] == i["vec_b"] == 4
    for bad in ({"vec_a": 4, "vec_b": 2}, {"vec_a": 8}, {"vec_b": 16}):
        with pytest.raises(rt.TTCError):
            rt.Plan(S, rt.BF16, [0, 2, 1], [64, 48, 80], config=dict(bad, family=3))
