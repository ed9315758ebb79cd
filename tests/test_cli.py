"""CPU tests of the drop-in Python surface: the ttc command line, TTCargs,
generateTransposition (database lookup, artefact emission), names and the
SQLite cache.  Mirrors how test-framework/ttc_test.py drives the reference
(ttc_test.py:124-156)."""
import json
import os
import sqlite3
import subprocess

import pytest

from ttc_amd import generator, sql_util, ttc, ttc_util

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture()
def workdir(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("TTC_ROOT", str(tmp_path / "root"))
    return tmp_path


def test_parse_args_matches_reference_flags():
    a = ttc.parse_args(["--perm=1,0,2", "--size=1000,768,16", "--beta=1.0", "--dataType=s", "--alpha=2",
                        "--lda=1000,768,16", "--ldb=768,1000,16", "--maxImplementations=7"])
    assert a.idxPerm == [1, 0, 2] and a.size == [1000, 768, 16]
    assert a.beta == 1.0 and a.alpha == 2.0 and a.maxNumImplementations == 7
    assert a.floatTypeA == "float" and a.floatTypeB == "float"
    assert a.lda == [1000, 768, 16] and a.ldb == [768, 1000, 16]
    a = ttc.parse_args(["--perm=1,0", "--size=4,5", "--dataType=zc", "--architecture=cuda"])
    assert (a.floatTypeA, a.floatTypeB) == ("double complex", "float complex")
    a = ttc.parse_args(["--perm=1,0", "--size=4,5", "--dataType=float->bfloat16"])
    assert (a.floatTypeA, a.floatTypeB) == ("float", "bfloat16")
    with pytest.raises(SystemExit):
        ttc.parse_args(["--perm=1,0", "--size=4,5", "--bogus"])
    with pytest.raises(SystemExit):
        ttc.parse_args(["--perm=1,0", "--size=4"])
    with pytest.raises(SystemExit):
        ttc.parse_args(["--perm=1,0", "--size=4,5", "--dataType=q"])


def test_names_follow_the_reference():
    # function name: ttc/CUDAtranspose.py:473-527; file name: ttc/ttc.py:41-95
    assert generator.function_name("float", "float", [1, 0, 2], [1000, 768, 16], 1.0) == \
        "sCuTranspose_102_1000x768x16"
    assert generator.function_name("float", "float", [1, 0], [4, 5], 0.0) == "sCuTranspose_10_4x5_bz"
    assert generator.function_name("double complex", "float complex", [1, 0], [4, 5], 1.0) == \
        "zcCuTranspose_10_4x5"
    assert generator.file_name("double complex", "float complex", [1, 0], [4, 5], 1.0) == "zsCuTranspose_10_4x5"
    assert generator.function_name("float", "double", [1, 0], [4, 5], 1.0) == "sdCuTranspose_10_4x5"
    assert generator.file_name("float", "double", [1, 0], [4, 5], 1.0) == "sdCuTranspose_10_4x5"
    assert generator.prototype("x", "double", "float", 1.0) == (
        'extern "C" void x(const double *A, float *B, const double alpha, const float beta, const int *size, '
        'const int *lda, const int *ldb)')
    assert generator.prototype("x", "float complex", "float complex", 0.0) == (
        'extern "C" void x(const hipFloatComplex *A, hipFloatComplex *B, const float alpha, const int *size, '
        'const int *lda, const int *ldb)')


def test_cost_loop_closed_form():
    # perm (1,0), size (2,3), loop order [0,1]:
    # i=0: idx 0, lda0 = 1, ldb0 = ldb[1] = 3; rest loop 1: lastA = 2*(3-1)=4, lastB = 1*(3-1)=2
    #      cost += 1*(2-1)*((4-1)^2 + (2-3)^2) = 10; total = 2
    # i=1: idx 1, lda0 = 2, ldb0 = 1, no rest: cost += 2*(3-1)*((0-2)^2 + (0-1)^2) = 20
    assert ttc_util.getCostLoop([0, 1], [1, 0], [2, 3]) == 30.0


def test_fuse_indices_keeps_empty_ld_lists_empty():
    size, perm, lda, ldb = [4, 5, 6], [2, 0, 1], [], []
    ttc_util.fuseIndices(size, perm, [], lda, ldb)
    assert size == [20, 6] and perm == [1, 0] and lda == [] and ldb == []
    loops = [[2, 1, 0]]
    size, perm = [4, 5, 6], [2, 0, 1]
    ttc_util.fuseIndices(size, perm, loops, [], [])
    assert loops == [[1, 0]]


def test_generate_only_emits_compilable_drop_in(workdir):
    a = ttc.parse_args(["--perm=1,0,2", "--size=1000,768,16", "--beta=1.0", "--dataType=s", "--generateOnly"])
    name, bw = ttc.generateTransposition(a)
    assert name == "sCuTranspose_102_1000x768x16" and bw == 0
    src = (workdir / "ttc_transpositions" / (name + ".hip")).read_text()
    hdr = (workdir / "ttc_transpositions" / (name + ".h")).read_text()
    assert '\nvoid sCuTranspose_102_1000x768x16(const float *A, float *B, const float alpha, ' \
           'const float beta, const int *size, const int *lda, const int *ldb);' in hdr
    assert '#ifdef __cplusplus\nextern "C" {\n#endif' in hdr  # C and C++ callers
    assert 'extern "C" void sCuTranspose_102_1000x768x16(const float *A' in src
    # the header compiles as C (TTC-C's language) as well as C++
    r = subprocess.run(["gcc", "-x", "c", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        str(workdir / "ttc_transpositions" / (name + ".h"))], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    assert "ttc_legacy_transpose(TTC_S, TTC_S" in src
    obj = workdir / "a.o"
    r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-c", str(workdir / "ttc_transpositions" / (name + ".hip")),
                        "-I", os.path.join(ROOT, "include"), "-o", str(obj)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    nm = subprocess.run(["nm", "-g", str(obj)], stdout=subprocess.PIPE, text=True).stdout
    assert " T sCuTranspose_102_1000x768x16" in nm and "ttc_legacy_transpose" in nm


def test_generate_fuses_before_naming(workdir):
    a = ttc.parse_args(["--perm=2,0,1", "--size=4,5,6", "--dataType=d", "--generateOnly", "--beta=0"])
    name, _ = ttc.generateTransposition(a)
    assert name == "dCuTranspose_10_20x6_bz"  # (0,1) fused, as ttc/ttc.py:679 before naming (770)


def test_invalid_problems_exit(workdir):
    for argv in (["--perm=1,1", "--size=4,5"], ["--perm=1,0", "--size=4,5", "--lda=3,5"],
                 ["--perm=1,0", "--size=4,5", "--ldb=4,4"], ["--perm=1,0", "--size=4,5", "--lda=4"],
                 ["--perm=1,0", "--size=4,5", "--architecture=avx"]):
        a = ttc.parse_args(argv + ["--generateOnly"])
        with pytest.raises(SystemExit):
            ttc.generateTransposition(a)


def test_database_roundtrip_serves_stored_solution(workdir):
    """A stored measurement is found by the reference's lookup sequence
    (size -> permutation -> measurement -> best variant) and its
    configuration is emitted without a GPU."""
    db = ttc.database_path()
    os.makedirs(os.path.dirname(db), exist_ok=True)
    con = sqlite3.connect(db)
    cur = con.cursor()
    ttc.createTables(cur)
    sid = ttc.insertIntoSize(cur, [100, 76, 16])
    pid = ttc.insertIntoPermutation(cur, [1, 0, 2])
    mid = ttc.insertIntoMeasurements(cur, 3, "h", 0, 1.0, 1.0, 0, "hipcc", "float", "float", 0, 0, 0, 0, sid, pid,
                                     -1, 0, 0, ttc_util.getArchitecture("hip"))
    lid = ttc.insertIntoLoopOrder(cur, [2, 1, 0])
    for rank, (tile, bw) in enumerate((([32, 32, 1], 900.0), ([64, 16, 2], 1200.0))):
        res = {"config": {"tile": tile, "vec_a": 4, "vec_b": 4}, "ms": 0.1, "GiBs": bw,
               "info": {"run_a": tile[0], "run_b": tile[1], "tile": tile, "vec_a": 4, "vec_b": 4, "threads": 256,
                        "nt_store": 0, "grid_order": 0}}
        ttc.insertIntoVariant(cur, mid, lid, rank, res)
    con.commit()
    assert sql_util.getSizeId(cur, [100, 76, 16]) == sid
    assert sql_util.getPermId(cur, [1, 0, 2]) == pid
    assert sql_util.getMeasurementId(cur, sid, pid, "float", "float", 2.0, 0, 0, 0,
                                     ttc_util.getArchitecture("hip")) == mid
    assert sql_util.getMeasurementId(cur, sid, pid, "float", "float", 0.0, 0, 0, 0,
                                     ttc_util.getArchitecture("hip")) == -1  # beta == 0 is a different key
    cfg, bw = sql_util.getBestConfig(cur, mid)
    assert cfg["tile"] == [64, 16, 2] and bw == 1200.0
    assert sql_util.getBestImplementation(cur, mid)[:2] == (64, 16)
    assert sql_util.getLoopPermFrom(cur, lid, 3) == [2, 1, 0]
    con.close()
    a = ttc.parse_args(["--perm=1,0,2", "--size=100,76,16", "--beta=1.0", "--dataType=s"])
    name, got_bw = ttc.generateTransposition(a)  # no --generateOnly: served from the database
    assert got_bw == 1200.0
    src = (workdir / "ttc_transpositions" / (name + ".hip")).read_text()
    assert "{{64, 16, 2, 0, 0, 0, 0, 0}, 4, 4" in src and src.count(",") > 16


def test_reference_style_database_gets_gpu_columns(tmp_path):
    """A database created with only the reference's columns is upgraded."""
    con = sqlite3.connect(str(tmp_path / "old.db"))
    cur = con.cursor()
    cur.execute("""CREATE TABLE 'variant' ('variant_id' INTEGER NOT NULL PRIMARY KEY AUTOINCREMENT,
        'blockA' INTEGER NULL, 'blockB' INTEGER NULL, 'prefetchDistance' INTEGER NULL, 'bandwidth' FLOAT NULL,
        'time' FLOAT NULL, 'tlbMisses' FLOAT NULL, 'l2misses' FLOAT NULL, 'rankLoopOrder' INTEGER NULL,
        'rankBlocking' INTEGER NULL, 'measurement_id' INTEGER NULL, 'loopOrder_id' INTEGER NULL)""")
    ttc.createTables(cur)
    cur.execute("PRAGMA table_info('variant')")
    cols = {r[1] for r in cur.fetchall()}
    assert {"tile", "vecA", "vecB", "config"} <= cols


def test_command_line_entry_point(workdir):
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run(["python", "-m", "ttc_amd", "--perm=0,2,1", "--size=8,9,10", "--dataType=sb",
                        "--generateOnly"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       cwd=str(workdir))
    assert r.returncode == 0, r.stdout
    assert (workdir / "ttc_transpositions" / "sbCuTranspose_021_8x9x10_bz.hip").exists()


def test_database_analytics(tmp_path):
    """Per-problem best bandwidths, winning configurations and the search's
    gain over the planner, from a database written by the insert helpers."""
    from ttc_amd import analytics
    db = str(tmp_path / "ttc.db")
    con = sqlite3.connect(db)
    cur = con.cursor()
    ttc.createTables(cur)
    lid = ttc.insertIntoLoopOrder(cur, [1, 0])
    cases = [([1, 0], [64, 48], "float", [(None, 1, 2, 1000.0), ({"tile": [32, 32]}, 1, 1, 1500.0)]),
             ([1, 0], [96, 80], "float", [(None, 2, 2, 2000.0), ({"tile": [16, 16]}, 1, 1, 1000.0)]),
             ([2, 0, 1], [8, 9, 10], "double", [(None, 2, 2, 800.0), ({"tile": [8, 9, 2]}, 2, 1, 1200.0)])]
    for perm, size, t, variants in cases:
        sid = ttc.insertIntoSize(cur, size)
        pid = ttc.insertIntoPermutation(cur, perm)
        mid = ttc.insertIntoMeasurements(cur, len(size), "h", 0, 1.0, 0.0, 0, "hipcc", t, t, 0, 0, 0, 0, sid, pid,
                                         -1, 0, 0, "gfx950")
        for rank, (cfg, fam, tord, bw) in enumerate(variants):
            info = {"run_a": 1, "run_b": 1, "tile": size, "vec_a": 4, "vec_b": 4, "threads": 256, "nt_store": 0,
                    "grid_order": 4, "family": fam, "tile_order": tord}
            ttc.insertIntoVariant(cur, mid, lid, rank, {"config": cfg or {}, "ms": 0.1, "GiBs": bw, "info": info})
    con.commit()
    con.close()
    s = analytics.summarize(db)
    assert s["measurements"] == 3
    rows = {(r["pair"], r["rank"]): r for r in s["by_problem"]}
    assert rows[("float->float", 2)]["n"] == 2 and rows[("float->float", 2)]["median"] == 1750.0
    assert rows[("double->double", 3)]["max"] == 1200.0
    assert s["planner"] == {"measurements": 3, "planner_fastest": 1, "median_gain": 1.5, "max_gain": 1.5}
    assert s["winners"][0]["count"] >= 1 and sum(w["count"] for w in s["winners"]) == 3
    text = analytics.render(s)
    assert "row x column" in text and "1 of 3 searches" in text


def test_stored_solution_for_other_leading_dims_is_not_reused(workdir):
    """The database key holds no lda/ldb (as the reference's); a
    configuration tuned on the dense problem whose vector width a padded
    lda rules out is dropped and the planner's choice emitted instead."""
    from ttc_amd import runtime as rt
    dense = rt.Plan(rt.S, rt.S, [1, 0], [1000, 768], beta_is_zero=False)
    cfg = {k: dense.info()[k] for k in ("tile", "vec_a", "vec_b", "family")}
    assert cfg["vec_a"] > 1  # a wide vector along A's 1000 rows
    with pytest.raises(rt.TTCError):
        rt.Plan(rt.S, rt.S, [1, 0], [1000, 768], [1001, 768], None, False, cfg)
    db = ttc.database_path()
    os.makedirs(os.path.dirname(db), exist_ok=True)
    con = sqlite3.connect(db)
    cur = con.cursor()
    ttc.createTables(cur)
    sid = ttc.insertIntoSize(cur, [1000, 768])
    pid = ttc.insertIntoPermutation(cur, [1, 0])
    mid = ttc.insertIntoMeasurements(cur, 2, "h", 0, 1.0, 1.0, 0, "hipcc", "float", "float", 0, 0, 0, 0, sid, pid,
                                     -1, 0, 0, ttc_util.getArchitecture("hip"))
    lid = ttc.insertIntoLoopOrder(cur, [1, 0])
    info = dict(dense.info(), threads=256)
    ttc.insertIntoVariant(cur, mid, lid, 0, {"config": cfg, "ms": 0.1, "GiBs": 999.0, "info": info})
    con.commit()
    con.close()
    # the dense call is served from the database ...
    name, bw = ttc.generateTransposition(ttc.parse_args(["--perm=1,0", "--size=1000,768", "--beta=1"]))
    assert bw == 999.0
    # ... the padded one falls back to the planner (no GPU: --generateOnly)
    a = ttc.parse_args(["--perm=1,0", "--size=1000,768", "--beta=1", "--lda=1001,768", "--generateOnly"])
    name, bw = ttc.generateTransposition(a)
    assert bw == 0.0
    src = (workdir / "ttc_transpositions" / (name + ".hip")).read_text()
    padded = rt.Plan(rt.S, rt.S, [1, 0], [1000, 768], [1001, 768], None, False).info()
    assert "{{%s}, %d, %d" % (", ".join(str(t) for t in padded["tile"] + [0] * 6), padded["vec_a"],
                               padded["vec_b"]) in src


def test_batch_rejects_short_buffers_before_any_device_call():
    """Batch checks every op's buffers against the extent its perm, size and
    lda/ldb span (the C ABI takes raw pointers and cannot), and keeps the
    tensors alive for replays."""
    import torch
    from oracle import oracle as orc
    from ttc_amd import runtime as rt
    perm, size, lda, ldb = [2, 0, 1], [5, 6, 7], [6, 6, 7], [7, 6, 8]
    sA, sB = orc.strides(perm, size, lda, ldb)
    last_a = sum((n - 1) * s for n, s in zip(size, sA))  # offset of the last element
    last_b = sum((n - 1) * s for n, s in zip(size, sB))
    assert rt.extent_bytes(rt.S, rt.D, perm, size, lda, ldb) == ((last_a + 1) * 4, (last_b + 1) * 8)
    na, nb = last_a + 1, last_b + 1
    assert rt.extent_bytes(rt.S, rt.S, [1, 0], [0, 9]) == (0, 0)
    good = dict(type_a=rt.S, type_b=rt.S, perm=perm, size=size, lda=lda, ldb=ldb,
                A=torch.zeros(na), B=torch.zeros(nb))
    with pytest.raises(ValueError, match="buffer B"):
        rt.Batch([good, dict(good, B=torch.zeros(nb - 1))])
    with pytest.raises(ValueError, match="buffer A"):
        rt.Batch([dict(good, A=torch.zeros(2 * na)[::2])])  # not contiguous
