"""Driver and command line of the MI355X transposition generator.

Counterpart of HPAC/TTC's ttc/ttc.py for its GPU path: the same flags
(ttc/ttc.py:1134-1141, 1181-1337), the same TTCargs record, the same driver
sequence — sanity checks, index fusion, database lookup, search, artefact
emission, database insert — and the same return value
``(transposeName, GiB/s)`` (ttc/ttc.py:614-1130).  What changes underneath:
no compiler runs per request and no subprocess is timed; the search times
prebuilt gfx950 kernels in-process (ttc_amd.autotune), and the emitted
artefact forwards to libttc_hip.so.

  python -m ttc_amd --perm=1,0,2 --size=1000,768,16 --beta=1 --dataType=s
"""
import copy
import itertools
import json
import os
import socket
import sqlite3
import sys
import time as _time
import traceback

from . import generator
from . import sql_util
from . import ttc_util

OKGREEN = '\033[92m'
FAIL = '\033[91m'
WARNING = '\033[93m'
ENDC = '\033[0m'

GPU_ARCHS = ("hip", "cuda")


def getTransposeName(ttcArgs):
    """Artefact (file) name, ttc/ttc.py:41-95 for the GPU compiler."""
    return generator.file_name(ttcArgs.floatTypeA, ttcArgs.floatTypeB, ttcArgs.idxPerm, ttcArgs.size,
                               ttcArgs.beta)


def printEpilog(transposeName, ttcArgs, functionName):
    """ttc/ttc.py:97-115, with the artefact's real signature."""
    print(OKGREEN + "[SUCCESS]" + ENDC + " Please find the generated code under ./ttc_transpositions/%s.hip\n"
          % transposeName)
    print("------------------ Usage ------------------")
    print("// 1) include header")
    print("#include \"%s.h\"" % transposeName)
    print("")
    print("// 2) execute transposition (A, B: device pointers; size, lda, ldb: host arrays)")
    if ttcArgs.beta != 0:
        print("%s(A, B, alpha, beta, size, lda, ldb);" % functionName)
    else:
        print("%s(A, B, alpha, size, lda, ldb);" % functionName)
    print("// 3) build: hipcc -c %s.hip -I<ttc-mi355x>/include ; link with -lttc_hip" % transposeName)
    print("-------------------------------------------")


def printHelp():
    print("Tensor Transpose Compiler for AMD Instinct MI355X (gfx950)\n")
    print("Usage: %s --perm=..., --size=... [optional arguments]\n" % sys.argv[0])
    print("Generates and tunes a GPU transposition of the form")
    print("   B_{Pi(i1,i2,..., iN)} = alpha * A_{i1,i2,..., iN} + beta * B_{Pi(i1,i2,..., iN)}\n")
    print("required arguments:")
    print("   --perm=<index1>,<index2>,...,<indexN>".ljust(60), "specify the permutation")
    print("   --size=<size1>,<size2>,...,<sizeN>".ljust(60), "size of the input tensor")
    print("")
    print("optional arguments:")
    print("   --lda=<lda1>,<lda2>,...,<ldaN>".ljust(60), "outer-size of each dimension of the input tensor")
    print("   --ldb=<ldb1>,<ldb2>,...,<ldbN>".ljust(60), "outer-size of each dimension of the output tensor")
    print("   --maxImplementations=<value>".ljust(60), "limit the number of measured variants")
    print("   ".ljust(14), "-> Default: 200; -1 denotes 'no limit'")
    print("   --alpha=<value>".ljust(60), "alpha value (default: 1.0)")
    print("   --beta=<value>".ljust(60), "beta value (default: 0.0)")
    print("   --dataType=[s,d,c,z,sd,ds,cz,zc,b,h,sb,sh,bs,hs,sq,sr,bq,br,hq,hr or typeA->typeB]".ljust(60),
          "element types ('xy': A has type x, B type y; b = bfloat16, h = half, q/r = fp8 e4m3/e5m2)")
    print("   --architecture=hip".ljust(60), "target (hip = gfx950; 'cuda' is accepted as an alias)")
    print("   --streamingStores".ljust(60), "try non-temporal B stores in the search (beta == 0 only)")
    print("   --threadsPerBlock=256".ljust(60), "work-group size (256 is the only value)")
    print("   --blockings=<a>x<b>[,...]".ljust(60), "restrict the search to tiles of these A x B run lengths")
    print("   --generateOnly".ljust(60), "emit the heuristic choice, no timing")
    print("   --keep".ljust(60), "accepted for compatibility")
    print("   --ignoreDatabase".ljust(60), "don't use the SQL database (i.e., no lookup)")
    print("   --hotA / --hotB".ljust(60), "time with A (B) kept in cache between launches (database keys)")
    print("   --verbose or -v".ljust(60), "verbose output")
    print("   --help".ljust(60), "prints this help")
    print("")
    print("Example: \"ttc --perm=1,0,2 --size=1000,768,16 --beta=1.0 --dataType=s\"")
    v = ttc_util.getVersion()
    print("Version:".ljust(20) + "v%d.%d.%d" % (v[0], v[1], v[2]))


# ------------------------------------------------------------------ database
def database_path():
    """$TTC_ROOT/ttc/ttc.db as the reference (ttc/ttc.py:636-655), else
    ~/.ttc_amd/ttc.db."""
    root = os.environ.get("TTC_ROOT")
    if root:
        return os.path.join(root, "ttc", "ttc.db")
    return os.path.join(os.path.expanduser("~"), ".ttc_amd", "ttc.db")


EXTRA_VARIANT_COLUMNS = [("tile", "VARCHAR(64)"), ("vecA", "INTEGER"), ("vecB", "INTEGER"),
                         ("threadsPerBlock", "INTEGER"), ("wgPerCU", "INTEGER"), ("ntStore", "INTEGER"),
                         ("gridOrder", "INTEGER"), ("runA", "INTEGER"), ("runB", "INTEGER"), ("config", "TEXT"),
                         ("family", "INTEGER"), ("tileOrder", "INTEGER")]


def createTables(cursor):
    """The reference schema (ttc/ttc.py:226-368) plus GPU configuration
    columns on 'variant'."""
    sizes = ",\n".join("'size%d' INTEGER NULL" % i for i in range(10))
    idxs = ",\n".join("'idx%d' INTEGER NULL" % i for i in range(10))
    cmds = [
        "CREATE TABLE IF NOT EXISTS 'size' ('size_id' INTEGER NOT NULL PRIMARY KEY AUTOINCREMENT, %s);" % sizes,
        "CREATE TABLE IF NOT EXISTS 'permutation' ('permutation_id' INTEGER NOT NULL PRIMARY KEY AUTOINCREMENT, "
        "%s);" % idxs,
        """CREATE TABLE IF NOT EXISTS 'measurements' (
          'measurement_id' INTEGER NOT NULL PRIMARY KEY AUTOINCREMENT, 'dim' INTEGER NULL,
          'host' VARCHAR(45) NULL, 'architecture' VARCHAR(45) NULL, 'version' VARCHAR(45) NULL,
          'alpha' FLOAT NULL, 'beta' FLOAT NULL, 'numThreads' INTEGER NULL, 'compiler_version' VARCHAR(60) NULL,
          'floatTypeA' VARCHAR(45) NULL, 'floatTypeB' VARCHAR(45) NULL, 'top1' FLOAT NULL, 'top5' FLOAT NULL,
          'compilationTime' FLOAT NULL, 'measuringTime' FLOAT NULL, 'referenceBandwidth' FLOAT NULL,
          'size_id' INTEGER NULL, 'permutation_id' INTEGER NULL, 'hotA' INTEGER NULL, 'hotB' INTEGER NULL,
          FOREIGN KEY ('size_id') REFERENCES 'size' ('size_id'),
          FOREIGN KEY ('permutation_id') REFERENCES 'permutation' ('permutation_id'));""",
        "CREATE TABLE IF NOT EXISTS 'loopOrder' ('loopOrder_id' INTEGER NOT NULL PRIMARY KEY AUTOINCREMENT, "
        "%s);" % idxs,
        """CREATE TABLE IF NOT EXISTS 'variant' (
          'variant_id' INTEGER NOT NULL PRIMARY KEY AUTOINCREMENT, 'blockA' INTEGER NULL, 'blockB' INTEGER NULL,
          'prefetchDistance' INTEGER NULL, 'bandwidth' FLOAT NULL, 'time' FLOAT NULL, 'tlbMisses' FLOAT NULL,
          'l2misses' FLOAT NULL, 'rankLoopOrder' INTEGER NULL, 'rankBlocking' INTEGER NULL,
          'measurement_id' INTEGER NULL, 'loopOrder_id' INTEGER NULL, %s,
          FOREIGN KEY ('measurement_id') REFERENCES 'measurements' ('measurement_id'),
          FOREIGN KEY ('loopOrder_id') REFERENCES 'loopOrder' ('loopOrder_id'));"""
        % ", ".join("'%s' %s NULL" % c for c in EXTRA_VARIANT_COLUMNS),
        """CREATE VIEW IF NOT EXISTS 'joined' AS SELECT * FROM measurements join size on measurements.size_id =
          size.size_id join permutation on measurements.permutation_id = permutation.permutation_id join variant
          on variant.measurement_id = measurements.measurement_id""",
        """CREATE VIEW IF NOT EXISTS 'fastest0Percent' AS select * from measurements join (
          select v2.* from (select *, max(bandwidth) as maxBandwidth from variant group by measurement_id) v1
          inner join variant v2 on v1.measurement_id = v2.measurement_id where v2.bandwidth >= (1.0 *
          v1.maxBandwidth)) fast on measurements.measurement_id = fast.measurement_id;""",
    ]
    for c in cmds:
        cursor.execute(c)
    # databases written by the reference lack the GPU columns
    cursor.execute("PRAGMA table_info('variant')")
    have = {r[1] for r in cursor.fetchall()}
    for name, typ in EXTRA_VARIANT_COLUMNS:
        if name not in have:
            cursor.execute("ALTER TABLE variant ADD COLUMN '%s' %s NULL" % (name, typ))


def _insert_dims(cursor, table, prefix, values):
    w, a = sql_util._where_dims(prefix, values)
    cursor.execute("SELECT %s_id from %s WHERE %s" % (table if table != "size" else "size", table, w), a)
    r = cursor.fetchall()
    if r:
        return r[0][0]
    cols = ", ".join("%s%d" % (prefix, i) for i in range(len(values)))
    cursor.execute("INSERT INTO %s (%s) VALUES (%s)" % (table, cols, ", ".join("?" * len(values))),
                   [int(v) for v in values])
    return cursor.lastrowid


def insertIntoSize(cursor, size):
    """ttc/ttc.py:452-495."""
    return _insert_dims(cursor, "size", "size", size)


def insertIntoPermutation(cursor, perm):
    """ttc/ttc.py:497-540."""
    w, a = sql_util._where_dims("idx", perm)
    cursor.execute("SELECT permutation_id from permutation WHERE %s" % w, a)
    r = cursor.fetchall()
    if r:
        return r[0][0]
    cols = ", ".join("idx%d" % i for i in range(len(perm)))
    cursor.execute("INSERT INTO permutation (%s) VALUES (%s)" % (cols, ", ".join("?" * len(perm))),
                   [int(v) for v in perm])
    return cursor.lastrowid


def insertIntoLoopOrder(cursor, loopOrder):
    """ttc/ttc.py:543-586."""
    w, a = sql_util._where_dims("idx", loopOrder)
    cursor.execute("SELECT loopOrder_id from loopOrder WHERE %s" % w, a)
    r = cursor.fetchall()
    if r:
        return r[0][0]
    cols = ", ".join("idx%d" % i for i in range(len(loopOrder)))
    cursor.execute("INSERT INTO loopOrder (%s) VALUES (%s)" % (cols, ", ".join("?" * len(loopOrder))),
                   [int(v) for v in loopOrder])
    return cursor.lastrowid


def insertIntoMeasurements(cursor, dim, host, version, alpha, beta, numThreads, compiler_version, floatTypeA,
                           floatTypeB, top1, top5, compilationTime, measuringTime, size_id, perm_id, referenceBw,
                           hotA, hotB, architecture):
    """ttc/ttc.py:423-450."""
    cursor.execute("""INSERT INTO measurements (dim, host, architecture, version, alpha, beta, numThreads,
        compiler_version, floatTypeA, floatTypeB, top1, top5, compilationTime, measuringTime, referenceBandwidth,
        size_id, permutation_id, hotA, hotB) VALUES (?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?)""",
                   (dim, host, architecture, version, alpha, beta, numThreads, compiler_version, floatTypeA,
                    floatTypeB, top1, top5, compilationTime, measuringTime, referenceBw, size_id, perm_id, hotA,
                    hotB))
    return cursor.lastrowid


def insertIntoVariant(cursor, measurement_id, loop_id, rank, result):
    """ttc/ttc.py:588-612 plus the GPU configuration.  blockA/blockB hold
    the tile's contiguous runs on A's and B's side."""
    info, cfg = result["info"], result["config"]
    cursor.execute("""INSERT INTO variant (blockA, blockB, prefetchDistance, bandwidth, time, tlbMisses, l2misses,
        rankLoopOrder, rankBlocking, measurement_id, loopOrder_id, tile, vecA, vecB, threadsPerBlock, wgPerCU,
        ntStore, gridOrder, runA, runB, config, family, tileOrder)
        VALUES (?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?,?)""",
                   (info["run_a"], info["run_b"], 0, result["GiBs"], result["ms"] * 1e-3, -1.0, -1.0, rank, rank,
                    measurement_id, loop_id, "x".join(str(t) for t in info["tile"]), info["vec_a"], info["vec_b"],
                    info["threads"], cfg.get("wg_per_cu", 0), info["nt_store"], info["grid_order"], info["run_a"],
                    info["run_b"], json.dumps(cfg), info.get("family", 0), info.get("tile_order", 0)))
    return cursor.lastrowid


# -------------------------------------------------------------------- driver
def _fail(msg):
    print(FAIL + "[TTC] ERROR: " + msg + ENDC)
    sys.exit(-1)


def generateTransposition(ttcArgs):
    """Returns (transposeName, bandwidth in GiB/s of the chosen variant, 0
    when nothing was measured).  Emits ./ttc_transpositions/<name>.hip/.h."""
    from . import runtime as rt
    if ttcArgs.architecture not in GPU_ARCHS:
        _fail("architecture '%s': ttc_amd generates GPU (gfx950) transpositions only; use "
              "--architecture=hip" % ttcArgs.architecture)
    if ttcArgs.architecture == "cuda" and not ttcArgs.silent:
        print(WARNING + "[TTC] WARNING: 'cuda' selected -> generating for HIP/gfx950 instead." + ENDC)
    ttcArgs.architecture = "hip"
    ttcArgs.compiler = "hipcc"
    try:
        ta = ttc_util.dtypeCode(ttcArgs.floatTypeA)
        tb = ttc_util.dtypeCode(ttcArgs.floatTypeB)
    except ValueError as e:
        _fail(str(e))
    if ttcArgs.maxNumImplementations == -1:
        ttcArgs.maxNumImplementations = 10000000
    if len(ttcArgs.vecLength) > 0 and any(v != 256 for v in ttcArgs.vecLength):
        _fail("--threadsPerBlock: ttc_amd's kernels run 256 threads per work-group")

    # ---------------------------------------------------- fuse indices
    sizeOld, permOld = copy.deepcopy(ttcArgs.size), copy.deepcopy(ttcArgs.idxPerm)
    ldaOld, ldbOld = copy.deepcopy(ttcArgs.lda), copy.deepcopy(ttcArgs.ldb)
    # sanity checks that the reference makes after fusion (ttc/ttc.py:733-768)
    # are made on the user's input here, where they are meaningful
    if len(ttcArgs.ldb) != 0 and len(ttcArgs.ldb) != len(ttcArgs.size):
        _fail("not all leading dimensions of B have been specified")
    if len(ttcArgs.lda) != 0 and len(ttcArgs.lda) != len(ttcArgs.size):
        _fail("not all leading dimensions of A have been specified")
    if len(ttcArgs.idxPerm) != len(ttcArgs.size):
        _fail("dimension of permutation and size does not match.")
    if sorted(ttcArgs.idxPerm) != list(range(len(ttcArgs.idxPerm))):
        _fail("permutation is invalid.")
    for i in range(len(ttcArgs.ldb)):
        if ttcArgs.size[ttcArgs.idxPerm[i]] > ttcArgs.ldb[i]:
            _fail("the leading dimension of B for dim %d is smaller than allowed (it must be >= %d)"
                  % (i, ttcArgs.size[ttcArgs.idxPerm[i]]))
    for i in range(len(ttcArgs.lda)):
        if ttcArgs.size[i] > ttcArgs.lda[i]:
            _fail("the leading dimension of A for dim %d is smaller than allowed (it must be >= %d)"
                  % (i, ttcArgs.size[i]))
    ttc_util.fuseIndices(ttcArgs.size, ttcArgs.idxPerm, ttcArgs.loopPermutations, ttcArgs.lda, ttcArgs.ldb)
    if len(sizeOld) != len(ttcArgs.size) and ttcArgs.silent != 1:
        print("Permutation and Size _before_ index-fusion: ", permOld, sizeOld, ldaOld, ldbOld)
        print("Permutation and Size _after_ index-fusion: ", ttcArgs.idxPerm, ttcArgs.size, ttcArgs.lda,
              ttcArgs.ldb)
    dim = len(ttcArgs.size)
    lda = ttcArgs.lda if ttcArgs.lda else None
    ldb = ttcArgs.ldb if ttcArgs.ldb else None
    bz = ttcArgs.beta == 0
    # --streamingStores: non-temporal B stores join the search (beta == 0
    # only, ttc/ttc_util.py:301-314); --blockings: the search only takes
    # tiles of these A x B run lengths (ttc/ttc.py:1296-1297, 827, 849)
    streaming = ttc_util.streamingStoresApplicable(ttcArgs.ldb, ttcArgs.size, ttcArgs.idxPerm, ttcArgs.beta, 1,
                                                   ttc_util.FLOAT_TYPES[ttcArgs.floatTypeB][1],
                                                   int(bool(ttcArgs.streamingStores)))
    blockings = [(int(a), int(b)) for a, b in (ttcArgs.blockings or [])]
    if any(a < 1 or b < 1 for a, b in blockings):
        _fail("--blockings: run lengths must be positive")
    try:
        base_plan = rt.Plan(ta, tb, ttcArgs.idxPerm, ttcArgs.size, lda, ldb, bz)
    except rt.TTCError as e:
        _fail("cannot plan this transposition: %s" % e)
    transposeName = getTransposeName(ttcArgs)
    functionName = generator.function_name(ttcArgs.floatTypeA, ttcArgs.floatTypeB, ttcArgs.idxPerm, ttcArgs.size,
                                           ttcArgs.beta)
    if ttcArgs.silent != 1:
        print("--------------Settings---------------------")
        print("Architecture: ".ljust(20) + ttc_util.getArchitecture("hip"))
        print("Data types: ".ljust(20) + "%s -> %s" % (ttcArgs.floatTypeA, ttcArgs.floatTypeB))
        print("Alpha / Beta:".ljust(20) + "%f / %f" % (ttcArgs.alpha, ttcArgs.beta))
        print("Database:".ljust(20) + database_path())
        v = ttc_util.getVersion()
        print("Version:".ljust(20) + "v%d.%d.%d" % (v[0], v[1], v[2]))
        print("-------------------------------------------")

    # --------------------------------------- check if a solution exists
    arch = ttc_util.getArchitecture("hip")
    dbp = database_path()
    os.makedirs(os.path.dirname(dbp), exist_ok=True)
    connection = sqlite3.connect(dbp)
    cursor = connection.cursor()
    createTables(cursor)
    best_cfg, bandwidth, solutionFound = None, 0.0, False
    if ttcArgs.ignoreDatabase == 0:
        sizeId = sql_util.getSizeId(cursor, ttcArgs.size)
        permId = sql_util.getPermId(cursor, ttcArgs.idxPerm) if sizeId != -1 else -1
        if permId != -1:
            mid = sql_util.getMeasurementId(cursor, sizeId, permId, ttcArgs.floatTypeA, ttcArgs.floatTypeB,
                                            ttcArgs.beta, 0, ttcArgs.hotA, ttcArgs.hotB, arch)
            if mid != -1:
                ret = sql_util.getBestConfig(cursor, mid)
                if ret != -1:
                    best_cfg, bandwidth = ret
                    solutionFound = True
                    # The key holds no lda/ldb (as the reference's, ttc/ttc.py:792-833):
                    # a stored tile / vector width may not fit this call's padding.
                    try:
                        rt.Plan(ta, tb, ttcArgs.idxPerm, ttcArgs.size, lda, ldb, bz, best_cfg)
                    except rt.TTCError:
                        best_cfg, bandwidth, solutionFound = None, 0.0, False
                        if ttcArgs.silent != 1:
                            print("A stored solution does not fit these leading dimensions; searching anew.")
                    if solutionFound and ttcArgs.silent != 1:
                        print("Solution already exists: generating the solution which was found previously.")

    # ------------------------------------------------------------ search
    results = []
    measuringTime = 0.0
    if not solutionFound and ttcArgs.generateOnly == 0:
        import torch
        if not torch.cuda.is_available():
            _fail("no GPU: use --generateOnly to emit the heuristic choice without timing")
        from . import autotune
        t0 = _time.time()

        def progress(i, n):
            if ttcArgs.silent != 1:
                sys.stdout.write("[TTC] %d out of %d implementations done.\r" % (i, n))
                sys.stdout.flush()
        cands_limit = ttcArgs.maxNumImplementations
        results = autotune.tune(ta, tb, ttcArgs.idxPerm, ttcArgs.size, lda, ldb, ttcArgs.alpha, ttcArgs.beta,
                                cands_limit, progress=progress, nt_stores=bool(streaming),
                                hot_a=bool(ttcArgs.hotA), hot_b=bool(ttcArgs.hotB),
                                blockings=blockings or None)
        measuringTime = _time.time() - t0
        if not results:
            _fail("no variant could be measured")
        best_cfg, bandwidth = results[0]["config"], results[0]["GiBs"]
        if ttcArgs.silent != 1:
            print("\nMeasuring took %f seconds" % measuringTime)
            print("\nThe fastest version (tile %s, vec %d/%d) attains %.2f GiB/s (%.1f GB/s).\n" % (
                results[0]["info"]["tile"], results[0]["info"]["vec_a"], results[0]["info"]["vec_b"],
                results[0]["GiBs"], results[0]["GBs"]))
    if best_cfg is None:  # --generateOnly without a stored solution: the planner's choice
        cfg0 = {"run_a": blockings[0][0], "run_b": blockings[0][1]} if blockings else {}
        if streaming:
            cfg0["nt_store"] = 1
        try:
            info = rt.Plan(ta, tb, ttcArgs.idxPerm, ttcArgs.size, lda, ldb, bz, cfg0 or None).info()
        except rt.TTCError:
            info = base_plan.info()
        best_cfg = {"tile": info["tile"], "vec_a": info["vec_a"], "vec_b": info["vec_b"], "family": info["family"],
                    "tile_order": info["tile_order"], "nt_store": 1 if info["nt_store"] else 0}

    # ----------------------------------------------------- emit artefact
    directory = os.path.join(os.getcwd(), "ttc_transpositions")
    generator.emit(directory, ttcArgs.floatTypeA, ttcArgs.floatTypeB, ttcArgs.idxPerm, ttcArgs.size, ttcArgs.beta,
                   best_cfg, note="tuned" if results or solutionFound else "heuristic")
    if ttcArgs.silent != 1:
        printEpilog(transposeName, ttcArgs, functionName)

    # --------------------------------------------------- update database
    if ttcArgs.updateDatabase and results:
        size_id = insertIntoSize(cursor, ttcArgs.size)
        perm_id = insertIntoPermutation(cursor, ttcArgs.idxPerm)
        mid = insertIntoMeasurements(cursor, dim, socket.gethostname(), 0, ttcArgs.alpha, ttcArgs.beta, 0,
                                     ttc_util.getCompilerVersion("hipcc"), ttcArgs.floatTypeA, ttcArgs.floatTypeB,
                                     0.0, 0.0, 0.0, measuringTime, size_id, perm_id, -1.0, ttcArgs.hotA,
                                     ttcArgs.hotB, arch)
        loop_id = insertIntoLoopOrder(cursor, list(range(dim))[::-1])
        for rank, r in enumerate(results):
            insertIntoVariant(cursor, mid, loop_id, rank, r)
    connection.commit()
    connection.close()
    return (transposeName, bandwidth)


# ---------------------------------------------------------------------- CLI
ALLOWED = ["--compiler", "--streamingStores", "--use-streamingStores", "--maxImplementations", "--help", "--alpha",
           "--beta", "--papi", "--size", "--perm", "--loopPerm", "--floatType", "--dataType", "--numThreads",
           "--generateOnly", "--prefetchDistances", "--keep", "--updateDatabase", "--dontCompile", "-v",
           "--blockings", "--noTest", "--no-align", "--no-vec", "--mpi", "--architecture", "--affinity", "--lda",
           "--ldb", "--ignoreDatabase", "--threadsPerBlock", "--hotA", "--hotB", "--verbose"]


def parse_args(argv):
    """argv (without the program name) -> TTCargs; mirrors
    ttc/ttc.py:1143-1374 for the flags that matter on the GPU."""
    perm, size, lda, ldb = [], [], [], []
    opts = {"alpha": 1.0, "beta": 0.0, "maxImpl": 200, "types": ("float", "float"), "arch": "hip",
            "generateOnly": 0, "keep": 0, "ignoreDatabase": 0, "hotA": 0, "hotB": 0, "debug": 0,
            "streamingStores": 0, "vecLength": [], "blockings": [], "loopPerms": [], "numThreads": 0}
    for arg in argv:
        key = arg.split("=")[0]
        if key not in ALLOWED:
            printHelp()
            print(FAIL + "Error: argument " + key + " not valid." + ENDC)
            sys.exit(-1)
        val = arg.split("=", 1)[1] if "=" in arg else None
        if arg == "--keep":
            opts["keep"] = 1
        elif arg == "--ignoreDatabase":
            opts["ignoreDatabase"] = 1
        elif arg in ("--use-streamingStores", "--streamingStores"):
            opts["streamingStores"] = 1
        elif arg in ("-v", "--verbose"):
            opts["debug"] = 1
        elif arg == "--generateOnly":
            opts["generateOnly"] = 1
        elif arg == "--hotA":
            opts["hotA"] = 1
        elif arg == "--hotB":
            opts["hotB"] = 1
        elif arg == "--help":
            printHelp()
            sys.exit(0)
        elif key == "--architecture":
            if val not in ("hip", "cuda", "avx", "avx512", "knc", "power"):
                print(FAIL + "[TTC] ERROR: Architecture unknown." + ENDC)
                sys.exit(-1)
            opts["arch"] = val
        elif key == "--compiler":
            if val not in ("hipcc", "nvcc", "g++", "icpc", "ibm"):
                print(FAIL + "[TTC] ERROR: unknown compiler choice." + ENDC)
                sys.exit(-1)
            if val in ("hipcc", "nvcc"):
                opts["arch"] = "hip"
        elif key in ("--dataType", "--floatType"):
            try:
                opts["types"] = ttc_util.dataTypeFromString(val)
            except ValueError:
                print("[TTC] ERROR: unknown precision. It needs to be one of s, d, c, z, sd, ds, cz, zc "
                      "(or ttc_amd's b, h, sb, sh, bs, hs, sq, sr, bq, br, hq, hr, typeA->typeB).")
                sys.exit(-1)
        elif key == "--numThreads":
            opts["numThreads"] = int(val)
        elif key == "--maxImplementations":
            opts["maxImpl"] = int(val)
        elif key == "--alpha":
            opts["alpha"] = float(val)
        elif key == "--beta":
            opts["beta"] = float(val)
        elif key == "--threadsPerBlock":
            for v in val.split(","):
                if int(v) not in (128, 256, 512):
                    print("[TTC] ERROR: value for --threadsPerBlock is not valid. It needs to be either 128, 256 "
                          "or 512.")
                    sys.exit(-1)
                opts["vecLength"].append(int(v))
        elif key == "--blockings":
            for b in val.split(","):
                opts["blockings"].append((int(b.split("x")[0]), int(b.split("x")[1])))
        elif key == "--size":
            size = [int(s) for s in val.split(",")]
        elif key == "--lda":
            lda = [int(s) for s in val.split(",")]
        elif key == "--ldb":
            ldb = [int(s) for s in val.split(",")]
        elif key == "--perm":
            perm = [int(s) for s in val.split(",")]
        elif key == "--loopPerm":
            for lp in val.split("-"):
                opts["loopPerms"].append([int(x) for x in lp.split(",")])
    if len(perm) != len(size):
        printHelp()
        print(FAIL + "[Error] dimension of permutation and size does not match." + ENDC)
        sys.exit(-1)
    a = ttc_util.TTCargs(perm, size)
    a.alpha, a.beta = opts["alpha"], opts["beta"]
    a.floatTypeA, a.floatTypeB = opts["types"]
    a.maxNumImplementations = opts["maxImpl"]
    a.lda, a.ldb = lda, ldb
    a.architecture = opts["arch"]
    a.generateOnly, a.keep, a.ignoreDatabase = opts["generateOnly"], opts["keep"], opts["ignoreDatabase"]
    a.hotA, a.hotB, a.debug = opts["hotA"], opts["hotB"], opts["debug"]
    a.streamingStores, a.vecLength, a.blockings = opts["streamingStores"], opts["vecLength"], opts["blockings"]
    a.loopPermutations, a.numThreads = opts["loopPerms"], opts["numThreads"]
    a.silent = 0
    return a


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 1:
        printHelp()
        sys.exit(0)
    args = parse_args(argv)
    return generateTransposition(args)


if __name__ == "__main__":
    main()
