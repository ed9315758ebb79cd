"""Python-3 counterpart of HPAC/TTC's ttc/ttc_util.py for the GPU path.

Same names and argument meaning as the reference so that code written
against ``ttc.ttc_util`` (e.g. test-framework/ttc_test.py:124-145) reads the
same.  Floating-point type strings follow the reference ("float", "double",
"float complex", "double complex") plus the types ttc_amd adds
("bfloat16", "half", "fp8_e4m3", "fp8_e5m2").
"""
import copy

OKGREEN = '\033[92m'
FAIL = '\033[91m'
WARNING = '\033[93m'
ENDC = '\033[0m'

# type string -> (ttc_dtype_t code, bytes, prefix letter)
FLOAT_TYPES = {
    "float": (0, 4, "s"),
    "double": (1, 8, "d"),
    "float complex": (2, 8, "c"),
    "double complex": (3, 16, "z"),
    "bfloat16": (4, 2, "b"),
    "half": (5, 2, "h"),
    "fp8_e4m3": (6, 1, "q"),
    "fp8_e5m2": (7, 1, "r"),
}


def getVersion():
    """ttc/ttc_util.py:10-11 (ttc_amd's own version)."""
    return [0, 2, 0]


def dtypeCode(floatType):
    """ttc_dtype_t code of a type string."""
    if floatType not in FLOAT_TYPES:
        raise ValueError("[TTC] ERROR: unknown data type %r" % (floatType,))
    return FLOAT_TYPES[floatType][0]


def listToString(l):
    """ttc/ttc_util.py:37-44."""
    if len(l) == 0:
        return ""
    return "(" + ",".join(str(x) for x in l) + ")"


# --dataType codes: the reference's (ttc/ttc.py:1253-1280) and ttc_amd's.
DATA_TYPES = {
    "s": ("float", "float"), "d": ("double", "double"), "c": ("float complex", "float complex"),
    "z": ("double complex", "double complex"), "sd": ("float", "double"), "ds": ("double", "float"),
    "cz": ("float complex", "double complex"), "zc": ("double complex", "float complex"),
    # additions (no counterpart in the reference)
    "b": ("bfloat16", "bfloat16"), "h": ("half", "half"), "sb": ("float", "bfloat16"), "sh": ("float", "half"),
    "bs": ("bfloat16", "float"), "hs": ("half", "float"),
    "sq": ("float", "fp8_e4m3"), "sr": ("float", "fp8_e5m2"), "bq": ("bfloat16", "fp8_e4m3"),
    "br": ("bfloat16", "fp8_e5m2"), "hq": ("half", "fp8_e4m3"), "hr": ("half", "fp8_e5m2"),
}


def dataTypeFromString(code):
    """Resolve a --dataType value: a code above, or "typeA->typeB"."""
    if code in DATA_TYPES:
        return DATA_TYPES[code]
    if "->" in code:
        a, b = [x.strip() for x in code.split("->", 1)]
        if a in FLOAT_TYPES and b in FLOAT_TYPES:
            return (a, b)
    raise ValueError("[TTC] ERROR: unknown precision %r" % code)


class TTCargs:
    """Argument record (ttc/ttc_util.py:46-114).  Same attributes; the
    defaults that only made sense for the CPU generator are kept so that
    callers setting them keep working.  architecture defaults to "hip"."""

    def __init__(self, idxPerm, size):
        self.size = copy.deepcopy(size)
        self.idxPerm = copy.deepcopy(idxPerm)
        self.alpha = 1.
        self.keep = 0
        self.generateOnly = 0
        self.beta = 0.
        self.affinity = ""
        self.numThreads = 0
        self.floatTypeA = "float"
        self.floatTypeB = "float"
        self.streamingStores = 0
        self.maxNumImplementations = 200
        self.lda = []
        self.ldb = []
        self.debug = 0
        self.architecture = "hip"
        self.align = 1
        self.blockings = []
        self.loopPermutations = []
        self.prefetchDistances = []
        self.scalar = 0
        self.silent = 0
        self.ignoreDatabase = 0
        self.compiler = "hipcc"
        self.hostName = ""
        self.vecLength = []
        self.hotA = 0
        self.hotB = 0
        self.updateDatabase = 1

    def getSize(self):
        return copy.deepcopy(self.size)

    def getPerm(self):
        return copy.deepcopy(self.idxPerm)

    def getCommandLineString(self):
        """ttc/ttc_util.py:84-114."""
        dataType = getFloatPrefix(self.floatTypeA, self.floatTypeB)
        streamingStores = " --streamingStores" if self.streamingStores else ""
        permStr = listToString(self.idxPerm)[1:-1]
        sizeStr = listToString(self.size)[1:-1]
        ldaStr = "--lda=%s" % listToString(self.lda)[1:-1] if len(self.lda) > 0 else ""
        ldbStr = "--ldb=%s" % listToString(self.ldb)[1:-1] if len(self.ldb) > 0 else ""
        return ("ttc --alpha=%f --beta=%f --perm=%s --size=%s --architecture=%s --maxImplementations=%d %s %s "
                "--dataType=%s%s" % (self.alpha, self.beta, permStr, sizeStr, self.architecture,
                                     self.maxNumImplementations, ldaStr, ldbStr, dataType, streamingStores))


def getFloatPrefix(floatTypeA, floatTypeB):
    """ttc/ttc_util.py:117-133; for the added types the prefix is the
    matching --dataType code."""
    for code, pair in DATA_TYPES.items():
        if pair == (floatTypeA, floatTypeB):
            return code
    return FLOAT_TYPES[floatTypeA][2] + FLOAT_TYPES[floatTypeB][2]


def getArchitecture(arch):
    """Architecture key of the database (ttc/ttc_util.py:145-165).  For the
    GPU path: the device's gfx name and compute-unit count."""
    if arch in ("hip", "cuda"):
        try:
            import torch
            if torch.cuda.is_available():
                p = torch.cuda.get_device_properties(0)
                return "%s:%dCU" % (getattr(p, "gcnArchName", "gfx").split(":")[0], p.multi_processor_count)
        except Exception:
            pass
        return "gfx950"
    return "host"


def walkCost(loops):
    """TTC's loop-order cost (ttc/ttc_util.py:167-209) of a loop nest given
    as (extent, A stride, B stride) per loop, outermost first.  Whenever
    loop i advances, every loop inside it wraps from its last iteration
    back to 0, so the address jumps by  stride_i - sum_inner stride_l*(n_l-1)
    in A and likewise in B; such advances happen  (n_i - 1) * prod_outer n
    times.  The cost is the sum of the squared A and B jumps over all of
    them: orders whose inner loops stay near the previous address score low."""
    cost = 0.0
    count = 1
    for i, (n, sa, sb) in enumerate(loops):
        wrap_a = sum(sa_l * (n_l - 1) for n_l, sa_l, _ in loops[i + 1:])
        wrap_b = sum(sb_l * (n_l - 1) for n_l, _, sb_l in loops[i + 1:])
        cost += count * (n - 1) * (float(sa - wrap_a) ** 2 + float(sb - wrap_b) ** 2)
        count *= n
    return cost


def getCostLoop(loopPerm, perm, size, lda=None, ldb=None):
    """Cost of visiting the elements of the transposition (perm, size) with
    the loops in loopPerm (A index labels, outermost first): walkCost over
    their element strides in A and in B (dense, or from lda / ldb as
    ttc/transpose.py:443-473 defines them)."""
    d = len(size)
    LA = list(lda) if lda else list(size)
    LB = list(ldb) if ldb else [size[perm[q]] for q in range(d)]
    sA = [1] * d
    for k in range(1, d):
        sA[k] = sA[k - 1] * LA[k - 1]
    sBpos = [1] * d
    for q in range(1, d):
        sBpos[q] = sBpos[q - 1] * LB[q - 1]
    sB = [sBpos[perm.index(k)] for k in range(d)]  # B stride of A index k
    return walkCost([(size[k], sA[k], sB[k]) for k in loopPerm])


def fuseIndices(size, perm, loopPermutations, lda, ldb):
    """Index fusion, in place on the caller's lists (ttc/ttc_util.py:214-276).

    Runs of A indices that appear consecutively in perm are merged when every
    dimension of the run but the last is unpadded in A and in B.  As in the
    reference, an empty lda/ldb list stays empty for the caller (the dense
    defaults are local), and loop orders are shifted to the fused rank.
    """
    if len(lda) == 0:
        lda = list(size)
    if len(ldb) == 0:
        ldb = [size[perm[i]] for i in range(len(size))]
    inputIndices = list(range(len(size)))
    n = len(inputIndices)
    for i in range(n):
        for j in range(n):
            l = 0
            newSize = 1
            while (i + l < len(inputIndices) and j + l < len(perm) and inputIndices[i + l] == perm[j + l]
                   and (l == 0 or lda[i + l - 1] == size[i + l - 1])
                   and (l == 0 or ldb[j + l - 1] == size[i + l - 1])):
                newSize *= size[i + l]
                l += 1
            if l >= 1:
                size[i] = newSize
            while l > 1:
                inputIndices.pop(i + 1)
                perm.pop(j + 1)
                size.pop(i + 1)
                lda[i] *= lda[i + 1]
                lda.pop(i + 1)
                ldb[j] *= ldb[j + 1]
                ldb.pop(j + 1)
                l -= 1
    # close the gaps left in the labels: each label becomes its rank
    for lst in (inputIndices, perm):
        order = sorted(lst)
        rank = {v: r for r, v in enumerate(order)}
        for q in range(len(lst)):
            lst[q] = rank[lst[q]]
    for loopOrder in loopPermutations:
        diff = len(loopOrder) - len(perm)
        removeList = []
        for i in range(len(loopOrder)):
            loopOrder[i] = loopOrder[i] - diff
            if loopOrder[i] < 0:
                removeList.append(loopOrder[i])
        for r in removeList:
            loopOrder.remove(r)


def getCompilerVersion(compiler):
    """First line of `<compiler> --version` (ttc/ttc_util.py:278-299)."""
    import subprocess
    comp = {"hipcc": "hipcc", "nvcc": "hipcc", "g++": "g++", "icpc": "icpc"}.get(compiler, compiler)
    try:
        out = subprocess.run([comp, "--version"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True).stdout
    except OSError:
        print(FAIL + "[TTC] ERROR: compiler (%s) does not seem to be available." % compiler + ENDC)
        return -1
    return out.split("\n")[0]


def streamingStoresApplicable(ldb, size, perm, beta, alignmentRequirementBytes, floatSizeBytes,
                              useStreamingStores):
    """1 when B's stores may be non-temporal (ttc/ttc_util.py:301-314): they
    must be asked for (--streamingStores) and beta must be 0, so that B is
    write-only; B's leading dimension must also be a multiple of the store
    alignment unless dimension 0 stays in place.  The GPU's non-temporal
    buffer stores take any alignment (alignmentRequirementBytes = 1)."""
    if not useStreamingStores or beta != 0:
        return 0
    lead = ldb[0] if len(ldb) > 0 else size[perm[0]]
    return 1 if perm[0] == 0 or (lead * floatSizeBytes) % alignmentRequirementBytes == 0 else 0
