"""Python-3 counterpart of HPAC/TTC's ttc/ttc_util.py for the GPU path.

Same names and argument meaning as the reference so that code written
against ``ttc.ttc_util`` (e.g. test-framework/ttc_test.py:124-145) reads the
same.  Floating-point type strings follow the reference ("float", "double",
"float complex", "double complex") plus the types ttc_amd adds
("bfloat16", "half", "fp8_e4m3", "fp8_e5m2").
"""
import copy
import os

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
    return [0, 1, 0]


def createTmpDirectory(base="tmp"):
    """ttc/ttc_util.py:13-20."""
    i = 0
    while os.path.exists(base + "%d" % i):
        i += 1
    d = base + "%d" % i
    os.makedirs(d)
    return d + "/"


def getFloatTypeSize(floatType):
    """Bytes per element (ttc/ttc_util.py:23-34)."""
    if floatType not in FLOAT_TYPES:
        raise ValueError("[TTC] ERROR: unknown data type %r" % (floatType,))
    return FLOAT_TYPES[floatType][1]


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


def getCostLoop(loopPerm, perm, size):
    """Loop-order cost model, restated from ttc/ttc_util.py:167-209:
    sum over loops of totalSize*(extent-1)*((lastOffsetA-lda0)^2 +
    (lastOffsetB-ldb0)^2) with dense strides."""
    lda = [1]
    ldb = [1]
    for i in range(1, len(size)):
        lda.append(lda[i - 1] * size[i - 1])
        ldb.append(ldb[i - 1] * size[perm[i - 1]])
    totalSize = 1
    cost = 0
    for i in range(len(loopPerm)):
        loopIdx = loopPerm[i]
        idxB = perm.index(loopIdx)
        lda0 = lda[loopIdx]
        ldb0 = ldb[idxB]
        lastOffsetA = 0
        lastOffsetB = 0
        for l in range(i + 1, len(loopPerm)):
            nxt = loopPerm[l]
            lastOffsetA += lda[nxt] * (size[nxt] - 1)
            lastOffsetB += ldb[perm.index(nxt)] * (size[nxt] - 1)
        cost += totalSize * (size[loopIdx] - 1) * ((lastOffsetA - lda0) ** 2. + (lastOffsetB - ldb0) ** 2.)
        totalSize *= size[loopIdx]
    return cost


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
    """ttc/ttc_util.py:301-314 (CPU rule; on the GPU, beta == 0 selects
    non-temporal stores unless disabled)."""
    leadingDimension = size[perm[0]]
    if len(ldb) > 0:
        leadingDimension = ldb[0]
    if useStreamingStores == 1 and beta == 0:
        if perm[0] == 0:
            return 1
        if (leadingDimension * floatSizeBytes) % alignmentRequirementBytes == 0:
            return 1
        return 0
    return 0
