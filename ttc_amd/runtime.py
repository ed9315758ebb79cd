"""ctypes binding of libttc_hip.so (include/ttc_hip.h).

The library is the product; this module only marshals arguments.  Device
memory and streams come from the caller (PyTorch-ROCm tensors and streams in
the tests and benchmark).  torch is imported before the library is loaded so
that both share one HIP runtime (the library's libamdhip64.so.7 dependency
resolves to the copy torch has already loaded).
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libttc_hip.so")

MAX_DIM = 8
MAX_INPUT_DIM = 16

# ttc_dtype_t
S, D, C, Z, BF16, F16, E4M3, E5M2 = range(8)
DTYPE_NAMES = {S: "s", D: "d", C: "c", Z: "z", BF16: "bf16", F16: "f16", E4M3: "e4m3", E5M2: "e5m2"}
DTYPE_BYTES = {S: 4, D: 8, C: 8, Z: 16, BF16: 2, F16: 2, E4M3: 1, E5M2: 1}

# ttc_status_t
SUCCESS, ERROR_INVALID_VALUE, ERROR_UNSUPPORTED, ERROR_HIP, ERROR_ALLOC, ERROR_NO_DEVICE, ERROR_WORKSPACE = range(7)

BATCH_WORKSPACE_BYTES = 16 << 20  # TTC_BATCH_WORKSPACE_BYTES

SCALE_DELAYED, SCALE_CURRENT = 0, 1


class TTCError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = _lib_status_string(status)
        super().__init__("%s: %s (status %d)" % (where or "ttc", msg, status))


class Config(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int * MAX_DIM), ("vec_a", ctypes.c_int), ("vec_b", ctypes.c_int),
                ("wg_per_cu", ctypes.c_int), ("nt_store", ctypes.c_int), ("grid_order", ctypes.c_int),
                ("run_a", ctypes.c_int), ("run_b", ctypes.c_int), ("family", ctypes.c_int),
                ("tile_order", ctypes.c_int), ("lds_buffers", ctypes.c_int)]

    @classmethod
    def make(cls, tile=None, vec_a=0, vec_b=0, wg_per_cu=0, nt_store=0, grid_order=0, run_a=0, run_b=0, family=0,
             tile_order=0, lds_buffers=0):
        c = cls()
        if tile:
            for k, t in enumerate(tile):
                c.tile[k] = int(t)
        c.vec_a, c.vec_b, c.wg_per_cu, c.nt_store = int(vec_a), int(vec_b), int(wg_per_cu), int(nt_store)
        c.grid_order, c.run_a, c.run_b, c.family = int(grid_order), int(run_a), int(run_b), int(family)
        c.tile_order, c.lds_buffers = int(tile_order), int(lds_buffers)
        return c

    def as_dict(self):
        return {"tile": list(self.tile), "vec_a": self.vec_a, "vec_b": self.vec_b, "wg_per_cu": self.wg_per_cu,
                "nt_store": self.nt_store, "grid_order": self.grid_order, "run_a": self.run_a, "run_b": self.run_b,
                "family": self.family, "tile_order": self.tile_order, "lds_buffers": self.lds_buffers}


class PlanInfo(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int), ("perm", ctypes.c_int * MAX_DIM), ("size", ctypes.c_int * MAX_DIM),
                ("stride_a", ctypes.c_int64 * MAX_DIM), ("stride_b", ctypes.c_int64 * MAX_DIM),
                ("tile", ctypes.c_int * MAX_DIM), ("vec_a", ctypes.c_int), ("vec_b", ctypes.c_int),
                ("threads", ctypes.c_int), ("slots_a", ctypes.c_int), ("slots_b", ctypes.c_int),
                ("tiles", ctypes.c_int64), ("tile_volume", ctypes.c_int), ("lds_bytes", ctypes.c_int),
                ("lds_pad", ctypes.c_int), ("nt_store", ctypes.c_int), ("grid_order", ctypes.c_int),
                ("wg_per_cu", ctypes.c_int), ("elements", ctypes.c_int64), ("bytes_algorithmic", ctypes.c_int64),
                ("type_a", ctypes.c_int), ("type_b", ctypes.c_int), ("beta_is_zero", ctypes.c_int),
                ("run_a", ctypes.c_int), ("run_b", ctypes.c_int), ("family", ctypes.c_int),
                ("rows", ctypes.c_int), ("cols", ctypes.c_int), ("tile_order", ctypes.c_int),
                ("lds_buffers", ctypes.c_int), ("lane_fixed", ctypes.c_int)]

    def as_dict(self):
        r = self.rank
        return {"rank": r, "perm": list(self.perm[:r]), "size": list(self.size[:r]),
                "stride_a": list(self.stride_a[:r]), "stride_b": list(self.stride_b[:r]),
                "tile": list(self.tile[:r]), "vec_a": self.vec_a, "vec_b": self.vec_b, "threads": self.threads,
                "slots_a": self.slots_a, "slots_b": self.slots_b, "tiles": self.tiles,
                "tile_volume": self.tile_volume, "lds_bytes": self.lds_bytes, "lds_pad": self.lds_pad,
                "nt_store": self.nt_store, "grid_order": self.grid_order, "wg_per_cu": self.wg_per_cu,
                "elements": self.elements, "bytes_algorithmic": self.bytes_algorithmic,
                "type_a": self.type_a, "type_b": self.type_b, "beta_is_zero": self.beta_is_zero,
                "run_a": self.run_a, "run_b": self.run_b, "family": self.family, "rows": self.rows,
                "cols": self.cols, "tile_order": self.tile_order, "lds_buffers": self.lds_buffers,
                "lane_fixed": self.lane_fixed}


class BatchOp(ctypes.Structure):
    _fields_ = [("type_a", ctypes.c_int), ("type_b", ctypes.c_int), ("dim", ctypes.c_int),
                ("perm", ctypes.c_int * MAX_INPUT_DIM), ("size", ctypes.c_int * MAX_INPUT_DIM),
                ("lda", ctypes.c_int * MAX_INPUT_DIM), ("ldb", ctypes.c_int * MAX_INPUT_DIM),
                ("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("alpha", ctypes.c_double),
                ("beta", ctypes.c_double), ("beta_is_zero", ctypes.c_int), ("cfg", Config)]


# Every symbol include/ttc_hip.h declares (checked by the CPU tests).
EXPORTS = [
    "ttc_version", "ttc_status_string", "ttc_dtype_size", "ttc_fuse_indices", "ttc_plan_create",
    "ttc_plan_get_info", "ttc_plan_destroy", "ttc_execute", "ttc_execute_fp8", "ttc_legacy_transpose",
    "ttc_batch_create", "ttc_batch_execute", "ttc_batch_graph_launch", "ttc_batch_execute_loop",
    "ttc_batch_get_info", "ttc_batch_destroy", "ttc_batch_execute_noalloc",
]

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libttc_hip.so is not built (%s); run `python -m ttc_amd.build`" % LIB_PATH)
    try:
        import torch  # noqa: F401  (shares torch's HIP runtime)
    except Exception:
        pass
    L = ctypes.CDLL(LIB_PATH)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    L.ttc_version.restype = ctypes.c_char_p
    L.ttc_status_string.restype = ctypes.c_char_p
    L.ttc_status_string.argtypes = [ctypes.c_int]
    L.ttc_dtype_size.argtypes = [ctypes.c_int]
    L.ttc_dtype_size.restype = ctypes.c_int
    L.ttc_fuse_indices.argtypes = [ctypes.c_int, ip, ip, ip, ip, ip, ip, ip, ip, ip]
    L.ttc_fuse_indices.restype = ctypes.c_int
    L.ttc_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, ip, ip,
                                  ctypes.c_int, ctypes.POINTER(Config)]
    L.ttc_plan_create.restype = ctypes.c_int
    L.ttc_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
    L.ttc_plan_get_info.restype = ctypes.c_int
    L.ttc_plan_destroy.argtypes = [vp]
    L.ttc_plan_destroy.restype = ctypes.c_int
    L.ttc_execute.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double, vp]
    L.ttc_execute.restype = ctypes.c_int
    L.ttc_execute_fp8.argtypes = [vp, vp, vp, ctypes.c_double, vp, ctypes.c_float, vp, vp, ctypes.c_int, vp]
    L.ttc_execute_fp8.restype = ctypes.c_int
    L.ttc_legacy_transpose.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ip, ip, ip, ip, ctypes.POINTER(Config),
                                       ctypes.c_char_p]
    L.ttc_legacy_transpose.restype = None
    L.ttc_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(BatchOp), ctypes.c_int]
    L.ttc_batch_create.restype = ctypes.c_int
    L.ttc_batch_execute.argtypes = [vp, vp]
    L.ttc_batch_execute.restype = ctypes.c_int
    L.ttc_batch_graph_launch.argtypes = [vp, vp, ctypes.c_int]
    L.ttc_batch_graph_launch.restype = ctypes.c_int
    L.ttc_batch_execute_loop.argtypes = [vp, vp]
    L.ttc_batch_execute_loop.restype = ctypes.c_int
    L.ttc_batch_get_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ip, ctypes.POINTER(ctypes.c_int64)]
    L.ttc_batch_get_info.restype = ctypes.c_int
    L.ttc_batch_destroy.argtypes = [vp]
    L.ttc_batch_destroy.restype = ctypes.c_int
    L.ttc_batch_execute_noalloc.argtypes = [ctypes.c_int, ctypes.POINTER(BatchOp), ctypes.c_int, vp]
    L.ttc_batch_execute_noalloc.restype = ctypes.c_int
    _lib = L
    return L


def lib():
    return _load()


def _lib_status_string(status):
    try:
        return _load().ttc_status_string(int(status)).decode()
    except Exception:
        return "status %d" % status


def _check(status, where):
    if status != SUCCESS:
        raise TTCError(status, where)


def _iarr(v):
    if v is None or len(v) == 0:
        return None
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def version():
    return _load().ttc_version().decode()


def fuse_indices(perm, size, lda=None, ldb=None):
    """C-ABI index fusion -> (perm, size, lda, ldb) of the fused problem."""
    d = len(perm)
    od = ctypes.c_int(0)
    op_, os_, ol, ob = [(ctypes.c_int * MAX_INPUT_DIM)() for _ in range(4)]
    _check(_load().ttc_fuse_indices(d, _iarr(perm), _iarr(size), _iarr(lda), _iarr(ldb), ctypes.byref(od),
                                    op_, os_, ol, ob), "ttc_fuse_indices")
    r = od.value
    return list(op_[:r]), list(os_[:r]), list(ol[:r]), list(ob[:r])


def _ptr(x):
    """Device pointer of a torch tensor, or an int address."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Plan:
    """An immutable transposition plan (ttc_plan_create)."""

    def __init__(self, type_a, type_b, perm, size, lda=None, ldb=None, beta_is_zero=False, config=None):
        L = _load()
        self._h = ctypes.c_void_p()
        cfg = config if isinstance(config, Config) else (Config.make(**config) if config else None)
        self.type_a, self.type_b = int(type_a), int(type_b)
        self.perm, self.size = list(perm), list(size)
        self.lda = list(lda) if lda else None
        self.ldb = list(ldb) if ldb else None
        self.beta_is_zero = bool(beta_is_zero)
        st = L.ttc_plan_create(ctypes.byref(self._h), self.type_a, self.type_b, len(perm), _iarr(perm),
                               _iarr(size), _iarr(lda), _iarr(ldb), int(self.beta_is_zero),
                               ctypes.byref(cfg) if cfg is not None else None)
        _check(st, "ttc_plan_create")
        self._lib = L

    def info(self):
        i = PlanInfo()
        _check(self._lib.ttc_plan_get_info(self._h, ctypes.byref(i)), "ttc_plan_get_info")
        return i.as_dict()

    def extent_bytes(self):
        """Bytes spanned by A and by B (last element offset + 1), from the
        plan's strides; the C ABI takes raw pointers and cannot check."""
        if getattr(self, "_extent", None) is None:
            i = self.info()
            if i["elements"] == 0:
                self._extent = (0, 0)
            else:
                ea = sum((n - 1) * s for n, s in zip(i["size"], i["stride_a"])) + 1
                eb = sum((n - 1) * s for n, s in zip(i["size"], i["stride_b"])) + 1
                self._extent = (ea * DTYPE_BYTES[self.type_a], eb * DTYPE_BYTES[self.type_b])
        return self._extent

    def _check_buffers(self, A, B):
        ea, eb = self.extent_bytes()
        for name, t, need in (("A", A, ea), ("B", B, eb)):
            if t is not None and not isinstance(t, int):
                have = t.numel() * t.element_size() - (0 if t.is_contiguous() else need)
                if not t.is_contiguous() or have < need:
                    raise ValueError("buffer %s holds %d bytes (contiguous: %s) but the plan spans %d bytes"
                                     % (name, t.numel() * t.element_size(), t.is_contiguous(), need))

    def execute(self, A, B, alpha=1.0, beta=0.0, stream=None):
        self._check_buffers(A, B)
        _check(self._lib.ttc_execute(self._h, _ptr(A), _ptr(B), float(alpha), float(beta), _stream(stream)),
               "ttc_execute")

    def execute_fp8(self, A, B, alpha=1.0, scale=1.0, scale_in=None, amax_out=None, scale_inv_out=None,
                    mode=SCALE_DELAYED, stream=None):
        self._check_buffers(A, B)
        _check(self._lib.ttc_execute_fp8(self._h, _ptr(A), _ptr(B), float(alpha), _ptr(scale_in), float(scale),
                                         _ptr(amax_out), _ptr(scale_inv_out), int(mode), _stream(stream)),
               "ttc_execute_fp8")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ttc_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def legacy_transpose(type_a, type_b, A, B, alpha, beta, perm, size, lda=None, ldb=None, beta_is_zero=None,
                     config=None, name=None):
    """Synchronous drop-in path (exits the process on error, as the reference)."""
    if beta_is_zero is None:
        beta_is_zero = beta == 0
    cfg = config if isinstance(config, Config) else (Config.make(**config) if config else None)
    _load().ttc_legacy_transpose(int(type_a), int(type_b), _ptr(A), _ptr(B), float(alpha), float(beta),
                                 int(bool(beta_is_zero)), len(perm), _iarr(perm), _iarr(size), _iarr(lda),
                                 _iarr(ldb), ctypes.byref(cfg) if cfg is not None else None,
                                 name.encode() if name else None)


def extent_bytes(type_a, type_b, perm, size, lda=None, ldb=None):
    """Bytes spanned by A and by B (last element offset + 1) of an unfused
    problem: A's stride of dimension k is prod_{m<k} lda[m], B's stride of
    position q is prod_{m<q} ldb[m] (NULL = dense), as the C ABI reads them."""
    d = len(size)
    if any(int(s) == 0 for s in size):
        return 0, 0
    LA = list(lda) if lda else list(size)
    LB = list(ldb) if ldb else [size[perm[q]] for q in range(d)]
    ea = eb = 1
    sa = sb = 1
    for k in range(d):
        ea += (int(size[k]) - 1) * sa
        sa *= int(LA[k])
        eb += (int(size[perm[k]]) - 1) * sb
        sb *= int(LB[k])
    return ea * DTYPE_BYTES[int(type_a)], eb * DTYPE_BYTES[int(type_b)]


def _check_tensor(name, t, need, where):
    """A torch tensor handed over as a raw pointer must be contiguous and
    hold `need` bytes (an int address cannot be checked)."""
    if t is None or isinstance(t, int):
        return
    have = t.numel() * t.element_size()
    if not t.is_contiguous() or have < need:
        raise ValueError("%s: buffer %s holds %d bytes (contiguous: %s) but the op spans %d bytes"
                         % (where, name, have, t.is_contiguous(), need))


def _batch_ops_array(ops, where):
    """ttc_batch_op_t array of op dicts, after checking every tensor against
    the extent its op spans."""
    for i, o in enumerate(ops):
        ea, eb = extent_bytes(o["type_a"], o.get("type_b", o["type_a"]), o["perm"], o["size"], o.get("lda"),
                              o.get("ldb"))
        _check_tensor("A", o.get("A"), ea, "%s op %d" % (where, i))
        _check_tensor("B", o.get("B"), eb, "%s op %d" % (where, i))
    arr = (BatchOp * len(ops))()
    for i, o in enumerate(ops):
        b = arr[i]
        b.type_a, b.type_b = int(o["type_a"]), int(o.get("type_b", o["type_a"]))
        b.dim = len(o["perm"])
        for k in range(b.dim):
            b.perm[k] = int(o["perm"][k])
            b.size[k] = int(o["size"][k])
            if o.get("lda"):
                b.lda[k] = int(o["lda"][k])
            if o.get("ldb"):
                b.ldb[k] = int(o["ldb"][k])
        b.A, b.B = _ptr(o["A"]), _ptr(o["B"])
        b.alpha, b.beta = float(o.get("alpha", 1.0)), float(o.get("beta", 0.0))
        b.beta_is_zero = int(bool(o.get("beta_is_zero", o.get("beta", 0.0) == 0)))
        cfg = o.get("config")
        if cfg:
            b.cfg = cfg if isinstance(cfg, Config) else Config.make(**cfg)
    return arr


def batch_execute_noalloc(ops, stream=None, wg_per_cu=0):
    """One grouped launch of `ops` (dicts as for Batch) through
    ttc_batch_execute_noalloc: no device allocation, the work table goes to
    the library's fixed per-GPU workspace (BATCH_WORKSPACE_BYTES).  Calls on
    one GPU that overlap without stream ordering race for that workspace."""
    L = _load()
    arr = _batch_ops_array(ops, "batch_execute_noalloc")
    _check(L.ttc_batch_execute_noalloc(len(ops), arr, int(wg_per_cu), _stream(stream)), "ttc_batch_execute_noalloc")


class Batch:
    """Grouped launch of many transpositions (ttc_batch_create).

    The device work table holds raw pointers to every op's A and B, and a
    captured graph replays them, so the Batch keeps a reference to each
    tensor it was given for as long as it lives."""

    def __init__(self, ops, wg_per_cu=0):
        """ops: list of dicts with type_a, type_b, perm, size, A, B, alpha, beta,
        optional lda, ldb, beta_is_zero, config.  Raises ValueError when a
        tensor is smaller than the extent its op spans."""
        L = _load()
        arr = _batch_ops_array(ops, "Batch")
        self._keep = [(o.get("A"), o.get("B")) for o in ops]
        self._h = ctypes.c_void_p()
        _check(L.ttc_batch_create(ctypes.byref(self._h), len(ops), arr, int(wg_per_cu)), "ttc_batch_create")
        self._lib = L
        self.nops = len(ops)

    def execute(self, stream=None):
        _check(self._lib.ttc_batch_execute(self._h, _stream(stream)), "ttc_batch_execute")

    def execute_loop(self, stream=None):
        _check(self._lib.ttc_batch_execute_loop(self._h, _stream(stream)), "ttc_batch_execute_loop")

    def graph_launch(self, stream, per_op_loop=False):
        _check(self._lib.ttc_batch_graph_launch(self._h, _stream(stream), int(bool(per_op_loop))),
               "ttc_batch_graph_launch")

    def info(self):
        t, b = ctypes.c_int64(), ctypes.c_int64()
        w = ctypes.c_int()
        _check(self._lib.ttc_batch_get_info(self._h, ctypes.byref(t), ctypes.byref(w), ctypes.byref(b)),
               "ttc_batch_get_info")
        return {"tiles": t.value, "workgroups": w.value, "bytes": b.value}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ttc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def torch_dtype(t):
    """torch storage dtype used for a ttc dtype in device buffers."""
    import torch
    return {S: torch.float32, D: torch.float64, C: torch.complex64, Z: torch.complex128,
            BF16: torch.bfloat16, F16: torch.float16, E4M3: torch.uint8, E5M2: torch.uint8}[t]
