"""Shared helpers for the tests: input generation, an independent numpy
restatement of the operation (used to pin the oracle), device round trips."""
import numpy as np

from oracle import oracle as orc

S, D, C, Z, BF16, F16, E4M3, E5M2 = range(8)
NAMES = {S: "s", D: "d", C: "c", Z: "z", BF16: "bf16", F16: "f16", E4M3: "e4m3", E5M2: "e5m2"}
CODE = {"s": (S, S), "d": (D, D), "c": (C, C), "z": (Z, Z), "sd": (S, D), "ds": (D, S), "cz": (C, Z), "zc": (Z, C),
        "b": (BF16, BF16), "h": (F16, F16), "sb": (S, BF16), "sh": (S, F16), "bs": (BF16, S), "hs": (F16, S)}
STORAGE = orc.NP_STORAGE


def extents(perm, size, lda=None, ldb=None):
    return orc.extent_elems(perm, size, lda, ldb)


def bf16_bits(x32):
    """float32 -> bfloat16 bits via torch on the CPU (round to nearest even)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x32, dtype=np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).numpy().view(np.uint16).copy()


def bf16_to_f32(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def f16_to_f32(bits):
    return bits.view(np.float16).astype(np.float32)


def rand_buffer(t, n, rng, dist="uniform"):
    """Flat storage buffer of n elements of type t."""
    def draw(m):
        if dist == "uniform":
            return rng.uniform(-1.0, 1.0, m)
        if dist == "normal":
            return rng.standard_normal(m)
        if dist == "outliers":  # N(0,1) with 1% of entries x1e4
            v = rng.standard_normal(m)
            v[rng.random(m) < 0.01] *= 1e4
            return v
        if dist == "heavy":  # N(0,1), 0.1% replaced by t(3) draws
            v = rng.standard_normal(m)
            mask = rng.random(m) < 0.001
            v[mask] = rng.standard_t(3, int(mask.sum()))
            return v
        if dist == "index":  # A[i] = i, the reference harness's init
            return np.arange(m, dtype=np.float64)
        raise ValueError(dist)
    if t == S:
        return draw(n).astype(np.float32)
    if t == D:
        return draw(n).astype(np.float64)
    if t == C:
        return (draw(n) + 1j * draw(n)).astype(np.complex64)
    if t == Z:
        return (draw(n) + 1j * draw(n)).astype(np.complex128)
    if t == BF16:
        return bf16_bits(draw(n).astype(np.float32))
    if t == F16:
        return draw(n).astype(np.float16).view(np.uint16)
    raise ValueError(t)


def _as_f32(t, v):
    if t == S:
        return v
    if t == BF16:
        return bf16_to_f32(v)
    if t == F16:
        return f16_to_f32(v)
    raise ValueError(t)


def _from_f32(t, r):
    if t == S:
        return r.astype(np.float32)
    if t == BF16:
        return bf16_bits(r)
    if t == F16:
        return r.astype(np.float32).astype(np.float16).view(np.uint16)
    raise ValueError(t)


def _real_update(tA, tB, a, b, alpha, beta, bz):
    """Element update on real arrays, with numpy scalar types standing for
    the C types of the reference's generated code."""
    f32 = (S, BF16, F16)
    if tA in f32 and tB in f32:
        x = np.float32(alpha) * _as_f32(tA, a)
        if not bz:
            x = x + np.float32(beta) * _as_f32(tB, b)
        return _from_f32(tB, x)
    if tA == D and tB == D:
        x = np.float64(alpha) * a
        return x if bz else x + np.float64(beta) * b
    if tA == S and tB == D:
        x = (np.float32(alpha) * a).astype(np.float64)
        return x if bz else x + np.float64(beta) * b
    if tA == D and tB == S:
        x = np.float64(alpha) * a
        if not bz:
            x = x + (np.float32(beta) * b).astype(np.float64)
        return x.astype(np.float32)
    raise ValueError((tA, tB))


def np_reference(tA, tB, A, B, alpha, beta, perm, size, lda=None, ldb=None, beta_is_zero=None):
    """B_out (flat) = the operation, computed with numpy's transpose on a
    Fortran-order view: an implementation independent of the oracle's loop
    nest and stride arithmetic."""
    if beta_is_zero is None:
        beta_is_zero = beta == 0
    d = len(size)
    LA = list(lda) if lda else list(size)
    LB = list(ldb) if ldb else [size[perm[q]] for q in range(d)]
    shapeA = LA[:-1] + [size[-1]]
    shapeB = LB[:-1] + [size[perm[-1]]]
    a = A.reshape(shapeA, order="F")[tuple(slice(0, s) for s in size)]
    at = np.transpose(a, perm)
    Bout = B.copy()
    bfull = Bout.reshape(shapeB, order="F")
    sl = tuple(slice(0, size[perm[q]]) for q in range(d))
    bold = bfull[sl]
    cplx = {C: S, Z: D}
    if tA in cplx:
        ra, rb = cplx[tA], cplx[tB]
        re = _real_update(ra, rb, np.ascontiguousarray(at.real), np.ascontiguousarray(bold.real), alpha, beta,
                          beta_is_zero)
        im = _real_update(ra, rb, np.ascontiguousarray(at.imag), np.ascontiguousarray(bold.imag), alpha, beta,
                          beta_is_zero)
        new = np.empty(re.shape, dtype=bfull.dtype)
        new.real = re
        new.imag = im
    else:
        new = _real_update(tA, tB, np.ascontiguousarray(at), np.ascontiguousarray(bold), alpha, beta, beta_is_zero)
    bfull[sl] = new
    return Bout


def oracle_run(tA, tB, A, B, alpha, beta, perm, size, lda=None, ldb=None, beta_is_zero=None, nthreads=0):
    out = B.copy()
    orc.transpose(tA, tB, A, out, alpha, beta, perm, size, lda, ldb, beta_is_zero, nthreads)
    return out


def make_case(tA, tB, perm, size, lda=None, ldb=None, seed=0, dist="uniform"):
    rng = np.random.default_rng(seed)
    na, nb = extents(perm, size, lda, ldb)
    A = rand_buffer(tA, na, rng, dist)
    B = rand_buffer(tB, nb, rng, dist)
    return A, B


def same_bits(x, y):
    return x.dtype == y.dtype and x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8))


def first_diff(x, y):
    xb, yb = x.view(np.uint8), y.view(np.uint8)
    w = x.itemsize
    bad = np.nonzero(xb.reshape(-1, w).any(axis=1) != yb.reshape(-1, w).any(axis=1))[0]
    idx = np.nonzero((xb.reshape(-1, w) != yb.reshape(-1, w)).any(axis=1))[0]
    return idx[:10], len(idx)


# ----------------------------------------------------------------- device
def to_device(arr, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).to(device)


def from_device(t, like):
    return t.cpu().numpy().view(like.dtype).reshape(like.shape)


# ---------------------------------------------------------------- fixtures
GOLDEN_CODES = ["s", "d", "c", "z", "sd", "ds", "cz", "zc", "b", "h", "sb", "sh", "bs", "hs"]


def load_golden(name):
    """(meta, arrays) of tests/golden/<name>.npz (written by make_golden.py)."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".npz")
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    meta = json.loads(bytes(arrays.pop("index")).decode())
    return meta, arrays
