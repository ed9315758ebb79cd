"""Benchmark workloads (shapes only; data is generated where it is used).

The reference sizes its benchmark tensors at 200 MiB (benchmark/benchmark.py:
39, 97) and counts bytes as N*(sA+sB) for beta == 0 and N*(sA+2*sB)
otherwise (ttc/transposeGenerator.py:665-668).  The shapes here:

  plain_c1():  fp32, perm (1,0,2), size (1000,768,16), alpha = beta = 1.
  plain_c2():  16 seeded random non-identity perms of rank 2..6, extents
               drawn log-uniform in [8, 4096], rescaled by a common factor so
               A is 200 MiB (+-0.5%) in fp32 and in fp64 (extents >= 8, powers
               of two bumped by one); case 15
               extracts a sub-tensor (lda/ldb padded by 1..31 per dimension).
  mixed_c3():  rank-4 perm (3,1,0,2), size (256,192,128,96), alpha 1.3,
               beta 0.7, for s->bf16, d->s, c and z.
  fp8_c4():    2-D (8192,14336) perm (1,0) and 4-D (128,4096,32,8) perm
               (0,2,1,3) from bf16 / fp32 to e4m3 / e5m2.
  batch_c5():  512 independent ops of rank 2..5, A 4 KiB..64 MiB log-uniform,
               fp32 / bf16 / fp64, random perms, alpha in [0.5,2], beta in {0,1}.
"""
import math

import numpy as np

MIB = 1 << 20


def algorithmic_bytes(n, size_a, size_b, beta_is_zero):
    """N*sA + N*sB (+ N*sB when B is read) — ttc/transposeGenerator.py:665-668."""
    return n * size_a + n * size_b * (1 if beta_is_zero else 2)


def _bump(e):
    e = int(e)
    if e >= 2 and (e & (e - 1)) == 0:
        e += 1
    return e


def _random_perm(rng, r):
    while True:
        p = [int(x) for x in rng.permutation(r)]
        if p != list(range(r)):
            return p


def _fit_extents(base, target, lo=8, hi=1 << 30, tol=0.005):
    """Scale the drawn extents by a common factor toward prod == target
    (keeping every extent >= lo), round, bump powers of two, then correct one
    or two extents to land within `tol` of target."""
    r = len(base)
    ext = [float(b) for b in base]
    for _ in range(64):
        ratio = target / float(np.prod(ext))
        if abs(ratio - 1.0) < 1e-12:
            break
        free = [i for i in range(r) if (ratio > 1 and ext[i] < hi) or (ratio < 1 and ext[i] > lo)]
        if not free:
            break
        f = ratio ** (1.0 / len(free))
        for i in free:
            ext[i] = min(hi, max(lo, ext[i] * f))
    ext = [_bump(max(lo, round(e))) for e in ext]

    def err(e):
        return abs(float(np.prod(e)) / target - 1.0)

    best = list(ext)
    for i in range(r):  # re-derive one extent from the others
        e = list(ext)
        rest = float(np.prod([x for j, x in enumerate(e) if j != i]))
        e[i] = _bump(max(lo, int(round(target / rest))))
        if err(e) < err(best):
            best = e
    if err(best) > tol and r >= 2:  # two extents: small search around the current values
        order = sorted(range(r), key=lambda i: -best[i])
        i, j = order[0], order[1]
        cur = list(best)
        for di in range(-6, 7):
            for dj in range(-6, 7):
                e = list(cur)
                e[i], e[j] = _bump(max(lo, cur[i] + di)), _bump(max(lo, cur[j] + dj))
                if err(e) < err(best):
                    best = e
    return [int(x) for x in best]


def plain_c1():
    return {"perm": [1, 0, 2], "size": [1000, 768, 16], "alpha": 1.0, "beta": 1.0}


def plain_c2(target_bytes=200 * MIB, elem_sizes=(4, 8)):
    """16 cases; each is {k, perm, rank, sizes: {elem_size: size}, lda, ldb}."""
    cases = []
    for k in range(16):
        rng = np.random.default_rng(1000 + k)
        r = 2 + (k % 5)
        perm = _random_perm(rng, r)
        base = np.exp(rng.uniform(math.log(8), math.log(4096), r))
        pad_a = [int(x) for x in rng.integers(1, 32, r)]
        pad_b = [int(x) for x in rng.integers(1, 32, r)]
        sizes, ldas, ldbs = {}, {}, {}
        for es in elem_sizes:
            ext = _fit_extents(base, target_bytes / es)
            sizes[es] = ext
            if k == 15:
                ldas[es] = [ext[i] + pad_a[i] for i in range(r)]
                ldbs[es] = [ext[perm[i]] + pad_b[i] for i in range(r)]
        cases.append({"k": k, "rank": r, "perm": perm, "sizes": sizes,
                      "lda": ldas if k == 15 else None, "ldb": ldbs if k == 15 else None})
    return cases


def mixed_c3():
    return {"perm": [3, 1, 0, 2], "size": [256, 192, 128, 96], "alpha": 1.3, "beta": 0.7,
            "dtypes": ["sb", "ds", "c", "z"]}


def fp8_c4():
    return [{"perm": [1, 0], "size": [8192, 14336]}, {"perm": [0, 2, 1, 3], "size": [128, 4096, 32, 8]}]


def batch_c5(nops=512, seed=5, min_bytes=4 << 10, max_bytes=64 << 20):
    """List of ops {type (0 s, 1 d, 4 bf16), perm, size, alpha, beta}."""
    rng = np.random.default_rng(seed)
    types = [0, 4, 1]
    ops = []
    for i in range(nops):
        r = int(rng.integers(2, 6))
        t = types[int(rng.integers(0, 3))]
        es = {0: 4, 1: 8, 4: 2}[t]
        nbytes = math.exp(rng.uniform(math.log(min_bytes), math.log(max_bytes)))
        n = max(1, int(nbytes / es))
        base = np.exp(rng.uniform(0, 1, r))
        f = (n / float(np.prod(base))) ** (1.0 / r)
        ext = [max(1, int(round(b * f))) for b in base]
        big = int(np.argmax(ext))
        rest = int(np.prod([e for j, e in enumerate(ext) if j != big]))
        ext[big] = max(1, int(round(n / rest)))
        perm = _random_perm(rng, r)
        alpha = float(rng.uniform(0.5, 2.0))
        beta = float(rng.integers(0, 2))
        ops.append({"type": t, "perm": perm, "size": ext, "alpha": alpha, "beta": beta})
    return ops
