"""Writes the golden fixtures under tests/golden/ from the oracle.

Each fixture file holds small seeded cases (inputs and the oracle's output)
for one A/B type pair: the permutations and sizes of the reference's test
sweep (test-framework/ttc_test.py:80-92) at reduced extents, beta in
{0, 1.2, -1}, a padded sub-tensor case (lda/ldb, ttc/ttc.py:733-746), and
the plain benchmark shapes at reduced extents: perm (1,0,2) (the
1000x768x16 case) and perm (3,1,0,2) with outliers, alpha 1.3, beta 0.7
(the 256x192x128x96 mixed-precision case).  fp8.npz holds cast-transpose
cases (scale, amax).  Files stay below 64 KiB.

  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from common import CODE, E4M3, E5M2, F16, S, BF16, make_case, oracle_run  # noqa: E402
from oracle import oracle as orc  # noqa: E402

CODES = ["s", "d", "c", "z", "sd", "ds", "cz", "zc", "b", "h", "sb", "sh", "bs", "hs"]
SHAPES = [([1, 0], [7, 5]), ([2, 1, 0], [3, 4, 5]), ([0, 2, 1], [4, 3, 5])]   # every beta
MORE_SHAPES = [([3, 1, 0, 2], [3, 2, 4, 3]), ([2, 0, 3, 1], [2, 3, 3, 4]), ([1, 0, 2], [5, 4, 3])]  # beta 1.2
BETAS = [0.0, 1.2, -1.0]


def cases():
    """(perm, size, lda, ldb, alpha, beta, dist) of every fixture case."""
    out = []
    for perm, size in SHAPES:
        for beta in BETAS:
            out.append((perm, size, None, None, 1.1, beta, "normal"))
    for perm, size in MORE_SHAPES:
        out.append((perm, size, None, None, 1.1, 1.2, "normal"))
    out.append(([2, 0, 1], [4, 3, 3], [6, 4, 3], [4, 6, 3], 0.9, -0.5, "normal"))  # padded sub-tensor
    out.append(([1, 0, 2], [11, 9, 2], None, None, 1.0, 1.0, "uniform"))          # plain 3-D shape, reduced
    out.append(([3, 1, 0, 2], [4, 3, 4, 3], None, None, 1.3, 0.7, "outliers"))     # mixed-precision shape, reduced
    return out


def write_pair(code):
    tA, tB = CODE[code]
    arrays, index = {}, []
    for n, (perm, size, lda, ldb, alpha, beta, dist) in enumerate(cases()):
        A, B = make_case(tA, tB, perm, size, lda, ldb, seed=7000 + n, dist=dist)
        out = oracle_run(tA, tB, A, B, alpha, beta, perm, size, lda, ldb)
        arrays["A%d" % n], arrays["B%d" % n], arrays["out%d" % n] = A, B, out
        index.append({"perm": perm, "size": size, "lda": lda, "ldb": ldb, "alpha": alpha, "beta": beta})
    arrays["index"] = np.frombuffer(json.dumps({"type_a": tA, "type_b": tB, "cases": index}).encode(), np.uint8)
    path = os.path.join(HERE, "%s.npz" % code)
    np.savez_compressed(path, **arrays)
    return path


def write_fp8():
    arrays, index = {}, []
    n = 0
    for tA in (S, BF16, F16):
        for tB in (E4M3, E5M2):
            for perm, size in (([1, 0], [12, 10]), ([0, 2, 1, 3], [4, 3, 5, 2])):
                for alpha, scale in ((1.0, 1.0), (0.5, 3.5), (1.0, 100.0)):
                    A, _ = make_case(tA, tA, perm, size, seed=8000 + n, dist="heavy")
                    out = np.zeros(int(np.prod(size)), np.uint8)
                    amax = orc.cast_fp8(tA, tB, A, out, alpha, scale, perm, size)
                    arrays["A%d" % n], arrays["out%d" % n] = A, out
                    index.append({"type_a": tA, "type_b": tB, "perm": perm, "size": size, "alpha": alpha,
                                  "scale": scale, "amax": amax})
                    n += 1
    arrays["index"] = np.frombuffer(json.dumps({"cases": index}).encode(), np.uint8)
    path = os.path.join(HERE, "fp8.npz")
    np.savez_compressed(path, **arrays)
    return path


def load(path):
    """(meta, arrays) of a fixture file."""
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    meta = json.loads(bytes(arrays.pop("index")).decode())
    return meta, arrays


def main():
    orc.build()
    for code in CODES:
        p = write_pair(code)
        print(p, os.path.getsize(p))
    p = write_fp8()
    print(p, os.path.getsize(p))


if __name__ == "__main__":
    main()
