"""Pins the oracle (CPU restatement of TTC's reference loop nest) against
checks that do not share its code: numpy's transpose with numpy scalar
arithmetic, brute-force index enumeration on tiny tensors, algebraic
special cases, and torch's CPU rounding for the new output types.

The reference ships no golden vectors or known-answer tests for this path
(its only check is the self-comparison in the generated harness,
ttc/transposeGenerator.py:612-635, 771-779); the matrix of perms, sizes,
dtypes, alpha and beta below is the reference's own sweep
(test-framework/ttc_test.py:80-92).
"""
import itertools

import numpy as np
import pytest

from common import (BF16, C, CODE, D, F16, S, Z, make_case, np_reference, oracle_run, rand_buffer,
                    same_bits)
from oracle import oracle as orc

TTC_TEST_DTYPES = ["d", "s", "c", "z", "sd", "ds", "cz", "zc"]
TTC_TEST_BETAS = [1, -1, 1.2, 0]
TTC_TEST_PERMS = [[2, 0, 1], [1, 0], [0, 2, 1], [1, 2, 0], [1, 0, 2], [2, 1, 0], [3, 1, 0, 2]]
TTC_TEST_SIZES = [
    [[11, 8, 3], [16, 4, 25]],
    [[7, 3], [7, 8], [8, 7], [8, 8], [16, 11], [11, 16], [17, 17], [17, 23]],
    [[11, 8, 3], [1, 4, 3], [11, 7, 17], [11, 17, 7], [7, 32, 32]],
    [[11, 8, 3], [4, 4, 3], [12, 8, 17], [33, 17, 7], [32, 32, 4]],
    [[11, 8, 3], [16, 16, 3], [12, 8, 17], [33, 17, 7], [32, 32, 4]],
    [[11, 8, 3], [16, 4, 25], [12, 8, 17], [33, 17, 9], [32, 32, 4]],
    [[17, 3, 7, 9]],
]


def ttc_test_cases():
    """(perm, size, dtype code, alpha, beta) of the reference's sweep."""
    out = []
    for beta in TTC_TEST_BETAS:          # alpha/beta sweep on the first two perms
        for code in TTC_TEST_DTYPES:
            for c in range(2):
                for size in TTC_TEST_SIZES[c]:
                    out.append((TTC_TEST_PERMS[c], size, code, 1.1, beta))
    for c in range(len(TTC_TEST_SIZES)):  # all perms and sizes, beta = 0
        for code in TTC_TEST_DTYPES:
            for size in TTC_TEST_SIZES[c]:
                out.append((TTC_TEST_PERMS[c], size, code, 1.1, 0.0))
    for c in range(len(TTC_TEST_SIZES)):  # alpha = 1, beta = 1.1
        for code in TTC_TEST_DTYPES:
            for size in TTC_TEST_SIZES[c]:
                out.append((TTC_TEST_PERMS[c], size, code, 1.0, 1.1))
    return out


def test_oracle_matches_numpy_on_reference_sweep(oracle_lib):
    n = 0
    for i, (perm, size, code, alpha, beta) in enumerate(ttc_test_cases()):
        tA, tB = CODE[code]
        A, B = make_case(tA, tB, perm, size, seed=i, dist="normal")
        got = oracle_run(tA, tB, A, B, alpha, beta, perm, size)
        want = np_reference(tA, tB, A, B, alpha, beta, perm, size)
        assert same_bits(got, want), (perm, size, code, alpha, beta)
        n += 1
    assert n > 600


@pytest.mark.parametrize("code", ["s", "d", "sd", "ds", "c", "z", "cz", "zc", "b", "h", "sb", "sh", "bs", "hs"])
def test_oracle_matches_numpy_random(oracle_lib, code):
    tA, tB = CODE[code]
    rng = np.random.default_rng(7)
    for trial in range(12):
        r = int(rng.integers(1, 6))
        size = [int(x) for x in rng.integers(1, 9, r)]
        perm = [int(x) for x in rng.permutation(r)]
        pad = trial % 3 == 0
        lda = [s + int(rng.integers(0, 4)) for s in size] if pad else None
        ldb = [size[perm[q]] + int(rng.integers(0, 4)) for q in range(r)] if pad else None
        alpha = float(rng.choice([1.0, -1.0, 1.3, 0.7]))
        beta = float(rng.choice([0.0, 1.0, -0.5, 0.7]))
        A, B = make_case(tA, tB, perm, size, lda, ldb, seed=trial, dist="outliers")
        got = oracle_run(tA, tB, A, B, alpha, beta, perm, size, lda, ldb)
        want = np_reference(tA, tB, A, B, alpha, beta, perm, size, lda, ldb)
        assert same_bits(got, want), (code, perm, size, lda, ldb, alpha, beta)


def _brute_offsets(perm, size, lda, ldb):
    """(offA, offB) pairs by direct enumeration of multi-indices."""
    d = len(size)
    LA = lda or size
    LB = ldb or [size[perm[q]] for q in range(d)]
    out = []
    for idx in itertools.product(*[range(s) for s in size]):
        oa, m = 0, 1
        for k in range(d):
            oa += idx[k] * m
            m *= LA[k]
        ob, m = 0, 1
        for q in range(d):
            ob += idx[perm[q]] * m  # B's dimension q is A's dimension perm[q]
            m *= LB[q]
        out.append((oa, ob))
    return out


@pytest.mark.parametrize("perm,size,pad", [
    ([1, 0], [3, 5], False), ([2, 0, 1], [2, 3, 4], False), ([0, 2, 1], [3, 2, 4], True),
    ([3, 1, 0, 2], [2, 3, 2, 3], True), ([4, 2, 0, 3, 1], [2, 2, 3, 2, 2], False), ([0], [7], True),
])
def test_oracle_index_map_brute_force(oracle_lib, perm, size, pad):
    """A[i] = offset of i in A, alpha = 1, beta = 0: B at every brute-force
    offset must hold exactly the matching A offset; padding is untouched."""
    lda = [s + 2 for s in size] if pad else None
    ldb = [size[perm[q]] + 1 for q in range(len(size))] if pad else None
    na, nb = orc.extent_elems(perm, size, lda, ldb)
    A = np.arange(na, dtype=np.float64)
    B = np.full(nb, -7.0)
    out = oracle_run(D, D, A, B, 1.0, 0.0, perm, size, lda, ldb)
    pairs = _brute_offsets(perm, size, lda, ldb)
    touched = np.zeros(nb, bool)
    for oa, ob in pairs:
        assert out[ob] == oa
        touched[ob] = True
    assert np.all(out[~touched] == -7.0)


def test_oracle_beta_zero_does_not_read_b(oracle_lib):
    A, _ = make_case(S, S, [1, 0], [9, 13], seed=1)
    B = np.full(9 * 13, np.nan, np.float32)
    out = oracle_run(S, S, A, B, 1.0, 0.0, [1, 0], [9, 13])
    assert not np.isnan(out).any()
    # with beta != 0 specialisation forced, NaN propagates (B is read)
    out2 = oracle_run(S, S, A, B, 1.0, 0.0, [1, 0], [9, 13], beta_is_zero=False)
    assert np.isnan(out2).all()


def test_oracle_algebra(oracle_lib):
    perm, size = [2, 0, 1], [5, 6, 7]
    A, B = make_case(S, S, perm, size, seed=3, dist="normal")
    neg = oracle_run(S, S, A, B, -1.0, 0.0, perm, size)
    pos = oracle_run(S, S, A, B, 1.0, 0.0, perm, size)
    assert same_bits(neg, -pos)  # sign of alpha
    keep = oracle_run(S, S, A, B, 0.0, 1.0, perm, size)
    assert np.array_equal(keep, B + np.float32(0.0))  # alpha = 0, beta = 1 leaves B's values
    # transposing back with the inverse perm restores A
    inv = list(np.argsort(perm))
    sizeB = [size[p] for p in perm]
    back = oracle_run(S, S, pos, np.zeros_like(A), 1.0, 0.0, inv, sizeB)
    assert same_bits(back, A)


def test_bf16_f16_rounding_matches_torch(oracle_lib):
    import torch
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 20000),
        np.array([0.0, -0.0, np.inf, -np.inf, 65504.0, 65519.0, 65520.0, 6e-8, 3e-8, 1e-40, -1e-40,
                  3.3895314e38, 3.39e38], np.float32),
    ]).astype(np.float32)
    got_bf = orc.convert(0, x)
    want_bf = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(got_bf, want_bf)
    got_h = orc.convert(1, x)
    want_h = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got_h, want_h)
    nan = orc.convert(0, np.array([np.nan], np.float32))
    assert np.isnan(((nan.astype(np.uint32) << 16).view(np.float32)))


@pytest.mark.parametrize("op,fmax,tdt", [(2, 448.0, "float8_e4m3fn"), (3, 57344.0, "float8_e5m2")])
def test_fp8_rounding_matches_torch(oracle_lib, op, fmax, tdt):
    """Saturating round-to-nearest-even against torch's cast of the clamped
    value (torch's own cast does not saturate)."""
    import torch
    rng = np.random.default_rng(12)
    x = (rng.standard_normal(50000) * 10.0 ** rng.uniform(-9, 6, 50000)).astype(np.float32)
    x = np.concatenate([x, np.array([0.0, -0.0, fmax, -fmax, fmax * 1.03, fmax * 1.1, 1e30, -1e30, np.inf,
                                     -np.inf, 2.0 ** -9, 2.0 ** -10, 3 * 2.0 ** -11, 2.0 ** -16, 2.0 ** -17,
                                     2.0 ** -6, 2.0 ** -14], np.float32)])
    got = orc.convert(op, x)
    clamped = np.clip(x, -fmax, fmax)
    want = torch.from_numpy(clamped).to(getattr(torch, tdt)).view(torch.uint8).numpy()
    assert np.array_equal(got, want)
    nan = orc.convert(op, np.array([np.nan, -np.nan], np.float32))
    assert (nan & 0x7f).tolist() == [0x7f, 0x7f]


def test_fp8_cast_oracle_closed_form(oracle_lib):
    import torch
    rng = np.random.default_rng(5)
    perm, size = [1, 0], [48, 40]
    A = rng.standard_normal(48 * 40).astype(np.float32)
    for tB, op, fmax in [(6, 2, 448.0), (7, 3, 57344.0)]:
        out = np.zeros(48 * 40, np.uint8)
        scale = np.float32(fmax / np.abs(A).max())
        amax = orc.cast_fp8(S, tB, A, out, 1.0, float(scale), perm, size)
        assert amax == float(np.abs(A).max())
        at = A.reshape(size, order="F").T.reshape(-1, order="F")
        want = orc.convert(op, (at * scale).astype(np.float32))
        assert np.array_equal(out, want)


def test_fastdiv_math():
    """The multiply-high division used by the kernels (ttc_params.h):
    s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1,
    n / d == (mulhi(n, m) + n) >> s for n < 2^31."""
    rng = np.random.default_rng(0)
    ds = list(range(1, 300)) + [int(x) for x in rng.integers(1, 2 ** 31 - 1, 300)] + [2 ** 30, 2 ** 31 - 1]
    for d in ds:
        s = 0
        while (1 << s) < d:
            s += 1
        m = ((1 << 32) * ((1 << s) - d)) // d + 1
        m &= 0xffffffff
        ns = [0, 1, d - 1, d, d + 1, 2 ** 31 - 1] + [int(x) for x in rng.integers(0, 2 ** 31, 200)]
        for n in ns:
            q = (((n * m) >> 32) + n) >> s
            assert q == n // d, (n, d)


def test_pinning_catches_planted_mistakes(oracle_lib):
    """The numpy pin has teeth: plausible mistakes planted in the oracle's
    inputs on the CPU (an inverted permutation, a dropped beta term, beta in
    the wrong precision, A's padding ignored) each produce a mismatch."""
    perm, size = [2, 0, 3, 1], [5, 4, 6, 3]
    inv = [perm.index(k) for k in range(len(perm))]
    A, B = make_case(S, S, perm, size, seed=11, dist="normal")
    want = np_reference(S, S, A, B, 1.1, 0.7, perm, size)
    assert same_bits(oracle_run(S, S, A, B, 1.1, 0.7, perm, size), want)
    assert not same_bits(oracle_run(S, S, A, B, 1.1, 0.7, inv, size), want)            # inverse permutation
    assert not same_bits(oracle_run(S, S, A, B, 1.1, 0.0, perm, size), want)           # beta term dropped
    Ad, Bd = make_case(D, S, perm, size, seed=12, dist="normal")
    want_ds = np_reference(D, S, Ad, Bd, 1.1, 0.1, perm, size)
    # beta rounded in double instead of B's float changes some results of ds
    wrong = np_reference(D, D, Ad, Bd.astype(np.float64), 1.1, 0.1, perm, size).astype(np.float32)
    assert same_bits(oracle_run(D, S, Ad, Bd, 1.1, 0.1, perm, size), want_ds)
    assert not same_bits(wrong, want_ds)
    lda = [7, 5, 6, 3]
    Ap, Bp = make_case(S, S, perm, size, lda, None, seed=13, dist="normal")
    want_p = np_reference(S, S, Ap, Bp, 1.0, 1.0, perm, size, lda)
    assert same_bits(oracle_run(S, S, Ap, Bp, 1.0, 1.0, perm, size, lda), want_p)
    dense = oracle_run(S, S, Ap[:int(np.prod(size))], Bp, 1.0, 1.0, perm, size)          # padding ignored
    assert not same_bits(dense, want_p)
