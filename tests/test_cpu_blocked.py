"""CPU tests of the blocked, vectorised CPU baseline (oracle/ttc_cpu_blocked.cpp,
TTC's blocked AVX / AVX-512 method, ttc/transposeGenerator.py:846-1474):
bit-identical to the oracle on random shapes, permutations, padded
sub-tensors and both beta classes, with each instruction set the host has;
its micro-kernels are exercised on blocks that fill them and on edges."""
import numpy as np
import pytest

from common import D, S, make_case, oracle_run, same_bits
from oracle import oracle as orc


def _isas():
    have = orc.blocked_isa()
    return [i for i in (1, 2) if i <= have]


def _check(t, perm, size, lda, ldb, alpha, beta, seed, isa):
    A, B = make_case(t, t, perm, size, lda, ldb, seed=seed, dist="normal")
    want = oracle_run(t, t, A, B, alpha, beta, perm, size, lda, ldb)
    got = B.copy()
    orc.blocked_transpose(t, A, got, alpha, beta, perm, size, lda, ldb, isa=isa, nthreads=4)
    return same_bits(got, want)


@pytest.mark.parametrize("t", [S, D])
def test_blocked_matches_oracle_on_random_problems(t):
    if not _isas():
        pytest.skip("host has neither AVX2 nor AVX-512")
    rng = np.random.default_rng(21)
    for case in range(120):
        r = int(rng.integers(1, 6))
        size = [int(x) for x in rng.integers(1, 40 if r <= 3 else 12, r)]
        perm = [int(x) for x in rng.permutation(r)]
        pad = rng.random() < 0.3
        lda = [s + int(rng.integers(0, 4)) for s in size] if pad else None
        ldb = [size[perm[q]] + int(rng.integers(0, 4)) for q in range(r)] if pad else None
        beta = float(rng.choice([0.0, 1.0, -0.75]))
        for isa in _isas():
            assert _check(t, perm, size, lda, ldb, 1.25, beta, 100 + case, isa), (case, perm, size, lda, ldb, beta,
                                                                                    isa)


@pytest.mark.parametrize("t", [S, D])
def test_blocked_full_micro_blocks_and_edges(t):
    """Extents that are multiples of 16 (every micro-block full), and one
    past (an edge row and column in every macro block)."""
    for size in ([128, 64], [129, 65], [32, 48, 3], [33, 17, 5]):
        perm = [1, 0] + list(range(2, len(size)))
        for beta in (0.0, 1.5):
            for isa in _isas():
                assert _check(t, perm, size, None, None, 0.5, beta, 7, isa), (size, beta, isa)


def test_blocked_rejects_other_types():
    A = np.zeros(4, np.complex64)
    with pytest.raises(ValueError):
        orc.blocked_transpose(2, A, A.copy(), 1.0, 0.0, [1, 0], [2, 2])
