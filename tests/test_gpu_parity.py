"""GPU parity: the HIP path (through the C ABI) against the oracle, bit for
bit, on the reference's own test sweep, the benchmark shapes and the edge
cases (padding, partial tiles, size-1 and empty dimensions, misaligned
pointers, every vector width and store mode)."""
import numpy as np
import pytest

from common import (BF16, C, CODE, D, F16, S, Z, extents, first_diff, from_device, make_case, oracle_run,
                    same_bits, to_device)
from test_oracle import ttc_test_cases

pytestmark = pytest.mark.gpu


def run_hip(device, tA, tB, A, B, alpha, beta, perm, size, lda=None, ldb=None, beta_is_zero=None, config=None,
            offset_a=0, offset_b=0):
    import torch
    from ttc_amd import runtime as rt
    if beta_is_zero is None:
        beta_is_zero = beta == 0
    dA = to_device(np.concatenate([np.zeros(offset_a, A.dtype), A]) if offset_a else A, device)
    dB = to_device(np.concatenate([np.zeros(offset_b, B.dtype), B]) if offset_b else B, device)
    plan = rt.Plan(tA, tB, perm, size, lda, ldb, beta_is_zero, config)
    plan.execute(dA.data_ptr() + offset_a * A.itemsize, dB.data_ptr() + offset_b * B.itemsize, alpha, beta)
    torch.cuda.synchronize()
    out = dB.cpu().numpy().view(B.dtype)
    return out[offset_b:].copy(), plan.info()


def check(device, tA, tB, perm, size, alpha, beta, lda=None, ldb=None, seed=0, dist="uniform", config=None,
          beta_is_zero=None, offset_a=0, offset_b=0):
    A, B = make_case(tA, tB, perm, size, lda, ldb, seed=seed, dist=dist)
    want = oracle_run(tA, tB, A, B, alpha, beta, perm, size, lda, ldb, beta_is_zero)
    got, info = run_hip(device, tA, tB, A, B, alpha, beta, perm, size, lda, ldb, beta_is_zero, config,
                        offset_a, offset_b)
    if not same_bits(got, want):
        idx, n = first_diff(got, want)
        raise AssertionError("mismatch %s->%s perm=%s size=%s lda=%s ldb=%s a=%s b=%s cfg=%s: %d elements differ, "
                             "first %s; got %s want %s; plan %s" % (tA, tB, perm, size, lda, ldb, alpha, beta, config,
                                                                    n, idx, got[idx], want[idx], info))
    return info


@pytest.mark.parametrize("code", ["d", "s", "c", "z", "sd", "ds", "cz", "zc"])
def test_reference_sweep(gpu, code):
    """test-framework/ttc_test.py's matrix (perms, sizes, alpha 1.1 / 1.0,
    beta 1, -1, 1.2, 0), all bit-exact against the oracle."""
    tA, tB = CODE[code]
    n = 0
    for i, (perm, size, c, alpha, beta) in enumerate(ttc_test_cases()):
        if c != code:
            continue
        check(gpu, tA, tB, perm, size, alpha, beta, seed=i, dist="normal")
        n += 1
    assert n > 70


def test_c1_full_size(gpu):
    """fp32 perm (1,0,2), size (1000,768,16), alpha = beta = 1, seed 0."""
    check(gpu, S, S, [1, 0, 2], [1000, 768, 16], 1.0, 1.0, seed=0)
    check(gpu, S, S, [1, 0, 2], [1000, 768, 16], 1.0, 0.0, seed=0)


RANDOM_SHAPES = [
    ([5, 3, 0, 4, 1, 2], [12, 10, 9, 11, 9, 17]),
    ([1, 0], [1531, 977]), ([2, 0, 1], [97, 130, 81]), ([0, 2, 1], [33, 257, 65]),
    ([3, 1, 0, 2], [41, 19, 37, 23]), ([0, 3, 2, 1], [8, 33, 29, 31]), ([4, 2, 0, 3, 1], [13, 17, 11, 9, 21]),
    ([2, 1, 0], [1, 600, 700]), ([1, 0, 2], [3, 5000, 40]), ([0, 1], [1000, 1000]), ([1, 0], [4096, 8]),
    ([1, 0], [8, 4096]), ([3, 2, 1, 0], [9, 10, 11, 12]), ([0, 2, 1, 3], [128, 512, 32, 2]),
]


def families(tA, tB, perm, size, lda=None, ldb=None):
    """Kernel families a shape can run on: the box tile always, the row x
    column tile when A's and B's leading dimensions differ, the row copy
    when they coincide (dimension 0 in place after fusion)."""
    from ttc_amd import runtime as rt
    out = [1]
    for fam in (2, 3):
        try:
            rt.Plan(tA, tB, perm, size, lda, ldb, True, dict(family=fam))
            out.append(fam)
        except rt.TTCError as e:
            assert e.status == rt.ERROR_UNSUPPORTED
    assert len(out) == 2 or rt.fuse_indices(perm, size, lda, ldb)[0] == [0]
    return out


@pytest.mark.parametrize("code", ["s", "d", "c", "z", "sd", "ds", "cz", "zc", "b", "h", "sb", "sh", "bs", "hs"])
def test_random_shapes(gpu, code):
    tA, tB = CODE[code]
    n_rc = n_row = 0
    for i, (perm, size) in enumerate(RANDOM_SHAPES):
        for fam in families(tA, tB, perm, size):
            n_rc += fam == 2
            n_row += fam == 3
            for beta in (0.0, 0.7):
                order = 1 + (i + 2 * int(beta > 0)) % 4  # row x column tiles: A, B and grouped B numbering
                info = check(gpu, tA, tB, perm, size, 1.3, beta, seed=100 + i, dist="outliers",
                             config=dict(family=fam, tile_order=order, lds_buffers=1 + (i + int(beta > 0)) % 2))
                assert info["family"] == fam and info["tile_order"] == (order if fam == 2 else 1)
    assert n_rc >= 9 and n_row >= 4


def test_padded_subtensor(gpu):
    """lda/ldb padded by 1..31 (rank <= 3) or 1..3 (higher ranks, to keep
    the padded buffers small): B's padding must stay untouched."""
    rng = np.random.default_rng(15)
    for i, (perm, size) in enumerate(RANDOM_SHAPES[:8]):
        hi = 32 if len(size) <= 3 else 4
        lda = [s + int(rng.integers(1, hi)) for s in size]
        ldb = [size[perm[q]] + int(rng.integers(1, hi)) for q in range(len(size))]
        na, nb = extents(perm, size, lda, ldb)
        assert max(na, nb) < 40_000_000, (size, lda, ldb)
        for code in ("s", "d", "z", "ds"):
            tA, tB = CODE[code]
            for fam in families(tA, tB, perm, size, lda, ldb):
                cfg = dict(family=fam)
                # row x column tiles: grouped B order and one LDS buffer on the padded cases too
                extra = dict(tile_order=3 + i % 2, lds_buffers=1 + i % 2) if fam == 2 else {}
                check(gpu, tA, tB, perm, size, 1.0, 1.0, lda, ldb, seed=i, config=dict(cfg, **extra))
                check(gpu, tA, tB, perm, size, 1.1, 0.0, lda, None, seed=i, config=cfg)
                check(gpu, tA, tB, perm, size, 0.9, -1.0, None, ldb, seed=i, config=dict(cfg, **extra))


def test_every_vector_width_and_mode(gpu):
    from ttc_amd import runtime as rt
    from ttc_amd.gen_instances import vec_widths
    perm, size = [2, 0, 1], [64, 48, 80]
    for code in ("s", "d", "b", "c", "sb", "ds"):
        tA, tB = CODE[code]
        for va in vec_widths(rt.DTYPE_BYTES[tA]):
            for vb in vec_widths(rt.DTYPE_BYTES[tB]):
                for nt in (1, 2):
                    for order in (1, 2, 3, 4):
                        for fam in (1, 2):
                            cfg = dict(vec_a=va, vec_b=vb, nt_store=nt, grid_order=order, wg_per_cu=1,
                                       family=fam)
                            for beta in (0.0, 1.5):
                                info = check(gpu, tA, tB, perm, size, 1.25, beta, config=cfg, seed=va * 10 + vb)
                                assert info["family"] == fam and info["vec_a"] == va and info["vec_b"] == vb


def test_explicit_tiles_with_two_partial_dims(gpu):
    perm, size = [1, 0, 2], [100, 90, 7]
    for tile in ([32, 32, 1], [36, 28, 1], [100, 7, 2], [4, 90, 1], [20, 50, 4], [12, 8, 7]):
        check(gpu, S, S, perm, size, 1.0, 1.0, config=dict(tile=tile, vec_a=4 if tile[0] % 4 == 0 else 1,
                                                              vec_b=1 if tile[1] % 2 else 2))


def test_lane_fixed_walk(gpu):
    """Row x column launches of >= 2048 tiles with a 2-byte side take the
    lane-fixed walk on full tiles (the general walk on partial ones): tiles
    with exact powers of two and with up to 1/16 padding on either walk,
    every vector width pair the kernels have, beta read or not, partial last
    tiles on both sides and padded leading dimensions, all bit-exact."""
    cases = [  # (tA, tB, perm, size, tile, vec pairs)
        (BF16, BF16, [1, 0], [4096, 4104], [64, 128], [(1, 1), (8, 8), (2, 8), (8, 1)]),
        (BF16, BF16, [1, 0], [4000, 4100], [62, 128], [(1, 1), (2, 2)]),   # rows 62 -> 64
        (S, BF16, [1, 0], [2048, 8200], [128, 64], [(4, 8), (1, 1), (4, 2)]),
        (BF16, S, [1, 0], [1920, 9000], [124, 62], [(2, 2), (1, 1)]),      # rows 124 -> 128, cols 62 -> 64
        (F16, F16, [2, 1, 0], [16, 128, 4200], [16, 4, 64], [(8, 8), (1, 2)]),  # rows: two dimensions
    ]
    n = 0
    for i, (tA, tB, perm, size, tile, vecs) in enumerate(cases):
        for va, vb in vecs:
            cfg = dict(family=2, tile=tile, vec_a=va, vec_b=vb, lds_buffers=1 + n % 2)
            for beta, bz in ((0.0, True), (0.75, False)):
                info = check(gpu, tA, tB, perm, size, 1.25, beta, seed=i, dist="normal", config=cfg, beta_is_zero=bz)
                assert info["lane_fixed"] == 1 and info["tiles"] >= 2048, info
                n += 1
        # padded leading dimensions
        lda = [s + 3 for s in size]
        ldb = [size[perm[q]] + 5 for q in range(len(size))]
        info = check(gpu, tA, tB, perm, size, 0.5, 1.0, lda, ldb, seed=i, config=dict(family=2, tile=tile))
        assert info["lane_fixed"] == 1, info
    assert n >= 26
    # grouped B order: 65 tiles along rows in groups of 8 / 16 leave padding
    # ids, which the walk processes (B read and written) as empty tiles
    for order, ids in ((3, 2376), (4, 2640)):
        for beta in (0.0, 0.75):
            cfg = dict(family=2, tile=[62, 128], vec_a=2, vec_b=2, tile_order=order)
            info = check(gpu, BF16, BF16, [1, 0], [4000, 4100], 1.5, beta, seed=order, config=cfg)
            assert info["lane_fixed"] == 1 and info["tiles"] == ids > 65 * 33, info


def test_explicit_row_column_tiles(gpu):
    """Row x column tiles with partial last tiles on the row group, the
    column group or both, multi-dimension groups and non-dense tables."""
    cases = [([1, 0, 2], [100, 90, 7], [36, 28, 1]), ([1, 0, 2], [100, 90, 7], [50, 7, 1]),
             ([3, 1, 0, 2], [12, 10, 9, 11], [12, 4, 1, 5]), ([3, 1, 0, 2], [12, 10, 9, 11], [12, 10, 3, 11]),
             ([3, 2, 1, 0], [9, 10, 11, 12], [9, 4, 1, 7]), ([1, 0], [1531, 977], [96, 80]),
             ([1, 0], [1531, 977], [1531, 4])]
    from ttc_amd import runtime as rt
    n = 0
    for perm, size, tile in cases:
        for code in ("s", "d", "z", "bs"):
            tA, tB = CODE[code]
            cfg = dict(tile=tile, family=2, vec_a=1, vec_b=1, tile_order=1 + n % 4, lds_buffers=1 + n % 2)
            # volume limit of the family: 256 threads x (128 bytes, at most 32 elements) each side
            limit = 256 * min(min(32, 128 // rt.DTYPE_BYTES[t]) for t in (tA, tB))
            if int(np.prod(tile)) > limit:
                with pytest.raises(rt.TTCError):
                    rt.Plan(tA, tB, perm, size, None, None, True, cfg)
                continue
            for beta in (0.0, 1.0):
                info = check(gpu, tA, tB, perm, size, 1.0, beta, config=cfg)
                assert info["family"] == 2 and info["tile"] == tile
                n += 1
    assert n >= 40


def test_row_copy_every_vector_width_and_mode(gpu):
    """Row-copy family (dimension 0 in place): every instantiated width,
    store mode and work-group count, dense and padded."""
    from ttc_amd import runtime as rt
    from ttc_amd.gen_instances import row_widths
    for perm, size in (([0, 2, 1], [64, 48, 80]), ([0, 3, 1, 2], [16, 5, 7, 9]), ([0, 1], [8, 999])):
        for code in ("s", "d", "b", "c", "z", "sb", "ds", "zc"):
            tA, tB = CODE[code]
            for v in row_widths(rt.DTYPE_BYTES[tA], rt.DTYPE_BYTES[tB]):
                for nt in (1, 2):
                    for wg in (0, 1):
                        cfg = dict(vec_a=v, vec_b=v, nt_store=nt, wg_per_cu=wg, family=3)
                        for beta in (0.0, 1.5):
                            info = check(gpu, tA, tB, perm, size, 1.25, beta, config=cfg, seed=v * 10 + nt)
                            assert info["family"] == 3 and info["vec_a"] == v
    lda, ldb = [68, 50, 81], [68, 83, 49]
    for code in ("s", "d", "sb"):
        tA, tB = CODE[code]
        check(gpu, tA, tB, [0, 2, 1], [64, 48, 80], 1.0, 1.0, lda, ldb, config=dict(family=3))


def test_misaligned_pointers_use_scalar_variant(gpu):
    for off in (1, 3):
        for fam in (1, 2):
            cfg = dict(family=fam)
            check(gpu, S, S, [1, 0], [256, 192], 1.0, 1.0, offset_a=off, offset_b=0, config=cfg)
            check(gpu, S, S, [1, 0], [256, 192], 1.0, 0.0, offset_a=0, offset_b=off, config=cfg)
            check(gpu, BF16, BF16, [1, 0], [256, 192], 1.0, 0.0, offset_a=off, offset_b=off, config=cfg)
        info = check(gpu, S, S, [0, 2, 1], [256, 6, 30], 1.0, 1.0, offset_a=off, offset_b=0)
        assert info["family"] == 3 and info["vec_a"] == 4
        check(gpu, BF16, S, [0, 2, 1], [256, 6, 30], 1.0, 0.0, offset_a=0, offset_b=off)


def test_edge_shapes(gpu):
    check(gpu, S, S, [0], [1], 2.0, 0.0)
    check(gpu, S, S, [0], [12345], 2.0, 1.0)
    check(gpu, D, D, [1, 0], [1, 1], 2.0, 1.0)
    check(gpu, S, S, [2, 1, 0], [1, 1, 5], 1.0, 0.0)
    check(gpu, Z, Z, [3, 2, 1, 0], [2, 1, 3, 1], 1.0, 1.0)
    check(gpu, S, S, [0, 1, 2, 3, 4, 5], [2, 3, 4, 5, 6, 7], 1.0, 0.0)  # identity -> fused rank 1
    # a stride-1 dimension of extent 1 with a padded leading dimension: the
    # remaining dimension 0 is in place but strided (B's padding must stay)
    check(gpu, BF16, BF16, [1, 0, 2], [4, 1, 16], 0.6, 1.0, [5, 1, 18], [3, 6, 18])
    check(gpu, S, S, [1, 0, 2], [4, 1, 16], 1.5, 0.0, [5, 1, 18], [3, 6, 18])
    check(gpu, D, D, [1, 0, 2], [1, 4, 16], 1.5, 0.5, [3, 4, 16], None)
    # rank 16 before fusion, several size-1 dimensions
    perm = [15 - i for i in range(16)]
    size = [2, 1, 2, 1, 2, 1, 2, 1, 1, 2, 1, 1, 1, 2, 1, 1]
    check(gpu, S, S, perm, size, 1.0, 1.0)


def test_empty_tensor_is_a_no_op(gpu):
    import torch
    from ttc_amd import runtime as rt
    plan = rt.Plan(S, S, [1, 0], [0, 5])
    plan.execute(0, 0, 1.0, 0.0)
    torch.cuda.synchronize()


def test_invalid_pointer_alignment_is_an_error(gpu):
    import torch
    from ttc_amd import runtime as rt
    buf = torch.zeros(4096, dtype=torch.uint8, device=gpu)
    plan = rt.Plan(S, S, [1, 0], [16, 16])
    with pytest.raises(rt.TTCError):
        plan.execute(buf.data_ptr() + 1, buf.data_ptr() + 2048, 1.0, 0.0)


def test_denormals_signed_zeros_and_infinities_are_copied_exactly(gpu):
    """alpha = 1, beta = 0 copies are bit copies for all non-NaN values."""
    import torch
    from ttc_amd import runtime as rt
    vals = np.array([0.0, -0.0, 1e-45, -1e-45, 1.17e-38, 1e-40, np.inf, -np.inf, 3.4e38, -2.5], np.float32)
    A = np.resize(vals, 64 * 48).astype(np.float32)
    B = np.zeros_like(A)
    want = oracle_run(S, S, A, B, 1.0, 0.0, [1, 0], [64, 48])
    got, _ = run_hip(gpu, S, S, A, B, 1.0, 0.0, [1, 0], [64, 48])
    assert same_bits(got, want)
    At = A.reshape(64, 48, order="F").T.reshape(-1, order="F")
    assert same_bits(got, At)
    # double precision and beta != 0 with denormal products
    Ad = np.resize(np.array([5e-324, -5e-324, 2.2e-308, 0.0, -0.0], np.float64), 64 * 48)
    Bd = np.resize(np.array([1e-310, -0.0, 1.0], np.float64), 64 * 48)
    want = oracle_run(D, D, Ad, Bd, 0.5, 0.5, [1, 0], [64, 48])
    got, _ = run_hip(gpu, D, D, Ad, Bd, 0.5, 0.5, [1, 0], [64, 48])
    assert same_bits(got, want)


def test_legacy_drop_in_entry(gpu):
    """ttc_legacy_transpose: synchronous, reference wrapper semantics."""
    import torch
    from ttc_amd import runtime as rt
    perm, size = [1, 0, 2], [100, 76, 16]
    A, B = make_case(S, S, perm, size, seed=9)
    want = oracle_run(S, S, A, B, 1.0, 1.0, perm, size, lda=size, ldb=[76, 100, 16])
    dA, dB = to_device(A, gpu), to_device(B, gpu)
    rt.legacy_transpose(S, S, dA, dB, 1.0, 1.0, perm, size, lda=size, ldb=[76, 100, 16],
                        name="sCuTranspose_102_100x76x16")
    got = from_device(dB, B)
    assert same_bits(got, want)


def test_concurrent_streams_share_a_plan(gpu):
    import torch
    from ttc_amd import runtime as rt
    perm, size = [2, 1, 0], [40, 50, 60]
    plan = rt.Plan(D, D, perm, size, beta_is_zero=False)
    cases = []
    streams = [torch.cuda.Stream() for _ in range(3)]
    for i, st in enumerate(streams):
        A, B = make_case(D, D, perm, size, seed=50 + i)
        dA, dB = to_device(A, gpu), to_device(B, gpu)
        torch.cuda.synchronize()
        plan.execute(dA, dB, 1.5, 0.5, stream=st)
        cases.append((A, B, dA, dB))
    torch.cuda.synchronize()
    for A, B, dA, dB in cases:
        assert same_bits(from_device(dB, B), oracle_run(D, D, A, B, 1.5, 0.5, perm, size))
