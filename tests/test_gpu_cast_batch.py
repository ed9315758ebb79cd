"""GPU parity of the two additions beyond the reference: the fp8
cast-transpose with fused amax, and the grouped (batched) launch, including
hipGraph replays."""
import numpy as np
import pytest

from common import (BF16, C, CODE, D, F16, S, Z, bf16_bits, extents, first_diff, from_device, make_case,
                    oracle_run, rand_buffer, same_bits, to_device)
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

E4M3, E5M2 = 6, 7
FP8_MAX = {E4M3: 448.0, E5M2: 57344.0}


def _heavy_input(tA, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    mask = rng.random(n) < 0.001
    v[mask] = rng.standard_t(3, int(mask.sum())) * 30.0
    if tA == S:
        return v.astype(np.float32)
    if tA == BF16:
        return bf16_bits(v.astype(np.float32))
    return v.astype(np.float16).view(np.uint16)


FP8_SHAPES = [([1, 0], [512, 448]), ([0, 2, 1, 3], [128, 64, 32, 8]), ([2, 0, 1], [33, 17, 40]),
              ([1, 0], [1000, 37]), ([0, 2, 1], [7, 9, 5])]


@pytest.mark.parametrize("tA", [S, BF16, F16])
@pytest.mark.parametrize("tB", [E4M3, E5M2])
def test_fp8_cast_transpose_delayed_and_current(gpu, tA, tB):
    import torch
    from ttc_amd import runtime as rt
    for i, (perm, size) in enumerate(FP8_SHAPES):
        for cfg in (None, dict(family=1)):  # the planner's family (row x column or row copy), box tiles
            n = int(np.prod(size))
            A = _heavy_input(tA, n, 40 + i)
            dA = to_device(A, gpu)
            dB = torch.zeros(n, dtype=torch.uint8, device=gpu)
            amax = torch.full((1,), -1.0, dtype=torch.float32, device=gpu)
            sinv = torch.zeros(1, dtype=torch.float32, device=gpu)
            plan = rt.Plan(tA, tB, perm, size, beta_is_zero=True, config=cfg)
            if cfg:
                assert plan.info()["family"] == 1
            for alpha in (1.0, 0.5):
                # delayed scaling with a host scale, then with a device scale
                for scale in (1.0, 3.5):
                    want = np.zeros(n, np.uint8)
                    want_amax = orc.cast_fp8(tA, tB, A, want, alpha, scale, perm, size)
                    plan.execute_fp8(dA, dB, alpha, scale=scale, amax_out=amax, scale_inv_out=sinv)
                    torch.cuda.synchronize()
                    got = dB.cpu().numpy()
                    assert same_bits(got, want), (tA, tB, perm, size, alpha, scale, first_diff(got, want))
                    assert float(amax.item()) == want_amax
                    assert float(sinv.item()) == float(np.float32(1.0) / np.float32(scale))
                    dscale = torch.tensor([scale], dtype=torch.float32, device=gpu)
                    dB.zero_()
                    plan.execute_fp8(dA, dB, alpha, scale=0.0, scale_in=dscale)
                    torch.cuda.synchronize()
                    assert same_bits(dB.cpu().numpy(), want)
                # current scaling: amax pass, then scale = fp8_max / amax
                ref = np.zeros(n, np.uint8)
                a_max = orc.cast_fp8(tA, tB, A, ref, alpha, 1.0, perm, size)
                scale = np.float32(FP8_MAX[tB]) / np.float32(a_max)
                want = np.zeros(n, np.uint8)
                orc.cast_fp8(tA, tB, A, want, alpha, float(scale), perm, size)
                plan.execute_fp8(dA, dB, alpha, amax_out=amax, scale_inv_out=sinv, mode=rt.SCALE_CURRENT)
                torch.cuda.synchronize()
                assert float(amax.item()) == a_max
                assert float(sinv.item()) == float(np.float32(1.0) / scale)
                got = dB.cpu().numpy()
                assert same_bits(got, want), (tA, tB, perm, size, alpha, "current", first_diff(got, want))


@pytest.mark.parametrize("tB", [E4M3, E5M2])
def test_fp8_conversion_every_binary32_input(gpu, tB):
    """Every binary32 bit pattern (NaN, Inf, subnormals and both zeros
    included) through the fp8 cast-transpose with alpha = scale = 1, in
    chunks of 2^28, against torch's cast after clamping to the format's
    range (round to nearest even, NaN -> sign | 0x7f): the saturating
    conversion the oracle restates, bit for bit."""
    import torch
    from ttc_amd import runtime as rt
    lim, tdt = (448.0, torch.float8_e4m3fn) if tB == E4M3 else (57344.0, torch.float8_e5m2)
    side = 1 << 14
    plan = rt.Plan(S, tB, [1, 0], [side, side], beta_is_zero=True)
    B = torch.empty(side * side, dtype=torch.uint8, device=gpu)
    for c in range(16):
        u = (torch.arange(side * side, dtype=torch.int64, device=gpu) + (c << 28)).to(torch.int32)
        x = u.view(torch.float32)
        plan.execute_fp8(x, B, 1.0, scale=1.0)
        want = torch.clamp(x, -lim, lim).to(tdt).view(torch.uint8)
        torch.cuda.synchronize()
        got2d, want2d = B.view(side, side), want.view(side, side).t()
        if not torch.equal(got2d, want2d):
            bad = (got2d != want2d).nonzero()[:8].cpu().tolist()
            raise AssertionError("chunk %d: first mismatches at %s" % (c, bad))
        del u, x, want


def test_fp8_padded_and_misaligned(gpu):
    import torch
    from ttc_amd import runtime as rt
    perm, size = [1, 0], [96, 80]
    lda, ldb = [101, 80], [83, 96]
    na, nb = extents(perm, size, lda, ldb)
    A = _heavy_input(S, na, 7)
    want = np.full(nb, 0x55, np.uint8)
    amax = orc.cast_fp8(S, E4M3, A, want, 1.0, 2.0, perm, size, lda, ldb)
    dA = to_device(A, gpu)
    dB = to_device(np.full(nb + 1, 0x55, np.uint8), gpu)
    damax = torch.zeros(1, dtype=torch.float32, device=gpu)
    plan = rt.Plan(S, E4M3, perm, size, lda, ldb, beta_is_zero=True)
    plan.execute_fp8(dA.data_ptr(), dB.data_ptr() + 1, 1.0, scale=2.0, amax_out=damax)  # B misaligned by 1 byte
    torch.cuda.synchronize()
    assert same_bits(dB.cpu().numpy()[1:], want)
    assert float(damax.item()) == amax


def _batch_ops(rng, n_ops, types, max_elems=60000):
    ops = []
    for i in range(n_ops):
        t = types[int(rng.integers(0, len(types)))]
        r = int(rng.integers(1, 6))
        size = [int(x) for x in rng.integers(1, 24, r)]
        while np.prod(size) > max_elems:
            size[int(np.argmax(size))] //= 2
        perm = [int(x) for x in rng.permutation(r)]
        beta = float(rng.choice([0.0, 1.0, 0.5]))
        alpha = float(rng.uniform(0.5, 2.0))
        pad = rng.random() < 0.2
        lda = [s + int(rng.integers(0, 3)) for s in size] if pad else None
        ldb = [size[perm[q]] + int(rng.integers(0, 3)) for q in range(r)] if pad else None
        A, B = make_case(t, t, perm, size, lda, ldb, seed=1000 + i, dist="normal")
        ops.append(dict(type_a=t, type_b=t, perm=perm, size=size, lda=lda, ldb=ldb, alpha=alpha, beta=beta,
                        hA=A, hB=B))
    # a few larger ops so that work-groups span op boundaries
    for j, (t, perm, size) in enumerate(((S, [1, 0], [700, 900]), (D, [2, 0, 1], [64, 80, 90]),
                                         (BF16, [0, 2, 1], [128, 40, 96]))):
        A, B = make_case(t, t, perm, size, seed=2000 + j)
        ops.append(dict(type_a=t, type_b=t, perm=perm, size=size, lda=None, ldb=None, alpha=1.25, beta=0.75,
                        hA=A, hB=B))
    return ops


def _expected(o):
    return oracle_run(o["type_a"], o["type_b"], o["hA"], o["hB"], o["alpha"], o["beta"], o["perm"], o["size"],
                      o["lda"], o["ldb"])


def test_batch_grouped_and_loop_match_oracle(gpu):
    import torch
    from ttc_amd import runtime as rt
    rng = np.random.default_rng(77)
    ops = _batch_ops(rng, 60, [S, D, BF16, F16, C, Z])
    for o in ops:
        o["A"], o["B"] = to_device(o["hA"], gpu), to_device(o["hB"], gpu)
        o["want"] = _expected(o)
    batch = rt.Batch(ops)
    info = batch.info()
    assert info["tiles"] > 0 and info["workgroups"] > 0
    batch.execute()
    torch.cuda.synchronize()
    for i, o in enumerate(ops):
        got = from_device(o["B"], o["hB"])
        assert same_bits(got, o["want"]), (i, o["type_a"], o["perm"], o["size"], o["alpha"], o["beta"],
                                           first_diff(got, o["want"]))
    # row x column ops staged through two LDS buffers inside the grouped launch
    for o in ops:
        o["B"].copy_(to_device(o["hB"], gpu))
    torch.cuda.synchronize()
    batch2 = rt.Batch([dict(o, config={"lds_buffers": 2}) for o in ops])
    batch2.execute()
    torch.cuda.synchronize()
    for i, o in enumerate(ops):
        assert same_bits(from_device(o["B"], o["hB"]), o["want"]), ("two LDS buffers", i)
    # the per-op launch loop with the same kernels gives the same bits
    for o in ops:
        o["B"].copy_(to_device(o["hB"], gpu))
    torch.cuda.synchronize()
    batch.execute_loop()
    torch.cuda.synchronize()
    for i, o in enumerate(ops):
        assert same_bits(from_device(o["B"], o["hB"]), o["want"]), i


def test_batch_graph_replays(gpu):
    """Each replay starts from the original B (restored eagerly in between)
    and must reproduce the eager result."""
    import torch
    from ttc_amd import runtime as rt
    rng = np.random.default_rng(78)
    ops = _batch_ops(rng, 40, [S, D, BF16])
    stream = torch.cuda.Stream()
    for o in ops:
        o["A"], o["B"] = to_device(o["hA"], gpu), to_device(o["hB"], gpu)
        o["B0"] = o["B"].clone()
        o["want"] = _expected(o)
    batch = rt.Batch(ops)
    for per_op_loop in (False, True):
        for replay in range(3):
            for o in ops:
                o["B"].copy_(o["B0"])
            torch.cuda.synchronize()
            batch.graph_launch(stream, per_op_loop=per_op_loop)
            stream.synchronize()
            for i, o in enumerate(ops):
                assert same_bits(from_device(o["B"], o["hB"]), o["want"]), (per_op_loop, replay, i)
            # unrelated eager work between replays
            junk = torch.randn(1 << 20, device=gpu)
            junk.mul_(2.0).sum().item()


def test_batch_with_tuned_configs_matches_oracle(gpu):
    """Per-op configurations from the batch tuner (every op tuned here) give
    the same bits, grouped and through a graph replay."""
    import torch
    from ttc_amd import autotune
    from ttc_amd import runtime as rt
    rng = np.random.default_rng(79)
    ops = _batch_ops(rng, 24, [S, D, BF16])
    for o in ops:
        o["A"], o["B"] = to_device(o["hA"], gpu), to_device(o["hB"], gpu)
        o["B0"] = o["B"].clone()
        o["want"] = _expected(o)
    cfgs, tuned = autotune.tune_batch(ops, max_impl=3, min_bytes=0, reps=1)
    assert tuned >= 1 and len(cfgs) == len(ops)
    for o in ops:  # the tuner ran the ops on its own buffers; B is untouched
        assert same_bits(from_device(o["B"], o["hB"]), o["hB"])
    batch = rt.Batch([dict(o, config=c) for o, c in zip(ops, cfgs)])
    batch.execute()
    torch.cuda.synchronize()
    for i, o in enumerate(ops):
        assert same_bits(from_device(o["B"], o["hB"]), o["want"]), (i, cfgs[i])
    stream = torch.cuda.Stream()
    for replay in range(2):
        for o in ops:
            o["B"].copy_(o["B0"])
        torch.cuda.synchronize()
        batch.graph_launch(stream)
        stream.synchronize()
        for i, o in enumerate(ops):
            assert same_bits(from_device(o["B"], o["hB"]), o["want"]), (replay, i)


def test_batch_rejects_unsupported_pairs(gpu):
    import torch
    from ttc_amd import runtime as rt
    a = torch.zeros(64, device=gpu)
    with pytest.raises(rt.TTCError):
        rt.Batch([dict(type_a=S, type_b=D, perm=[1, 0], size=[4, 4], A=a, B=a)])
