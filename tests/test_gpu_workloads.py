"""GPU parity at the full sizes of the mixed-precision / complex workload
(perm (3,1,0,2), size (256,192,128,96), alpha 1.3, beta 0.7, inputs N(0,1)
with 1% outliers x1e4) and of the fp8 cast-transpose workload (2-D
8192x14336 perm (1,0) and 4-D (128,4096,32,8) perm (0,2,1,3), heavy-tailed
inputs), through a size-independent property: every output element depends
on exactly one input element, so a random sample of output positions is
checked bit for bit against the element formula, and the fused amax against
an independent reduction.  Errors against FP64 are reported (max-abs,
max-ulp)."""
import numpy as np
import pytest

from common import BF16, C, D, S, Z, _real_update

pytestmark = pytest.mark.gpu

C3_PERM, C3_SIZE = [3, 1, 0, 2], [256, 192, 128, 96]
NSAMPLE = 1 << 20


def _offsets(perm, size, rng, n):
    """Sampled A-index tuples -> (flat A offsets, flat B offsets), dense strides."""
    idx = np.stack([rng.integers(0, s, n) for s in size])
    offA = np.zeros(n, np.int64)
    m = 1
    for k, s in enumerate(size):
        offA += idx[k] * m
        m *= s
    offB = np.zeros(n, np.int64)
    m = 1
    for q in range(len(size)):
        offB += idx[perm[q]] * m
        m *= size[perm[q]]
    return offA, offB


def _draw(torch, n, dev, gen, dtype):
    x = torch.randn(n, generator=gen, device=dev, dtype=torch.float64)
    mask = torch.rand(n, generator=gen, device=dev) < 0.01
    x[mask] *= 1e4
    return x.to(dtype)


def _ulp_error(got, exact, dtype):
    """max |got - exact| / ulp(exact) in the output type."""
    e = np.abs(exact)
    ulp = np.spacing(e.astype(dtype)).astype(np.float64)
    ulp[ulp == 0] = np.finfo(dtype).tiny
    return float(np.max(np.abs(got.astype(np.float64) - exact) / ulp))


@pytest.mark.parametrize("code", ["sb", "ds", "c", "z"])
def test_mixed_and_complex_full_size_sampled(gpu, code):
    import torch
    from ttc_amd import runtime as rt
    n = int(np.prod(C3_SIZE))
    gen = torch.Generator(device=gpu)
    gen.manual_seed(3)
    alpha, beta = 1.3, 0.7
    if code == "sb":
        tA, tB, dA, dB = S, BF16, torch.float32, torch.bfloat16
    elif code == "ds":
        tA, tB, dA, dB = D, S, torch.float64, torch.float32
    elif code == "c":
        tA, tB, dA, dB = C, C, torch.complex64, torch.complex64
    else:
        tA, tB, dA, dB = Z, Z, torch.complex128, torch.complex128
    if code in ("c", "z"):
        rd = torch.float32 if code == "c" else torch.float64
        A = torch.complex(_draw(torch, n, gpu, gen, rd), _draw(torch, n, gpu, gen, rd))
        B = torch.complex(_draw(torch, n, gpu, gen, rd), _draw(torch, n, gpu, gen, rd))
    else:
        A = _draw(torch, n, gpu, gen, dA)
        B = _draw(torch, n, gpu, gen, dB)
    rng = np.random.default_rng(33)
    offA, offB = _offsets(C3_PERM, C3_SIZE, rng, NSAMPLE)
    tiA, tiB = torch.from_numpy(offA).to(gpu), torch.from_numpy(offB).to(gpu)
    a_s, b_old = A[tiA].cpu(), B[tiB].cpu()
    plan = rt.Plan(tA, tB, C3_PERM, C3_SIZE, beta_is_zero=False)
    plan.execute(A, B, alpha, beta)
    torch.cuda.synchronize()
    got = B[tiB].cpu()
    del A, B
    if code == "sb":
        want = _real_update(S, BF16, a_s.numpy(), b_old.view(torch.int16).numpy().view(np.uint16), alpha, beta, False)
        assert np.array_equal(got.view(torch.int16).numpy().view(np.uint16), want)
        exact = alpha * a_s.double().numpy() + beta * b_old.double().numpy()
        g = got.float().numpy()
        err_abs = float(np.max(np.abs(g - exact)))
        # bf16 ulp of |exact|: 2^(e-7)
        e = np.floor(np.log2(np.maximum(np.abs(exact), 1e-38)))
        err_ulp = float(np.max(np.abs(g - exact) / np.exp2(e - 7)))
    elif code == "ds":
        want = _real_update(D, S, a_s.numpy(), b_old.numpy(), alpha, beta, False)
        assert np.array_equal(got.numpy().view(np.uint32), want.view(np.uint32))
        exact = alpha * a_s.numpy() + beta * b_old.double().numpy()
        err_abs = float(np.max(np.abs(got.double().numpy() - exact)))
        err_ulp = _ulp_error(got.numpy(), exact, np.float32)
    else:
        r = S if code == "c" else D
        an, bn = a_s.numpy(), b_old.numpy()
        re = _real_update(r, r, np.ascontiguousarray(an.real), np.ascontiguousarray(bn.real), alpha, beta, False)
        im = _real_update(r, r, np.ascontiguousarray(an.imag), np.ascontiguousarray(bn.imag), alpha, beta, False)
        gn = got.numpy()
        assert np.array_equal(gn.real, re) and np.array_equal(gn.imag, im)
        exact = alpha * an.astype(np.complex128) + beta * bn.astype(np.complex128)
        err_abs = float(np.max(np.abs(gn.astype(np.complex128) - exact)))
        dt = np.float32 if code == "c" else np.float64
        err_ulp = max(_ulp_error(gn.real, exact.real, dt), _ulp_error(gn.imag, exact.imag, dt))
    # Reported, not bounded: the bit-exact reference rounds alpha*a and beta*b
    # in the declared types before adding, so where the two cancel the error
    # against the exact FP64 sum is not bounded by an ulp of the result.
    print("%s: %d sampled outputs bit-exact; vs FP64: max-abs %.3g, max-ulp %.3f" % (code, NSAMPLE, err_abs,
                                                                                       err_ulp))
    assert np.isfinite(err_abs)


def test_lane_fixed_walk_with_a_offsets_past_2_30(gpu):
    """fp32 -> bf16 at the full mixed-workload size on [128, 1, 1, 64] tiles:
    A's dimension 3 (25 MB stride) spans 1.6 GB per tile, so the lane-fixed
    walk's A-side table entries pass 2^30; its A walk has no padding slots,
    which is what lets it run.  Sampled outputs bit for bit."""
    import torch
    from ttc_amd import runtime as rt
    n = int(np.prod(C3_SIZE))
    gen = torch.Generator(device=gpu)
    gen.manual_seed(31)
    alpha, beta = 1.3, 0.7
    A = _draw(torch, n, gpu, gen, torch.float32)
    B = _draw(torch, n, gpu, gen, torch.bfloat16)
    rng = np.random.default_rng(34)
    offA, offB = _offsets(C3_PERM, C3_SIZE, rng, NSAMPLE)
    tiA, tiB = torch.from_numpy(offA).to(gpu), torch.from_numpy(offB).to(gpu)
    a_s, b_old = A[tiA].cpu(), B[tiB].cpu()
    plan = rt.Plan(S, BF16, C3_PERM, C3_SIZE, beta_is_zero=False, config={"family": 2, "tile": [128, 1, 1, 64]})
    info = plan.info()
    assert info["lane_fixed"] == 1 and info["tile"][:4] == [128, 1, 1, 64], info
    plan.execute(A, B, alpha, beta)
    torch.cuda.synchronize()
    got = B[tiB].cpu()
    del A, B
    want = _real_update(S, BF16, a_s.numpy(), b_old.view(torch.int16).numpy().view(np.uint16), alpha, beta, False)
    assert np.array_equal(got.view(torch.int16).numpy().view(np.uint16), want)


FP8_CASES = [([1, 0], [8192, 14336]), ([0, 2, 1, 3], [128, 4096, 32, 8])]


@pytest.mark.parametrize("tA", [S, BF16])
@pytest.mark.parametrize("case", [0, 1])
def test_fp8_full_size_current_scaling_sampled(gpu, tA, case):
    import torch
    from ttc_amd import runtime as rt
    perm, size = FP8_CASES[case]
    n = int(np.prod(size))
    gen = torch.Generator(device=gpu)
    gen.manual_seed(4)
    x = torch.randn(n, generator=gen, device=gpu)
    heavy = torch.rand(n, generator=gen, device=gpu) < 0.001
    x[heavy] = torch.distributions.StudentT(3.0).sample((int(heavy.sum().item()),)).to(gpu)
    A = x if tA == S else x.to(torch.bfloat16)
    a32 = A.float()
    amax_ref = float(a32.abs().max().item())  # alpha = 1: |fl32(1 * a)| = |a|
    rng = np.random.default_rng(44 + case)
    offA, offB = _offsets(perm, size, rng, NSAMPLE)
    a_s = a32[torch.from_numpy(offA).to(gpu)].cpu()
    for tB, fmax, tdt in ((rt.E4M3, 448.0, torch.float8_e4m3fn), (rt.E5M2, 57344.0, torch.float8_e5m2)):
        B = torch.zeros(n, dtype=torch.uint8, device=gpu)
        amax = torch.zeros(1, dtype=torch.float32, device=gpu)
        sinv = torch.zeros(1, dtype=torch.float32, device=gpu)
        plan = rt.Plan(tA, tB, perm, size, beta_is_zero=True)
        plan.execute_fp8(A, B, 1.0, amax_out=amax, scale_inv_out=sinv, mode=rt.SCALE_CURRENT)
        torch.cuda.synchronize()
        assert float(amax.item()) == amax_ref
        scale = np.float32(fmax) / np.float32(amax_ref)
        assert float(sinv.item()) == float(np.float32(1.0) / scale)
        got = B[torch.from_numpy(offB).to(gpu)].cpu().numpy()
        y = (a_s * torch.tensor(float(scale), dtype=torch.float32)).clamp(-fmax, fmax)  # fl32(a * scale), saturated
        want = y.to(tdt).view(torch.uint8).numpy()
        assert np.array_equal(got, want), (tA, tB, case)
        # against quantising the FP64 product: differs only where fl32 rounding moves a tie
        y64 = torch.from_numpy(a_s.double().numpy() * float(scale)).clamp(-fmax, fmax)
        q64 = y64.to(tdt).view(torch.uint8).numpy()
        frac = float(np.mean(q64 != got))
        print("fp8 %s->%s case %d: %d sampled bit-exact; differs from FP64 quantisation at %.2e of samples" % (
            tA, tB, case, NSAMPLE, frac))
        assert frac < 1e-3
        del B
