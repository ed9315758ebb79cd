"""The allocation-free grouped launch, ttc_batch_execute_noalloc, against the
allocating one (ttc_batch_create + ttc_batch_execute) through the C ABI:
bit-identical results (and both equal to the oracle) on two calls in a
row on one stream, the second with a smaller set of ops, so that the
second call rewrites the shared workspace after the first call's kernels;
and a work table beyond the workspace is refused with its own status."""
import numpy as np
import pytest

from common import BF16, C, D, F16, S, Z, make_case, oracle_run, same_bits, to_device

pytestmark = pytest.mark.gpu


def _ops(rng, n_ops, seed0):
    ops = []
    types = [S, D, BF16, F16, C, Z]
    for i in range(n_ops):
        t = types[int(rng.integers(0, len(types)))]
        r = int(rng.integers(1, 6))
        size = [int(x) for x in rng.integers(1, 24, r)]
        while np.prod(size) > 60000:
            size[int(np.argmax(size))] //= 2
        perm = [int(x) for x in rng.permutation(r)]
        pad = rng.random() < 0.2
        lda = [s + int(rng.integers(0, 3)) for s in size] if pad else None
        ldb = [size[perm[q]] + int(rng.integers(0, 3)) for q in range(r)] if pad else None
        A, B = make_case(t, t, perm, size, lda, ldb, seed=seed0 + i, dist="normal")
        ops.append(dict(type_a=t, type_b=t, perm=perm, size=size, lda=lda, ldb=ldb,
                        alpha=float(rng.uniform(0.5, 2.0)), beta=float(rng.choice([0.0, 1.0, 0.5])), hA=A, hB=B))
    # larger ops, so that work-group slices cross op boundaries
    for j, (t, perm, size) in enumerate(((S, [1, 0], [700, 900]), (D, [2, 0, 1], [64, 80, 90]))):
        A, B = make_case(t, t, perm, size, seed=seed0 + 500 + j)
        ops.append(dict(type_a=t, type_b=t, perm=perm, size=size, lda=None, ldb=None, alpha=1.25, beta=0.75,
                        hA=A, hB=B))
    return ops


def test_noalloc_matches_allocating_batch_twice(gpu):
    import torch
    from ttc_amd import runtime as rt
    rng = np.random.default_rng(91)
    stream = torch.cuda.Stream()
    calls = [_ops(rng, 60, 9000), _ops(rng, 12, 9600)]  # the second call is smaller
    assert len(calls[1]) < len(calls[0])
    for ops in calls:
        for o in ops:
            o["A"] = to_device(o["hA"], gpu)
            o["Bref"], o["Bnoa"] = to_device(o["hB"], gpu), to_device(o["hB"], gpu)
            o["want"] = oracle_run(o["type_a"], o["type_b"], o["hA"], o["hB"], o["alpha"], o["beta"], o["perm"],
                                   o["size"], o["lda"], o["ldb"])
    torch.cuda.synchronize()
    print("noalloc: two calls of %d and %d ops" % (len(calls[0]), len(calls[1])), flush=True)
    # the allocating entry point
    for ops in calls:
        b = rt.Batch([dict(o, B=o["Bref"]) for o in ops])
        b.execute(stream)
        stream.synchronize()
        b.close()
    # the allocation-free one: two calls back to back on one stream
    for ops in calls:
        rt.batch_execute_noalloc([dict(o, B=o["Bnoa"]) for o in ops], stream)
    stream.synchronize()
    for c, ops in enumerate(calls):
        for i, o in enumerate(ops):
            ref = o["Bref"].cpu().numpy().view(o["hB"].dtype)
            got = o["Bnoa"].cpu().numpy().view(o["hB"].dtype)
            assert same_bits(ref, o["want"]), ("allocating batch vs oracle", c, i)
            assert same_bits(got, ref), ("no-alloc vs allocating batch", c, i, o["type_a"], o["perm"], o["size"])


def test_noalloc_refuses_a_table_beyond_the_workspace(gpu):
    """Every op record takes about 1 KiB of the workspace: 20000 tiny ops
    (sharing one pair of buffers; nothing is launched) do not fit."""
    import torch
    from ttc_amd import runtime as rt
    a = torch.zeros(64, device=gpu)
    b = torch.zeros(64, device=gpu)
    ops = [dict(type_a=S, type_b=S, perm=[1, 0], size=[8, 8], A=a, B=b, alpha=1.0, beta=0.0)] * 20000
    with pytest.raises(rt.TTCError) as e:
        rt.batch_execute_noalloc(ops)
    assert e.value.status == rt.ERROR_WORKSPACE
    torch.cuda.synchronize()
    assert float(b.abs().sum().item()) == 0.0  # nothing ran
