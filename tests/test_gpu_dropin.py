"""GPU checks of the drop-in boundary as the reference's callers use it.

* Emitted per-transposition symbols (<p>CuTranspose_<perm>_<size>[_bz],
  ttc/CUDAtranspose.py:395-451, names 473-527), built into a shared library
  against libttc_hip.so and called through ctypes with device A/B and HOST
  int arrays size/lda/ldb, exactly as the reference wrapper's callers pass
  them; results bit-exact against the oracle.
* A configuration baked in for the dense problem still serves a call with a
  padded lda (the wrapper takes lda/ldb at call time; the reference
  kernels accept any).
* A C program linking the artefacts and libttc_hip.so (tests/c/artefact_check.c).
* The reference test driver's sequence (test-framework/ttc_test.py:124-156,
  `import ttc; ttc.ttc.generateTransposition(args)`) with the search on the GPU.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from common import C, S, Z, make_case, oracle_run, same_bits

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "c", "libdropin_check.so")
EXE = os.path.join(HERE, "c", "artefact_check")
FIXTURE = os.path.join(HERE, "c", "golden_s_102_11x9x2.bin")


def _lib():
    from ttc_amd import runtime as rt
    rt.lib()  # libttc_hip.so first (shares torch's HIP runtime)
    if not os.path.exists(LIB):
        pytest.fail("drop-in checks not built (tests/c/dropin_build.py; build() builds them)")
    return ctypes.CDLL(LIB)


def _ints(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def _call(fn, dA, dB, alpha, beta, size, lda, ldb, real_a, real_b):
    vp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    fn.restype = None
    if beta is None:
        fn.argtypes = [vp, vp, real_a, ip, ip, ip]
        fn(dA.data_ptr(), dB.data_ptr(), alpha, _ints(size), _ints(lda), _ints(ldb))
    else:
        fn.argtypes = [vp, vp, real_a, real_b, ip, ip, ip]
        fn(dA.data_ptr(), dB.data_ptr(), alpha, beta, _ints(size), _ints(lda), _ints(ldb))


@pytest.mark.parametrize("bz", [False, True])
def test_emitted_symbol_c1_shape(gpu, bz):
    import torch
    from common import from_device, to_device
    L = _lib()
    perm, size = [1, 0, 2], [1000, 768, 16]
    A, B = make_case(S, S, perm, size, seed=11, dist="uniform")
    alpha, beta = 1.3, (0.0 if bz else 0.7)
    want = oracle_run(S, S, A, B, np.float32(alpha), np.float32(beta), perm, size)
    dA, dB = to_device(A, gpu), to_device(B, gpu)
    fn = getattr(L, "sCuTranspose_102_1000x768x16" + ("_bz" if bz else ""))
    _call(fn, dA, dB, alpha, None if bz else beta, size, size, [768, 1000, 16], ctypes.c_float, ctypes.c_float)
    torch.cuda.synchronize()  # the symbol is synchronous already
    assert same_bits(from_device(dB, B), want)


def test_emitted_symbol_zc_mixed_complex(gpu):
    """double complex A -> float complex B: file zsCuTranspose_10_300x200.hip,
    symbol zcCuTranspose_10_300x200, alpha double, beta float."""
    from common import from_device, to_device
    L = _lib()
    assert not hasattr(L, "zsCuTranspose_10_300x200")
    perm, size = [1, 0], [300, 200]
    A, B = make_case(Z, C, perm, size, seed=12, dist="normal")
    want = oracle_run(Z, C, A, B, 0.9, np.float32(-1.25), perm, size)
    dA, dB = to_device(A, gpu), to_device(B, gpu)
    _call(L.zcCuTranspose_10_300x200, dA, dB, 0.9, -1.25, size, size, [200, 300], ctypes.c_double, ctypes.c_float)
    assert same_bits(from_device(dB, B), want)


def test_baked_configuration_with_padded_lda(gpu):
    """sCuTranspose_10_1000x768 carries the dense problem's vector widths;
    a padded lda (1001) rules them out, and the call still succeeds."""
    from common import from_device, to_device
    L = _lib()
    perm, size = [1, 0], [1000, 768]
    for lda, ldb in (([1000, 768], [768, 1000]), ([1001, 768], [768, 1000]), ([1003, 768], [771, 1000])):
        A, B = make_case(S, S, perm, size, lda, ldb, seed=13, dist="uniform")
        want = oracle_run(S, S, A, B, np.float32(2.0), np.float32(-0.5), perm, size, lda, ldb)
        dA, dB = to_device(A, gpu), to_device(B, gpu)
        _call(L.sCuTranspose_10_1000x768, dA, dB, 2.0, -0.5, size, lda, ldb, ctypes.c_float, ctypes.c_float)
        assert same_bits(from_device(dB, B), want), (lda, ldb)


def test_c_program_links_artefacts(gpu):
    if not os.path.exists(EXE):
        pytest.fail("artefact_check not built (tests/c/dropin_build.py; build() builds it)")
    print("running", EXE, flush=True)
    r = subprocess.run([EXE, FIXTURE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "artefact_check ok" in r.stdout


def test_reference_driver_generates_on_gpu(gpu, tmp_path, monkeypatch):
    """ttc_test.py's call sequence, searching on the GPU (4 variants)."""
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("TTC_ROOT", str(tmp_path / "root"))
    import ttc
    from ttc import ttc_util
    args = ttc_util.TTCargs([2, 1, 0], [64, 40, 48])
    args.alpha, args.beta = 1.0, 1.1
    args.affinity = ""
    args.numThreads = 8
    args.floatTypeA, args.floatTypeB = "float", "float"
    args.streamingStores = 0
    args.maxNumImplementations = 4
    args.ignoreDatabase = 1
    args.lda, args.ldb = [], []
    args.debug = 0
    args.architecture = "cuda"
    args.align = 1
    args.blockings, args.loopPermutations, args.prefetchDistances = [], [], []
    args.scalar = 0
    args.silent = 1
    args.hotA = args.hotB = 0
    (name, bw) = ttc.ttc.generateTransposition(args)
    assert name == "sCuTranspose_210_64x40x48" and bw > 0
    assert (tmp_path / "ttc_transpositions" / (name + ".hip")).exists()
