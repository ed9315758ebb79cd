"""Builds the drop-in checks of the emitted artefacts (run by build(), on the CPU).

Emits, with ttc_amd.generator, the per-transposition artefacts a TTC user
links against (ttc/CUDAtranspose.py:395-451 wrapper, names 473-527; files
ttc/ttc.py:1035-1067), and builds them the way a user would:

* libdropin_check.so: the artefacts below in one shared library linked
  against libttc_hip.so, loaded by tests/test_gpu_dropin.py and smoke()
  through ctypes (host int* size/lda/ldb, device A/B, as the reference's
  callers pass them);
* artefact_check: a C program (artefact_check.c) that includes the emitted
  headers and links the artefacts, libttc_hip.so (Cflags/Libs of its
  pkg-config file) and the oracle, as a TTC-C style C caller would.

Artefacts (floatTypeA, floatTypeB, perm, size, beta, configuration):
  sCuTranspose_102_1000x768x16      float, beta != 0      planner's choice
  sCuTranspose_102_1000x768x16_bz   float, beta == 0      planner's choice
  zcCuTranspose_10_300x200          double complex -> float complex, beta != 0
                                    (file zsCuTranspose_10_300x200.hip: the
                                    reference's file-name quirk, ttc/ttc.py:66-70)
  sCuTranspose_10_1000x768          float, beta != 0, with the vector widths
                                    the dense problem allows baked in (called
                                    with a padded lda in the tests)
  sCuTranspose_102_11x9x2           float, beta != 0, the golden fixture's
                                    shape (tests/golden/s.npz case 13)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

LIB = os.path.join(HERE, "libdropin_check.so")
EXE = os.path.join(HERE, "artefact_check")
FIXTURE = os.path.join(HERE, "golden_s_102_11x9x2.bin")
EMIT_DIR = os.path.join(ROOT, "ttc_amd", "_build", "artefacts")
ORACLE_DIR = os.path.join(ROOT, "oracle")

# (floatTypeA, floatTypeB, perm, size, beta, config) — config None: the planner's choice
CASES = [
    ("float", "float", [1, 0, 2], [1000, 768, 16], 1.0, None),
    ("float", "float", [1, 0, 2], [1000, 768, 16], 0.0, None),
    ("double complex", "float complex", [1, 0], [300, 200], 1.0, None),
    ("float", "float", [1, 0], [1000, 768], 1.0, "dense"),
    ("float", "float", [1, 0, 2], [11, 9, 2], 1.0, None),
]


def case_names():
    from ttc_amd import generator
    return [(generator.file_name(a, b, p, s, be), generator.function_name(a, b, p, s, be))
            for a, b, p, s, be, _ in CASES]


def _config(a, b, perm, size, beta, how):
    if how != "dense":
        return {}
    from ttc_amd import runtime as rt
    from ttc_amd import ttc_util
    info = rt.Plan(ttc_util.dtypeCode(a), ttc_util.dtypeCode(b), perm, size, beta_is_zero=beta == 0).info()
    return {k: info[k] for k in ("tile", "vec_a", "vec_b", "family")}


def _write_fixture():
    """Case 13 of tests/golden/s.npz (perm (1,0,2), size (11,9,2),
    alpha = beta = 1) as raw float32: A, B before, B after."""
    import json

    import numpy as np
    with np.load(os.path.join(ROOT, "tests", "golden", "s.npz"), allow_pickle=False) as z:
        meta = json.loads(bytes(z["index"]).decode())
        c = meta["cases"][13]
        assert c["perm"] == [1, 0, 2] and c["size"] == [11, 9, 2] and c["alpha"] == 1.0 and c["beta"] == 1.0
        data = np.concatenate([z["A13"], z["B13"], z["out13"]]).astype(np.float32)
    data.tofile(FIXTURE)


def _pkg_flags(pc_path):
    """Cflags and Libs of a pkg-config file (variables expanded)."""
    vars_, flags = {}, {}
    with open(pc_path) as f:
        for line in f:
            line = line.strip()
            if "=" in line and ":" not in line.split("=")[0]:
                k, v = line.split("=", 1)
                vars_[k] = v
            elif ":" in line:
                k, v = line.split(":", 1)
                flags[k.strip()] = v.strip()
    def expand(v):
        for _ in range(4):
            for k, x in vars_.items():
                v = v.replace("${%s}" % k, x)
        return v.split()
    return expand(flags.get("Cflags", "")), expand(flags.get("Libs", ""))


def build():
    import subprocess

    from ttc_amd import build as hb
    from ttc_amd import generator
    os.makedirs(EMIT_DIR, exist_ok=True)
    srcs = []
    for a, b, perm, size, beta, how in CASES:
        path, _ = generator.emit(EMIT_DIR, a, b, perm, size, beta, _config(a, b, perm, size, beta, how),
                                 note="drop-in check")
        srcs.append(path)
    # the shared library resolves libttc_hip.so next to the package on either machine
    hb.compile_artefacts(srcs, LIB, rpath="$ORIGIN/../../ttc_amd/lib")
    # pkg-config file of the in-tree library, as `install` writes one for a prefix
    pc = os.path.join(hb.BUILD, "ttc_hip.pc")
    with open(pc, "w") as f:
        f.write(hb.pkgconfig_text(ROOT, "${prefix}/ttc_amd/lib", "${prefix}/include"))
    cflags, libs = _pkg_flags(pc)
    libs = [x for x in libs if not x.startswith("-Wl,-rpath")]
    _write_fixture()
    cmd = (["gcc", "-O2", "-std=c11", "-Wall", "-o", EXE, os.path.join(HERE, "artefact_check.c"), "-I" + EMIT_DIR]
           + cflags + ["-L" + HERE, "-ldropin_check"] + libs + ["-L" + ORACLE_DIR, "-lttc_oracle",
                                      "-Wl,-rpath,$ORIGIN:$ORIGIN/../../ttc_amd/lib:$ORIGIN/../../oracle"])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("artefact_check build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return LIB, EXE


if __name__ == "__main__":
    print(build())
