"""Emission of drop-in artefacts: <name>.hip and <name>.h.

The reference writes ./ttc_transpositions/<name>.cu (kernels and an
`extern "C"` host wrapper) and <name>.h (its prototype) for the fastest
variant (ttc/ttc.py:1035-1067, ttc/GPUtransposeGenerator.py:303-314,
ttc/CUDAtranspose.py:395-451).  Here the kernels live in libttc_hip.so, so
the emitted .hip holds the same `extern "C"` symbol with the same signature,
with the tuned configuration baked in, forwarding to ttc_legacy_transpose
(synchronous, exit(-1) on error — the reference wrapper's contract).  Build it
with  hipcc -c <name>.hip -I<repo>/include  and link against libttc_hip.so.
"""
import os

from . import ttc_util

# C types of the emitted signature.  Complex types keep the float2/double2
# layout of cuFloatComplex/cuDoubleComplex.
C_TYPES = {"float": "float", "double": "double", "float complex": "hipFloatComplex",
           "double complex": "hipDoubleComplex", "bfloat16": "uint16_t", "half": "uint16_t",
           "fp8_e4m3": "uint8_t", "fp8_e5m2": "uint8_t"}
ENUM = {"float": "TTC_S", "double": "TTC_D", "float complex": "TTC_C", "double complex": "TTC_Z",
        "bfloat16": "TTC_BF16", "half": "TTC_F16", "fp8_e4m3": "TTC_E4M3", "fp8_e5m2": "TTC_E5M2"}
_CUDA_LETTER = {"float": "s", "double": "d", "float complex": "c", "double complex": "z"}


def real_type(floatType):
    """alpha/beta type: the real type of the element (ttc/CUDAtranspose.py:22-32)."""
    return "double" if floatType in ("double", "double complex") else "float"


def function_name(floatTypeA, floatTypeB, perm, size, beta):
    """<p>CuTranspose_<perm>_<size>[_bz] as ttc/CUDAtranspose.py:473-527
    builds it for the emitted wrapper (A's letter, then B's if different)."""
    la = _CUDA_LETTER.get(floatTypeA, ttc_util.FLOAT_TYPES[floatTypeA][2])
    lb = _CUDA_LETTER.get(floatTypeB, ttc_util.FLOAT_TYPES[floatTypeB][2])
    name = la + ("" if floatTypeA == floatTypeB else lb) + "CuTranspose_"
    name += "".join(str(p) for p in perm) + "_" + "x".join(str(s) for s in size)
    if beta == 0:
        name += "_bz"
    return name


def file_name(floatTypeA, floatTypeB, perm, size, beta):
    """File name as ttc/ttc.py:41-95 builds it for the CUDA compiler:
    s, sd, d, ds, c, cz, z and (sic) "zs" for double complex -> float
    complex, then "Cu"."""
    quirk = {("float", "float"): "s", ("float", "double"): "sd", ("double", "double"): "d",
             ("double", "float"): "ds", ("float complex", "float complex"): "c",
             ("float complex", "double complex"): "cz", ("double complex", "double complex"): "z",
             ("double complex", "float complex"): "zs"}
    p = quirk.get((floatTypeA, floatTypeB), ttc_util.getFloatPrefix(floatTypeA, floatTypeB))
    name = p + "CuTranspose_" + "".join(str(x) for x in perm) + "_" + "x".join(str(s) for s in size)
    if beta == 0:
        name += "_bz"
    return name


def prototype(fname, floatTypeA, floatTypeB, beta):
    ta, tb = C_TYPES[floatTypeA], C_TYPES[floatTypeB]
    al, be = real_type(floatTypeA), real_type(floatTypeB)
    if beta != 0:
        return ('extern "C" void %s(const %s *A, %s *B, const %s alpha, const %s beta, const int *size, '
                'const int *lda, const int *ldb)' % (fname, ta, tb, al, be))
    return ('extern "C" void %s(const %s *A, %s *B, const %s alpha, const int *size, const int *lda, '
            'const int *ldb)' % (fname, ta, tb, al))


def config_initializer(cfg):
    tile = list(cfg.get("tile", [])) + [0] * (8 - len(cfg.get("tile", [])))
    return "{{%s}, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d}" % (", ".join(str(int(t)) for t in tile[:8]),
                                                               cfg.get("vec_a", 0), cfg.get("vec_b", 0),
                                                               cfg.get("wg_per_cu", 0), cfg.get("nt_store", 0),
                                                               cfg.get("grid_order", 0), cfg.get("run_a", 0),
                                                               cfg.get("run_b", 0), cfg.get("family", 0),
                                                               cfg.get("tile_order", 0), cfg.get("lds_buffers", 0))


def render(floatTypeA, floatTypeB, perm, size, beta, cfg, note=""):
    """(hip source, header) texts for one transposition (fused perm/size)."""
    fname = function_name(floatTypeA, floatTypeB, perm, size, beta)
    proto = prototype(fname, floatTypeA, floatTypeB, beta)
    guard = (fname + "_H").upper()
    # The prototype sits in an extern "C" block so that C callers (TTC-C)
    # can include the header as well as C++ ones.
    hdr = ["#ifndef %s" % guard, "#define %s" % guard, "#include <hip/hip_complex.h>", "#include <stdint.h>",
           "/* B(%s) <- alpha * A(%s)%s */" % (",".join("i%d" % p for p in perm),
                                              ",".join("i%d" % i for i in range(len(perm))),
                                              " + beta * B(...)" if beta != 0 else ""),
           "#ifdef __cplusplus", 'extern "C" {', "#endif",
           proto[len('extern "C" '):] + ";",
           "#ifdef __cplusplus", "}", "#endif", "#endif", ""]
    src = ["// Generated by ttc_amd %s%s" % (".".join(str(v) for v in ttc_util.getVersion()),
                                            (" -- " + note) if note else ""),
           "// B[pi(i)] = alpha * A[i]%s with pi = (%s), size = (%s)." % (
               " + beta * B[pi(i)]" if beta != 0 else "", ",".join(map(str, perm)), ",".join(map(str, size))),
           '#include "ttc_hip.h"', '#include "%s.h"' % file_name(floatTypeA, floatTypeB, perm, size, beta), "",
           proto, "{",
           "  static const int perm[%d] = {%s};" % (len(perm), ", ".join(str(p) for p in perm)),
           "  static const ttc_config_t cfg = %s;" % config_initializer(cfg or {}),
           "  ttc_legacy_transpose(%s, %s, A, B, (double)alpha, %s, %d, %d, perm, size, lda, ldb, &cfg, \"%s\");" % (
               ENUM[floatTypeA], ENUM[floatTypeB], "(double)beta" if beta != 0 else "0.0", 1 if beta == 0 else 0,
               len(perm), fname),
           "}", ""]
    return "\n".join(src), "\n".join(hdr)


def emit(directory, floatTypeA, floatTypeB, perm, size, beta, cfg, note=""):
    """Writes <file>.hip and <file>.h under `directory`; returns (path, function name)."""
    os.makedirs(directory, exist_ok=True)
    base = file_name(floatTypeA, floatTypeB, perm, size, beta)
    src, hdr = render(floatTypeA, floatTypeB, perm, size, beta, cfg, note)
    with open(os.path.join(directory, base + ".hip"), "w") as f:
        f.write(src)
    with open(os.path.join(directory, base + ".h"), "w") as f:
        f.write(hdr)
    return os.path.join(directory, base + ".hip"), function_name(floatTypeA, floatTypeB, perm, size, beta)
