"""CPU tests of the drop-in surface beyond the command line: the `ttc`
import name and command (the reference's package and scripts/ttc), the
reference test driver's calling sequence, the installed C library (header,
pkg-config and CMake files), and the emitted artefacts built for the GPU
checks (their names and exported symbols)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture()
def workdir(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("TTC_ROOT", str(tmp_path / "root"))
    return tmp_path


def test_ttc_import_name_is_the_reference_surface():
    import ttc
    from ttc import sql_util, ttc_util
    import ttc_amd
    assert ttc.ttc.generateTransposition is ttc_amd.ttc.generateTransposition
    assert ttc.main is ttc_amd.ttc.main and ttc.ttc.main is ttc_amd.ttc.main
    assert ttc_util.TTCargs is ttc_amd.ttc_util.TTCargs and ttc_util.fuseIndices is ttc_amd.ttc_util.fuseIndices
    assert sql_util.getBestConfig is ttc_amd.sql_util.getBestConfig


def test_reference_test_driver_sequence(workdir):
    """The attribute set and call of test-framework/ttc_test.py:124-156,
    unchanged except --generateOnly (no GPU here)."""
    import ttc
    from ttc import ttc_util
    args = ttc_util.TTCargs([2, 1, 0], [16, 8, 12])
    args.alpha, args.beta = 1.0, 1.1
    args.affinity = ""
    args.numThreads = 8
    args.floatTypeA, args.floatTypeB = "double", "double"
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
    args.generateOnly = 1
    (name, bw) = ttc.ttc.generateTransposition(args)
    assert name == "dCuTranspose_210_16x8x12" and bw == 0
    assert (workdir / "ttc_transpositions" / (name + ".hip")).exists()
    assert "--perm=2,1,0" in args.getCommandLineString()


def test_python_m_ttc_and_console_script(workdir):
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-m", "ttc", "--perm=1,0", "--size=40,30", "--dataType=z", "--beta=1",
                        "--generateOnly"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
    assert r.returncode == 0, r.stdout
    assert (workdir / "ttc_transpositions" / "zCuTranspose_10_40x30.hip").exists()
    try:
        import tomllib
    except ImportError:
        import tomli as tomllib
    with open(os.path.join(ROOT, "pyproject.toml"), "rb") as f:
        meta = tomllib.load(f)
    target = meta["project"]["scripts"]["ttc"]
    assert target == "ttc.ttc:main"  # scripts/ttc:2 of the reference
    import importlib
    mod, fn = target.split(":")
    assert callable(getattr(importlib.import_module(mod), fn))
    assert set(meta["tool"]["setuptools"]["packages"]) == {"ttc", "ttc_amd"}


def test_install_prefix_pkgconfig_and_cmake(tmp_path):
    from ttc_amd import build as hb
    if not os.path.exists(hb.LIB):
        pytest.skip("libttc_hip.so not built")
    prefix = hb.install(str(tmp_path / "prefix"))
    for rel in ("lib/libttc_hip.so", "include/ttc_hip.h", "lib/pkgconfig/ttc_hip.pc",
                "lib/cmake/ttc_hip/ttc_hip-config.cmake"):
        assert os.path.exists(os.path.join(prefix, rel)), rel
    pc = open(os.path.join(prefix, "lib/pkgconfig/ttc_hip.pc")).read()
    assert "prefix=%s" % prefix in pc and "Libs: -L${libdir}" in pc and "-lttc_hip" in pc
    assert "Cflags: -I${includedir}" in pc
    # a CMake project finds the package and its imported target
    proj = tmp_path / "proj"
    proj.mkdir()
    (proj / "CMakeLists.txt").write_text(
        "cmake_minimum_required(VERSION 3.16)\nproject(p NONE)\n"
        "find_package(ttc_hip CONFIG REQUIRED)\n"
        "get_target_property(loc ttc::ttc_hip IMPORTED_LOCATION)\n"
        "get_target_property(inc ttc::ttc_hip INTERFACE_INCLUDE_DIRECTORIES)\n"
        "message(STATUS \"TTCLOC=${loc}\")\nmessage(STATUS \"TTCINC=${inc}\")\n")
    r = subprocess.run(["cmake", "-S", str(proj), "-B", str(tmp_path / "b"), "-DCMAKE_PREFIX_PATH=" + prefix],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    assert "TTCLOC=%s/lib/libttc_hip.so" % prefix in r.stdout and "TTCINC=%s/include" % prefix in r.stdout
    # the installed library exports every symbol the installed header declares
    from ttc_amd import runtime as rt
    nm = subprocess.run(["nm", "-D", "--defined-only", os.path.join(prefix, "lib/libttc_hip.so")],
                        stdout=subprocess.PIPE, text=True).stdout
    for sym in rt.EXPORTS:
        assert " T %s\n" % sym in nm, sym


def test_emitted_artefacts_of_the_gpu_checks():
    """The artefacts the GPU checks link: the zc one lives in a file named
    "zs..." (ttc/ttc.py:66-70) while its symbol is "zc..."
    (ttc/CUDAtranspose.py:485-493); the built library exports each symbol."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "c"))
    import dropin_build
    names = dropin_build.case_names()
    assert ("zsCuTranspose_10_300x200", "zcCuTranspose_10_300x200") in names
    assert ("sCuTranspose_102_1000x768x16_bz", "sCuTranspose_102_1000x768x16_bz") in names
    if not os.path.exists(dropin_build.LIB):
        pytest.skip("drop-in checks not built")
    nm = subprocess.run(["nm", "-D", "--defined-only", dropin_build.LIB], stdout=subprocess.PIPE, text=True).stdout
    for _, fn in names:
        assert " T %s\n" % fn in nm, fn
    assert "zsCuTranspose" not in nm
    assert os.path.exists(dropin_build.EXE) and os.path.exists(dropin_build.FIXTURE)


def test_library_reports_the_digest_of_its_sources():
    """ttc_version() carries a digest of the sources the library was built
    from; a library older than the checkout's sources reports another."""
    from ttc_amd import build as hb
    from ttc_amd import runtime as rt
    if not os.path.exists(hb.LIB):
        pytest.skip("libttc_hip.so not built")
    v = rt.version()
    assert v.startswith("ttc-mi355x %s (gfx950) src:" % hb.VERSION) and len(v.split("src:")[1]) == 16
    assert v == hb.version_string(), "libttc_hip.so is stale: rebuild with python -m ttc_amd.build"


def test_throughput_mode_rejects_what_it_cannot_check():
    from ttc_amd import runtime as rt
    from ttc_amd import throughput
    with pytest.raises(ValueError):
        throughput.run_streams(rt.C, rt.C, [1, 0], [4, 4])
    with pytest.raises(ValueError):
        throughput.run_processes(rt.S, rt.S, [1, 0], [4, 4], processes=17)
