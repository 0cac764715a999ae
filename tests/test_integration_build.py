"""The drop-in headers compile and link against the reference's own code.

CPU only, and only where the reference sources and MPICH are mounted: these
tests build, they never run what they build (no device here).

* The engine drop-in: the reference's main.cpp with the include switch a
  maintainer applies (the xxtric-gpu.hpp branch, main.cpp:21-33, replaced by
  xxtric_single.hpp), compiled with -DUSE_SINGLE_DEVICE -DUSE_XTRIC in a
  temporary copy and linked against libxtric.so.  main.cpp builds
  TriangleCountingSingle from the reference's Graph (main.cpp:148-149).
* The launcher drop-in: tests/integration/reference_callsites.cpp calls the
  counting.cuh launchers through counting_xtric.hpp with the reference's
  argument shapes (xxtric-gpu.hpp:195-263), stream addresses included.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib/libmpi.so.12"
LIBDIR = os.path.join(ROOT, "xxtric_amd", "lib")

needs_ref = pytest.mark.skipif(
    not (os.path.isfile(os.path.join(REF, "main.cpp")) and os.path.exists(MPI_LIB)
         and os.path.exists(os.path.join(LIBDIR, "libxtric.so"))),
    reason="needs the reference sources, MPICH and a built libxtric.so")

SWITCH_FROM = '#if defined(USE_CUDA)\n#include "xxtric-gpu.hpp"'
SWITCH_TO = ('#if defined(USE_SINGLE_DEVICE) && defined(USE_XTRIC)\n#include "xxtric_single.hpp"\n'
             '#elif defined(USE_CUDA)\n#include "xxtric-gpu.hpp"')


def _cxx(args, cwd):
    # g++ with MPICH's header and library named directly, not its mpicxx
    # wrapper: the wrapper puts /opt/conda/lib first on the library path,
    # where an older libstdc++ lacks what libxtric.so and the ROCm runtime need.
    r = subprocess.run(["g++", "-I" + MPI_INC] + args, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def _link_flags():
    return [MPI_LIB, "-L" + LIBDIR, "-lxtric", "-Wl,-rpath," + LIBDIR + ":/opt/conda/lib",
            "-Wl,-rpath-link,/opt/rocm/lib"]


@needs_ref
def test_reference_main_builds_on_engine_dropin(tmp_path):
    src = open(os.path.join(REF, "main.cpp")).read()
    assert SWITCH_FROM in src
    main = tmp_path / "main.cpp"
    main.write_text(src.replace(SWITCH_FROM, SWITCH_TO, 1))
    exe = tmp_path / "tric_xtric"
    _cxx(["-std=c++20", "-O1", "-w", "-DUSE_SINGLE_DEVICE", "-DUSE_XTRIC", "-iquote", REF,
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "xxtric_amd", "host"),
          str(main), "-o", str(exe)] + _link_flags(), cwd=str(tmp_path))
    assert exe.exists()
    # The engine came from libxtric, not from the reference's CUDA engine.
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True, text=True).stdout
    for sym in ("xt_ctx_create", "xt_graph_from_csr", "xt_graph_count", "xt_graph_destroy"):
        assert sym in nm, sym


@needs_ref
def test_reference_launcher_callsites_build(tmp_path):
    exe = tmp_path / "callsites"
    _cxx(["-std=c++20", "-O1", "-w", "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          "-I" + REF, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "xxtric_amd", "host"),
          os.path.join(ROOT, "tests", "integration", "reference_callsites.cpp"), "-o", str(exe)]
         + _link_flags() + ["-L/opt/rocm/lib", "-lamdhip64"], cwd=str(tmp_path))
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True, text=True).stdout
    for sym in ("xt_local_count_gpu_inner", "xt_single_gpu_count_masked", "xt_single_gpu_count"):
        assert sym in nm, sym


@needs_ref
def test_engine_header_brings_no_graph_type(tmp_path):
    """xxtric_single.hpp declares no Graph/Edge of its own: with only the
    reference's graph.hpp before it, a second definition would not compile,
    and without any Graph the class still parses (it is a template on G)."""
    f = tmp_path / "t.cpp"
    f.write_text('#include "xxtric_single.hpp"\nstruct Graph;\nstruct Edge;\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "xxtric_amd", "host"), str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert shutil.which("g++")
