"""Builds libxtric.so (HIP for gfx950 + host C++) and the `tric` driver in-tree.

    python -m xxtric_amd.build          # or: python xxtric_amd/build.py

Objects go to xxtric_amd/_build/, the library to xxtric_amd/lib/libxtric.so,
the command-line driver to xxtric_amd/bin/tric.  Every HIP source is compiled
with --offload-arch=gfx950 and packed-fp32 code generation disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIBDIR, "libxtric.so")
TRIC = os.path.join(BINDIR, "tric")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_SOURCES = ["api.hip", "orient.hip", "count.hip", "gen.hip", "batch.hip", "serve.hip"]
HOST_SOURCES = ["ingest.cpp", "partition.cpp"]

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off",
              "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r.stdout + r.stderr


def _compile(src):
    name = os.path.basename(src)
    obj = os.path.join(BUILD, name + ".o")
    if name.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    else:
        cmd = [CXX] + HOST_FLAGS + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    srcs += [os.path.join(CSRC, s) for s in HOST_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    tric_src = os.path.join(HOST, "tric.cpp")
    if os.path.exists(tric_src):
        _run([CXX, "-O3", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
              tric_src, "-o", TRIC, "-L" + LIBDIR, "-lxtric", "-Wl,-rpath,$ORIGIN/../lib",
              "-Wl,-rpath," + os.path.join("/opt/rocm", "lib"), "-lpthread"])
    if verbose:
        print("built", LIB)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
