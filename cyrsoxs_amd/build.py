"""Builds the gfx950 HIP library cyrsoxs_amd/lib/libcyrsoxs.so (and, when the
pybind11 and Python headers are present, the CyRSoXS extension module and the
CyRSoXS command-line executable) in-tree with hipcc.

    python -m cyrsoxs_amd.build [--jobs N] [--only lib]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CYRSOXS_ARCH", "gfx950")

LIB_SOURCES = ["runtime.cpp", "host_math.cpp", "fft.hip", "polarization.hip", "ewald.hip", "ebasis.hip",
               "mbasis.hip", "rotate.hip", "pipeline.hip", "noalloc.hip", "engine.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          # packed FP32 instructions are excluded from all device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I", os.path.join(REPO, "include")]
# host geometry must be evaluated exactly as written (bit-exact masks / NN indices)
PER_FILE = {"host_math.cpp": ["-ffp-contract=off"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _headers_mtime():
    """Newest modification time of the library's headers (csrc/*.hpp, include/*.h)."""
    t = 0.0
    for d, ext in ((CSRC, ".hpp"), (os.path.join(REPO, "include"), ".h")):
        for f in os.listdir(d):
            if f.endswith(ext):
                t = max(t, os.path.getmtime(os.path.join(d, f)))
    return t


def _compile(src, force=True, hdr_mtime=0.0):
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if not force and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t >= os.path.getmtime(path) and t >= hdr_mtime and t >= os.path.getmtime(__file__):
            return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *COMMON, *PER_FILE.get(os.path.basename(src), []), *lang, "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_lib(jobs=8, force=True):
    """Compiles every source (force) or only those older than their source, a
    header or this script, then links lib/libcyrsoxs.so."""
    hdr = _headers_mtime()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force, hdr), LIB_SOURCES))
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libcyrsoxs.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


HOST_SOURCES = ["host/config.cpp", "host/input_data.cpp", "host/h5io.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-I", os.path.join(REPO, "include")]


def _compile_host(src, extra=()):
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, src.replace("/", "_") + ".o")
    _run([CXX, *HOST_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_cli(jobs=8):
    """The CyRSoXS executable (reference src/main.cpp) next to the library."""
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_host, HOST_SOURCES + ["host/main.cpp"]))
    out = os.path.join(LIBDIR, "CyRSoXS")
    _run([CXX, "-o", out + ".tmp", *objs, "-L", LIBDIR, "-lcyrsoxs", "-Wl,-rpath,$ORIGIN", "-ldl"])
    os.replace(out + ".tmp", out)
    return out


def build_pymodule(jobs=8):
    """The pybind11 module CyRSoXS (reference src/pymain.cpp) in the package directory."""
    try:
        import pybind11
    except ImportError:
        return None
    pyinc = sysconfig.get_paths()["include"]
    if not os.path.exists(os.path.join(pyinc, "Python.h")):
        return None
    extra = ["-I", pybind11.get_include(), "-I", pyinc, "-fvisibility=hidden"]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_host, HOST_SOURCES))
        objs.append(ex.submit(_compile_host, "host/pymain.cpp", extra).result())
    out = os.path.join(HERE, "CyRSoXS" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run([CXX, "-shared", "-o", out + ".tmp", *objs, "-L", LIBDIR, "-lcyrsoxs", "-Wl,-rpath,$ORIGIN/lib", "-ldl"])
    os.replace(out + ".tmp", out)
    return out


def build_all(jobs=8, only=None, force=True):
    lib = build_lib(jobs, force)
    if only == "lib":
        return lib
    build_cli(jobs)
    build_pymodule(jobs)
    return lib


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 4))
    ap.add_argument("--only", default=None)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources older than their source file or a header")
    a = ap.parse_args(argv)
    print(build_all(a.jobs, a.only, force=not a.incremental))


if __name__ == "__main__":
    main()
