"""Build recipe of the HIP libraries (in-tree, for gfx950).

    python -m astaroth_amd.build          # both precisions
    python -m astaroth_amd.build --fp32   # one precision

Each precision is one shared library, astaroth_amd/lib/libastaroth_{fp32,fp64}.so, exporting the
C ABI of include/astaroth.h (AC_DOUBLE_PRECISION selects AcReal). Translation units compile in
parallel; packed FP32 instructions are disabled for every HIP file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(REPO, "_build")
INCLUDE = os.path.join(REPO, "include")

HIP_SOURCES = ["device.hip", "solve.hip", "boundconds.hip", "reductions.hip", "init.hip", "halo.hip",
               "ensemble.hip"]
HOST_SOURCES = ["host.cc", "legacy.cc", "simulation.cc", "fortran.cc", "node.cc"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def lib_path(double: bool) -> str:
    return os.path.join(LIBDIR, "libastaroth_%s.so" % ("fp64" if double else "fp32"))


def _compile(args):
    cmd, src = args
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (src, " ".join(cmd), proc.stderr))
    return src, proc.stderr


def build_precision(double: bool, jobs: int = 8, verbose: bool = False) -> str:
    tag = "fp64" if double else "fp32"
    objdir = os.path.join(BUILD, tag)
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hip = hipcc()
    common = [
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-I" + INCLUDE,
        "-I" + CSRC,
        "-DAC_DOUBLE_PRECISION=%d" % int(double),
        "-Wall",
        "-Wno-unused-function",
    ]
    hipflags = [
        "--offload-arch=" + ARCH,
        "-x",
        "hip",
        # packed FP32 (v_pk_*_f32) disabled for all device code
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-munsafe-fp-atomics",
    ]
    jobs_list = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs_list.append(([hip] + hipflags + common + ["-c", os.path.join(CSRC, src), "-o", obj], src))
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src.replace(".cc", ".o"))
        objs.append(obj)
        jobs_list.append(([hip] + common + ["-c", os.path.join(CSRC, src), "-o", obj], src))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for src, err in ex.map(_compile, jobs_list):
            if verbose and err.strip():
                print("[%s] %s" % (src, err.strip()[-2000:]))
    out = lib_path(double)
    tmp = out + ".tmp"
    link = [hip, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    proc = subprocess.run(link, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(link), proc.stderr))
    os.replace(tmp, out)
    return out


def build_oracle(verbose: bool = False) -> None:
    """The oracle (test infrastructure) is built by its own Makefile."""
    proc = subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "-j2"], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (proc.stdout, proc.stderr))
    if verbose:
        print(proc.stdout.strip())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    precs = []
    if a.fp32 or not a.fp64:
        precs.append(False)
    if a.fp64 or not a.fp32:
        precs.append(True)
    from . import fortran_include

    for p in fortran_include.write():
        print("wrote", p)
    for d in precs:
        print("built", build_precision(d, a.j, a.v))
    if not a.no_oracle:
        build_oracle(a.v)
        print("built oracle")


if __name__ == "__main__":
    sys.exit(main())
