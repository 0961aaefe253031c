"""Writes the Fortran include files of the two builds, include/fortran/{fp32,fp64}/astaroth.f90.

A Fortran program that calls the binding of include/astaroth_fortran.h does

    use, intrinsic :: iso_c_binding
    include "astaroth.f90"

with the directory of its build on the include path. The file declares the binding's calls
`external`, every enum value of include/user_defines.h as an integer(c_int) parameter (the
tables the reference's DSL generator writes into its astaroth.f90, acc/src/code_generator.c:
604-772), the stream and reduction constants, and the AcMeshInfo derived type with the C
layout of this library: int3 parameters as int3_params(1:3, i) (x, y, z contiguous, as in C)
and real parameters of the build's AcReal kind (c_float or c_double). The reference's type
declares int3_params(N, 3), which interleaves the components differently from its C struct, and
plain `real`, which is 4 bytes also in an FP64 build.

    python -m astaroth_amd.fortran_include      # (re)write both files
"""
from __future__ import annotations

import os
import sys

from . import abi

CALLS = ("achostupdatebuiltinparams", "acgetdevicecount", "acdevicecreate", "acdevicedestroy",
         "acdeviceprintinfo", "acdeviceloadmeshinfo", "acdeviceloadmesh", "acdevicestoremesh",
         "acdeviceintegratesubstep", "acdeviceperiodicboundconds", "acdeviceswapbuffers",
         "acdevicereducescal", "acdevicereducevec", "acdevicesynchronizestream")


def path(double: bool) -> str:
    return os.path.join(abi.INCLUDE, "fortran", "fp64" if double else "fp32", "astaroth.f90")


def text(double: bool) -> str:
    kind = "c_double" if double else "c_float"
    out = ["!  -*-f90-*-",
           "! astaroth_amd Fortran include, %s build (AcReal = real(%s)); written by" % ("FP64" if double else "FP32", kind),
           "! astaroth_amd/fortran_include.py from include/user_defines.h. Calls: include/astaroth_fortran.h",
           "! (every argument by reference). Needs `use, intrinsic :: iso_c_binding` before the include.",
           ""]
    for i in range(0, len(CALLS), 4):
        out.append("external " + ", ".join(CALLS[i:i + 4]))
    out.append("")

    def table(names, count_name):
        for k, name in enumerate(names):
            out.append("integer(c_int), parameter :: %s = %d" % (name, k))
        out.append("integer(c_int), parameter :: %s = %d" % (count_name, len(names)))
        out.append("")

    table(abi.INT_PARAMS, "AC_NUM_INT_PARAMS")
    table(abi.INT3_PARAMS, "AC_NUM_INT3_PARAMS")
    table(abi.REAL_PARAMS, "AC_NUM_REAL_PARAMS")
    table(abi.VTXBUF_NAMES, "AC_NUM_VTXBUF_HANDLES")
    for k in range(abi.NUM_STREAMS):
        out.append("integer(c_int), parameter :: STREAM_%d = %d" % (k, k))
    out.append("integer(c_int), parameter :: NUM_STREAMS = %d" % abi.NUM_STREAMS)
    out.append("integer(c_int), parameter :: STREAM_DEFAULT = STREAM_0")
    out.append("integer(c_int), parameter :: STREAM_ALL = NUM_STREAMS")
    out.append("")
    table(abi.RTYPES, "NUM_REDUCTION_TYPES")
    out += ["integer(c_int), parameter :: AC_BOUNDCOND_PERIODIC = %d" % abi.AC_BOUNDCOND_PERIODIC,
            "integer(c_int), parameter :: AC_BOUNDCOND_SYMMETRIC = %d" % abi.AC_BOUNDCOND_SYMMETRIC,
            "integer(c_int), parameter :: AC_BOUNDCOND_ANTISYMMETRIC = %d" % abi.AC_BOUNDCOND_ANTISYMMETRIC,
            ""]
    out += ["type, bind(C) :: AcMeshInfo",
            "  integer(c_int), dimension(AC_NUM_INT_PARAMS) :: int_params",
            "  integer(c_int), dimension(3, AC_NUM_INT3_PARAMS) :: int3_params",
            "  real(%s), dimension(AC_NUM_REAL_PARAMS) :: real_params" % kind,
            "end type AcMeshInfo", ""]
    return "\n".join(out)


def write() -> list[str]:
    written = []
    for double in (False, True):
        p = path(double)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        new = text(double)
        if not os.path.exists(p) or open(p).read() != new:
            with open(p, "w") as fh:
                fh.write(new)
        written.append(p)
    return written


if __name__ == "__main__":
    for p in write():
        print("wrote", p)
    sys.exit(0)
