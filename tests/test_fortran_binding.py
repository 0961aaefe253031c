"""The Fortran binding (reference src/core/astaroth_fortran.{h,cc}): by-reference C symbols and
the generated include files include/fortran/{fp32,fp64}/astaroth.f90.

No Fortran compiler is installed here, so a Fortran program is not compiled: the include files
are checked for agreement with the C tables and layout (every parameter value, the bind(C)
type's size under C layout rules against sizeof(AcMeshInfo)), and the symbols are called the
way a Fortran caller calls them, every argument by reference, from C (compile and link) and
through ctypes (GPU).
"""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from astaroth_amd import abi, fortran_include
from astaroth_amd.api import HostMesh, Library
from astaroth_amd.initial import mhd_fields

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def parameters(text):
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r"integer\(c_int\), parameter :: (\w+) = (\d+)", text)}


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_include_file_is_current_and_matches_tables(double):
    text = open(fortran_include.path(double)).read()
    assert text == fortran_include.text(double), "include file stale: python -m astaroth_amd.fortran_include"
    p = parameters(text)
    for table, count in ((abi.INT_PARAMS, "AC_NUM_INT_PARAMS"), (abi.INT3_PARAMS, "AC_NUM_INT3_PARAMS"),
                         (abi.REAL_PARAMS, "AC_NUM_REAL_PARAMS"), (abi.VTXBUF_NAMES, "AC_NUM_VTXBUF_HANDLES"),
                         (abi.RTYPES, "NUM_REDUCTION_TYPES")):
        assert p[count] == len(table)
        for k, name in enumerate(table):
            assert p[name] == k, name
    assert p["NUM_STREAMS"] == abi.NUM_STREAMS and p["STREAM_31"] == 31
    for name in fortran_include.CALLS:
        assert re.search(r"^external .*\b%s\b" % name, text, re.M), name
    # the derived type under C layout rules has the size of the C struct
    m = re.search(r"type, bind\(C\) :: AcMeshInfo(.*?)end type", text, re.S)
    off, align = 0, 1
    for kind, dims in re.findall(r"(integer\(c_int\)|real\(c_float\)|real\(c_double\)), dimension\(([^)]*)\)",
                                 m.group(1)):
        size = 8 if kind == "real(c_double)" else 4
        n = 1
        for d in dims.split(","):
            d = d.strip()
            n *= int(d) if d.isdigit() else p[d]
        off = (off + size - 1) // size * size + n * size
        align = max(align, size)
    off = (off + align - 1) // align * align
    MI, _ = abi.structs(double)
    assert off == ctypes.sizeof(MI) == (728 if double else 424)
    assert "dimension(3, AC_NUM_INT3_PARAMS)" in m.group(1)  # x, y, z contiguous as in C


C_SAMPLE = r"""
#include <stdio.h>
#include "astaroth_fortran.h"
/* the calls a Fortran program makes, with every argument by reference */
int main(void) {
    int count = 0, id = 0, nf = NUM_VTXBUF_HANDLES, step = 0;
    Stream s = STREAM_DEFAULT;
    ReductionType rt = RTYPE_MAX;
    VertexBufferHandle h = VTXBUF_LNRHO;
    AcMeshInfo info = {0};
    Device dev = NULL;
    AcReal dt = 0, r = 0;
    int3 a = {3, 3, 3}, b = {8, 8, 8};
    static AcReal f[8 * 14 * 14 * 14];
    (void)f; (void)count;
    if (0) {
        achostupdatebuiltinparams_(&info); acgetdevicecount_(&count);
        acdevicecreate_(&id, &info, &dev); acdeviceprintinfo_(&dev); acdeviceloadmeshinfo_(&dev, &info);
        acdeviceloadmesh_(&dev, &s, &info, &nf, f); acdevicestoremesh_(&dev, &s, &info, &nf, f);
        acdeviceintegratesubstep_(&dev, &s, &step, &a, &b, &dt); acdeviceperiodicboundconds_(&dev, &s, &a, &b);
        acdeviceswapbuffers_(&dev); acdevicereducescal_(&dev, &s, &rt, &h, &r);
        acdevicereducevec_(&dev, &s, &rt, &h, &h, &h, &r); acdevicesynchronizestream_(&dev, &s);
        acdevicedestroy_(&dev);
    }
    printf("ok\n");
    return 0;
}
"""


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_c_caller_compiles_and_links(tmp_path, double):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "f.c"
    src.write_text(C_SAMPLE)
    libdir = os.path.join(REPO, "astaroth_amd", "lib")
    lib = "astaroth_fp64" if double else "astaroth_fp32"
    exe = str(tmp_path / "f")
    cmd = [cc, "-std=c11", "-Wall", "-Werror", "-DAC_DOUBLE_PRECISION=%d" % int(double),
           "-I", os.path.join(REPO, "include"), str(src), "-o", exe, "-L", libdir, "-l" + lib,
           "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib", "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    assert subprocess.run([exe], capture_output=True, text=True, timeout=60).stdout == "ok\n"


@pytest.mark.gpu
@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_binding_calls_match_device_interface(double, monkeypatch):
    # both sides run the library's default kernel configuration (no autotuning at creation)
    monkeypatch.setenv("AC_AUTOTUNE", "0")
    lib = Library(double)
    d = lib.dll
    P = ctypes.POINTER
    I, R, I3, MI, vp = ctypes.c_int, lib.real, abi.Int3, lib.MeshInfo, ctypes.c_void_p
    sig = {"achostupdatebuiltinparams_": [P(MI)], "acgetdevicecount_": [P(I)],
           "acdevicecreate_": [P(I), P(MI), P(vp)], "acdevicedestroy_": [P(vp)],
           "acdeviceloadmesh_": [P(vp), P(I), P(MI), P(I), P(R)],
           "acdevicestoremesh_": [P(vp), P(I), P(MI), P(I), P(R)],
           "acdeviceintegratesubstep_": [P(vp), P(I), P(I), P(I3), P(I3), P(R)],
           "acdeviceperiodicboundconds_": [P(vp), P(I), P(I3), P(I3)], "acdeviceswapbuffers_": [P(vp)],
           "acdevicereducescal_": [P(vp), P(I), P(I), P(I), P(R)],
           "acdevicereducevec_": [P(vp), P(I), P(I), P(I), P(I), P(I), P(R)],
           "acdevicesynchronizestream_": [P(vp), P(I)]}
    for name, args in sig.items():
        getattr(d, name).argtypes = args
        getattr(d, name).restype = None
    ref = ctypes.byref
    count = I(0)
    d.acgetdevicecount_(ref(count))
    assert count.value >= 1
    n = (20, 14, 10)  # not cubic: each field holds mx * my * mz values
    info = lib.config(n=n[0], ny=n[1], nz=n[2])
    d.achostupdatebuiltinparams_(ref(info))
    fields = mhd_fields(*n, np.float64, config_id=1)
    # a Fortran array f(mx, my, mz, 8): x fastest, field slowest
    farray = np.ascontiguousarray(np.stack(fields).astype(lib.dtype))
    handle = vp()
    d.acdevicecreate_(ref(I(0)), ref(info), ref(handle))
    s = I(abi.STREAM_DEFAULT)
    nf = I(8)
    d.acdeviceloadmesh_(ref(handle), ref(s), ref(info), ref(nf), farray.ctypes.data_as(P(R)))
    m = I3(n[0] + 6, n[1] + 6, n[2] + 6)
    lo, hi = I3(3, 3, 3), I3(n[0] + 3, n[1] + 3, n[2] + 3)
    dt = R(1e-3)
    for st in range(3):
        d.acdeviceperiodicboundconds_(ref(handle), ref(s), ref(I3(0, 0, 0)), ref(m))
        d.acdeviceintegratesubstep_(ref(handle), ref(s), ref(I(st)), ref(lo), ref(hi), ref(dt))
        d.acdeviceswapbuffers_(ref(handle))
    d.acdeviceperiodicboundconds_(ref(handle), ref(s), ref(I3(0, 0, 0)), ref(m))
    out = np.zeros_like(farray)
    d.acdevicestoremesh_(ref(handle), ref(s), ref(info), ref(nf), out.ctypes.data_as(P(R)))
    d.acdevicesynchronizestream_(ref(handle), ref(s))
    rmax, rvec = R(0), R(0)
    d.acdevicereducescal_(ref(handle), ref(s), ref(I(abi.RT["RTYPE_MAX"])), ref(I(abi.VTX["VTXBUF_LNRHO"])),
                          ref(rmax))
    d.acdevicereducevec_(ref(handle), ref(s), ref(I(abi.RT["RTYPE_RMS"])), ref(I(1)), ref(I(2)), ref(I(3)),
                         ref(rvec))
    d.acdevicedestroy_(ref(handle))
    assert not handle.value
    # the same through the Device interface
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(HostMesh(lib, info, fields))
        dev.integrate_step(1e-3)
        dev.periodic_boundconds()
        res = HostMesh(lib, info)
        dev.store_mesh(res)
        assert rmax.value == dev.reduce_scal("RTYPE_MAX", "VTXBUF_LNRHO")
        assert rvec.value == dev.reduce_vec("RTYPE_RMS", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ")
    for q in range(8):
        assert np.array_equal(out[q], res.fields[q])
