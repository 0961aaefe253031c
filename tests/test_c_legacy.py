"""The legacy single-call interface (acInit ... acQuit, reference include/astaroth.h:208-283) and
the Node interface (:355-500) from a C program built against include/ with the system C compiler.

CPU: the program compiles and links against both libraries (FP32 and FP64 headers).
GPU: it runs, and its numbers equal the same sequence driven through the Device interface:
2 x acIntegrate == 2 x acNodeIntegrate == 2 x acDeviceIntegrateStep (bitwise), and acReduce* /
acNodeReduce* == the Device reductions finalised as sqrt(sum / N) for RMS types
(node.cc:932-957).
"""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "c", "legacy_sample.c")


def build_sample(tmp_path, double):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lib = "astaroth_fp64" if double else "astaroth_fp32"
    libdir = os.path.join(REPO, "astaroth_amd", "lib")
    if not os.path.exists(os.path.join(libdir, "lib%s.so" % lib)):
        pytest.fail("library not built: run python -m astaroth_amd.build")
    exe = str(tmp_path / ("legacy_%s" % ("fp64" if double else "fp32")))
    cmd = [cc, "-std=c11", "-O2", "-Wall", "-Werror", "-DAC_DOUBLE_PRECISION=%d" % int(double),
           "-I", os.path.join(REPO, "include"), SRC, "-o", exe, "-L", libdir, "-l" + lib,
           "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib", "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    return exe


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_c_program_builds_against_headers(tmp_path, double):
    build_sample(tmp_path, double)


@pytest.mark.gpu
@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_legacy_interface_matches_device_interface(tmp_path, double):
    from astaroth_amd.api import Library, HostMesh
    from astaroth_amd.abi import IP

    exe = build_sample(tmp_path, double)
    n = 24
    # both sides use the default kernel configuration (no autotuning), so the same kernel runs
    env = dict(os.environ, AC_AUTOTUNE="0")
    out = subprocess.run([exe, str(n)], capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr
    got = [float(v) for v in out.stdout.split()]
    assert len(got) == 5, out.stdout
    # the Node interface (inner region and outer slabs on separate streams) prints the same
    # numbers to rounding (its launches start their z-march columns at other planes)
    node = subprocess.run([exe, str(n), "node"], capture_output=True, text=True, timeout=120,
                          env=env)
    assert node.returncode == 0, node.stderr
    got_node = [float(v) for v in node.stdout.split()]
    assert got_node == pytest.approx(got, rel=1e-5 if not double else 1e-12)

    # the same sequence through the Device interface (same host initial state)
    lib = Library(double)
    info = lib.config(n=n, AC_cs2_sound=1.0)
    mesh = HostMesh(lib, info)
    assert lib.dll.acHostMeshRandomize(ctypes.byref(mesh.struct)) == 0
    for q in range(8):
        f = mesh.fields[q]
        f[...] = 0 if q == 7 else (f - lib.dtype(0.5)) * lib.dtype(2e-3)
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(mesh)
        dev.periodic_boundconds()
        for _ in range(2):
            dev.integrate_step(1e-3)
        dev.periodic_boundconds()
        res = HostMesh(lib, info)
        dev.store_mesh(res)
        nn = float(n ** 3)
        umax = dev.reduce_vec("RTYPE_MAX", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ")
        urms = math.sqrt((1.0 / nn) * dev.reduce_vec("RTYPE_RMS", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ"))
        lrmin = dev.reduce_scal("RTYPE_MIN", "VTXBUF_LNRHO")
        arms = math.sqrt((1.0 / nn) * dev.reduce_vec_scal("RTYPE_ALFVEN_RMS", "VTXBUF_AX", "VTXBUF_AY",
                                                          "VTXBUF_AZ", "VTXBUF_LNRHO"))
    probe = float(res.fields[1][7, 6, 5])
    rtol = 1e-6 if not double else 1e-14
    assert got[0] == pytest.approx(umax, rel=0, abs=0)
    assert got[2] == pytest.approx(lrmin, rel=0, abs=0)
    assert got[1] == pytest.approx(urms, rel=rtol)
    assert got[3] == pytest.approx(arms, rel=rtol)
    assert got[4] == probe  # the stored state is bitwise the Device-interface state
