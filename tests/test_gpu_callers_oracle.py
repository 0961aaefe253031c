"""Every caller-facing step interface against the oracle on the 64^3 uniform-field box (dt = 1e-3,
5 RK3 steps), as the reference verifies its callers' steps against the host model
(samples/mpitest/main.cc:357-363, samples/ctest/main.c:36-55):

* acNodeIntegrate (the Device step on one GPU; reference node.cc:729-806);
* the legacy acIntegrate (astaroth.cc:28-188);
* acEnsembleIntegrateStepGraph: the box inside a three-box ensemble (persistent batched launch,
  periodic images, one graph per step);
* the Fortran binding's by-reference symbols (astaroth_fortran.cc): acdeviceperiodicboundconds_,
  acdeviceintegratesubstep_ and acdeviceswapbuffers_ three times per step.

Oracle: the CPU restatement of the host model solver in the two-buffer RK3 form of the GPU path
(integration.cuh:14-37), FP64. Bars: FP64 library, acEvalError of the reference
(verification.c:78-102) on every field at the point of maximum absolute error; FP32 library,
range-normalised error <= 1e-5 per field.
"""
import ctypes

import numpy as np
import pytest

import pyoracle
from astaroth_amd import abi
from astaroth_amd.abi import int3
from astaroth_amd.api import HostMesh, Library, real_params_double
from astaroth_amd.initial import mhd_fields

pytestmark = pytest.mark.gpu

N, STEPS, DT = 64, 5, 1e-3
NAMES = abi.VTXBUF_NAMES


@pytest.fixture(scope="module")
def reference():
    pyoracle.set_threads(16)
    f = mhd_fields(N, N, N, np.float64, config_id=1)
    pyoracle.integrate_steps_gpu_form(f, real_params_double(Library(True).config(n=N)), DT, STEPS)
    return [x[3:-3, 3:-3, 3:-3].copy() for x in f]


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


def initial(lib):
    info = lib.config(n=N)
    return info, HostMesh(lib, info, mhd_fields(N, N, N, np.float64, config_id=1))


def check(lib, ref, got, label):
    print("\n%s (%s):" % (label, "fp64" if lib.double else "fp32"), flush=True)
    for q in range(8):
        g = got[q]
        if lib.double:
            v = pyoracle.verify_field(ref[q], g, 53, np.finfo(np.float64).eps)
            print("  %-16s %6.3g ulp %s" % (NAMES[q], v["ulp"], "OK" if v["ok"] else "FAIL"), flush=True)
            assert v["ok"], (label, NAMES[q], v)
        else:
            err = float(np.max(np.abs(g.astype(np.float64) - ref[q]))) / float(np.ptp(ref[q]))
            print("  %-16s range-normalised %.3g" % (NAMES[q], err), flush=True)
            assert err <= 1e-5, (label, NAMES[q], err)


def test_node_integrate_against_oracle(lib, reference):
    info, mesh = initial(lib)
    with lib.node_create(info, autotune=False) as node:
        node.load_mesh(mesh)
        for _ in range(STEPS):
            node.integrate(DT)
        out = node.store_mesh(HostMesh(lib, info))
    check(lib, reference, [out.interior(q) for q in range(8)], "acNodeIntegrate")


def test_legacy_integrate_against_oracle(lib, reference):
    info, mesh = initial(lib)
    d = lib.dll
    d.acIntegrate.argtypes, d.acIntegrate.restype = [lib.real], ctypes.c_int
    d.acStore.argtypes, d.acStore.restype = [ctypes.POINTER(lib.Mesh)], ctypes.c_int
    assert d.acInit(info) == abi.AC_SUCCESS
    try:
        assert d.acLoad(mesh.struct) == abi.AC_SUCCESS
        for _ in range(STEPS):
            assert d.acIntegrate(lib.real(DT)) == abi.AC_SUCCESS
        out = HostMesh(lib, info)
        assert d.acStore(ctypes.byref(out.struct)) == abi.AC_SUCCESS
    finally:
        assert d.acQuit() == abi.AC_SUCCESS
    check(lib, reference, [out.interior(q) for q in range(8)], "acIntegrate")


def test_ensemble_graph_step_against_oracle(lib, reference):
    info, mesh = initial(lib)
    others = [lib.config(n=40, ny=24, nz=56, AC_nu_visc=2e-3), lib.config(n=32, ny=72, nz=16)]
    infos = [others[0], info, others[1]]
    with lib.ensemble_create(infos) as ens:
        for b, inf in enumerate(infos):
            nx, ny, nz = (inf.int_params[abi.IP[k]] for k in ("AC_nx", "AC_ny", "AC_nz"))
            fields = mesh.fields if b == 1 else mhd_fields(nx, ny, nz, np.float64, config_id=6, box_id=b)
            ens.load_mesh(b, HostMesh(lib, inf, fields))
        for _ in range(STEPS):
            ens.integrate_step(DT, mode="graph")
        out = HostMesh(lib, info)
        ens.store_mesh(1, out)
    check(lib, reference, [out.interior(q) for q in range(8)], "acEnsembleIntegrateStepGraph")


def test_fortran_symbols_against_oracle(lib, reference, monkeypatch):
    monkeypatch.setenv("AC_AUTOTUNE", "0")
    info, mesh = initial(lib)
    d = lib.dll
    P, I, R, I3, MI, vp = ctypes.POINTER, ctypes.c_int, lib.real, abi.Int3, lib.MeshInfo, ctypes.c_void_p
    sig = {"acdevicecreate_": [P(I), P(MI), P(vp)], "acdevicedestroy_": [P(vp)],
           "acdeviceloadmesh_": [P(vp), P(I), P(MI), P(I), P(R)],
           "acdevicestoremesh_": [P(vp), P(I), P(MI), P(I), P(R)],
           "acdeviceintegratesubstep_": [P(vp), P(I), P(I), P(I3), P(I3), P(R)],
           "acdeviceperiodicboundconds_": [P(vp), P(I), P(I3), P(I3)], "acdeviceswapbuffers_": [P(vp)],
           "acdevicesynchronizestream_": [P(vp), P(I)]}
    for name, args in sig.items():
        getattr(d, name).argtypes = args
        getattr(d, name).restype = None
    ref = ctypes.byref
    farray = np.ascontiguousarray(np.stack(mesh.fields))  # f(mx, my, mz, 8) in Fortran order
    handle, s, nf = vp(), I(abi.STREAM_DEFAULT), I(8)
    d.acdevicecreate_(ref(I(0)), ref(info), ref(handle))
    d.acdeviceloadmesh_(ref(handle), ref(s), ref(info), ref(nf), farray.ctypes.data_as(P(R)))
    m, lo, hi, dt = I3(N + 6, N + 6, N + 6), I3(3, 3, 3), I3(N + 3, N + 3, N + 3), R(DT)
    for _ in range(STEPS):
        for st in range(3):
            d.acdeviceperiodicboundconds_(ref(handle), ref(s), ref(I3(0, 0, 0)), ref(m))
            d.acdeviceintegratesubstep_(ref(handle), ref(s), ref(I(st)), ref(lo), ref(hi), ref(dt))
            d.acdeviceswapbuffers_(ref(handle))
    out = np.zeros_like(farray)
    d.acdevicestoremesh_(ref(handle), ref(s), ref(info), ref(nf), out.ctypes.data_as(P(R)))
    d.acdevicesynchronizestream_(ref(handle), ref(s))
    d.acdevicedestroy_(ref(handle))
    check(lib, reference, [out[q][3:-3, 3:-3, 3:-3] for q in range(8)], "Fortran symbols")
