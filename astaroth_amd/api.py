"""Python mirror of the Device interface of include/astaroth.h.

Loads astaroth_amd/lib/libastaroth_{fp32,fp64}.so with ctypes and wraps the C entry points with
the reference's names and argument meaning (src/core/device.cc, include/astaroth.h:501-642), so a
test reads like the reference's own samples:

    lib  = Library(double=True)
    info = lib.config(n=64)                    # AcMeshInfo of a 64^3 box
    mesh = HostMesh(lib, info)                 # 8 numpy arrays of m^3 cells
    dev  = lib.device_create(0, info)          # acDeviceCreate
    dev.load_mesh(mesh); dev.integrate_step(dt); dev.store_mesh(mesh)

The library is required: a missing or unloadable .so raises, there is no fallback path.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import abi
from .abi import IP, RP, VTX, RT, int3

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")


class AstarothError(RuntimeError):
    pass


def lib_path(double: bool) -> str:
    return os.path.join(LIBDIR, "libastaroth_%s.so" % ("fp64" if double else "fp32"))


_LIBS: dict[bool, "Library"] = {}


class Library:
    """One precision build of the C library."""

    def __new__(cls, double: bool = False):
        if double in _LIBS:
            return _LIBS[double]
        self = super().__new__(cls)
        _LIBS[double] = self
        self._init(double)
        return self

    def _init(self, double: bool):
        path = lib_path(double)
        if not os.path.exists(path):
            raise AstarothError(
                "%s is not built; run `python -m astaroth_amd.build` (hipcc, gfx950)" % path)
        self.path = path
        self.double = double
        self.dll = ctypes.CDLL(path)
        self.real = abi.real_ctype(double)
        self.dtype = abi.real_dtype(double)
        self.MeshInfo, self.Mesh = abi.structs(double)
        self.SimulationResult = abi.simulation_result(double)
        self._declare()
        size_mi = ctypes.c_size_t()
        size_m = ctypes.c_size_t()
        rsize = self.dll.acAbiRealSize(ctypes.byref(size_mi), ctypes.byref(size_m))
        if rsize != ctypes.sizeof(self.real) or size_mi.value != ctypes.sizeof(self.MeshInfo) \
                or size_m.value != ctypes.sizeof(self.Mesh):
            raise AstarothError("ABI mismatch between %s and astaroth_amd.abi" % path)

    # ---- ctypes declarations -------------------------------------------------------------
    def _declare(self):
        d = self.dll
        R = self.real
        MI = self.MeshInfo
        M = self.Mesh
        vp = ctypes.c_void_p
        i = ctypes.c_int
        I3 = abi.Int3
        decl = {
            "acAbiRealSize": (i, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
            "acDeviceCreate": (i, [i, MI, ctypes.POINTER(vp)]),
            "acDeviceDestroy": (i, [vp]),
            "acDevicePrintInfo": (i, [vp]),
            "acDeviceAutoOptimize": (i, [vp]),
            "acDeviceSynchronizeStream": (i, [vp, i]),
            "acDeviceSwapBuffers": (i, [vp]),
            "acDeviceLoadScalarUniform": (i, [vp, i, i, R]),
            "acDeviceLoadIntUniform": (i, [vp, i, i, i]),
            "acDeviceLoadInt3Uniform": (i, [vp, i, i, I3]),
            "acDeviceLoadMeshInfo": (i, [vp, MI]),
            "acDeviceLoadDefaultUniforms": (i, [vp]),
            "acDeviceLoadVertexBufferWithOffset": (i, [vp, i, M, i, I3, I3, i]),
            "acDeviceLoadMeshWithOffset": (i, [vp, i, M, I3, I3, i]),
            "acDeviceLoadVertexBuffer": (i, [vp, i, M, i]),
            "acDeviceLoadMesh": (i, [vp, i, M]),
            "acDeviceSetVertexBuffer": (i, [vp, i, i, R]),
            "acDeviceStoreVertexBufferWithOffset": (i, [vp, i, i, I3, I3, i, ctypes.POINTER(M)]),
            "acDeviceStoreMeshWithOffset": (i, [vp, i, I3, I3, i, ctypes.POINTER(M)]),
            "acDeviceStoreVertexBuffer": (i, [vp, i, i, ctypes.POINTER(M)]),
            "acDeviceStoreMesh": (i, [vp, i, ctypes.POINTER(M)]),
            "acDeviceTransferVertexBuffer": (i, [vp, i, i, vp]),
            "acDeviceTransferMesh": (i, [vp, i, vp]),
            "acDeviceIntegrateSubstep": (i, [vp, i, i, I3, I3, R]),
            "acDevicePeriodicBoundcondStep": (i, [vp, i, i, I3, I3]),
            "acDevicePeriodicBoundconds": (i, [vp, i, I3, I3]),
            "acDeviceGeneralBoundconds": (i, [vp, i, I3, I3, MI, I3]),
            "acDeviceGeneralBoundcondStep": (i, [vp, i, i, I3, I3, MI, I3]),
            "acDeviceReduceScal": (i, [vp, i, i, i, ctypes.POINTER(R)]),
            "acDeviceReduceVec": (i, [vp, i, i, i, i, i, ctypes.POINTER(R)]),
            "acDeviceReduceVecScal": (i, [vp, i, i, i, i, i, i, ctypes.POINTER(R)]),
            "acDeviceReduceScalDouble": (i, [vp, i, i, i, ctypes.POINTER(ctypes.c_double)]),
            "acDeviceReduceVecDouble": (i, [vp, i, i, i, i, i, ctypes.POINTER(ctypes.c_double)]),
            "acDeviceReduceVecScalDouble": (i, [vp, i, i, i, i, i, i, ctypes.POINTER(ctypes.c_double)]),
            "acDeviceReduceNoAlloc": (i, [vp, i, i, i, ctypes.POINTER(i), vp]),
            "acDeviceRunMPITest": (i, []),
            "acNodeCreate": (i, [i, MI, ctypes.POINTER(vp)]),
            "acNodeDestroy": (i, [vp]),
            "acNodePrintInfo": (i, [vp]),
            "acNodeQueryDeviceConfiguration": (i, [vp, ctypes.POINTER(abi.DeviceConfiguration)]),
            "acNodeAutoOptimize": (i, [vp]),
            "acNodeSynchronizeStream": (i, [vp, i]),
            "acNodeSynchronizeVertexBuffer": (i, [vp, i, i]),
            "acNodeSynchronizeMesh": (i, [vp, i]),
            "acNodeSwapBuffers": (i, [vp]),
            "acNodeLoadConstant": (i, [vp, i, i, R]),
            "acNodeLoadVertexBufferWithOffset": (i, [vp, i, M, i, I3, I3, i]),
            "acNodeLoadMeshWithOffset": (i, [vp, i, M, I3, I3, i]),
            "acNodeLoadVertexBuffer": (i, [vp, i, M, i]),
            "acNodeLoadMesh": (i, [vp, i, M]),
            "acNodeSetVertexBuffer": (i, [vp, i, i, R]),
            "acNodeStoreVertexBufferWithOffset": (i, [vp, i, i, I3, I3, i, ctypes.POINTER(M)]),
            "acNodeStoreMeshWithOffset": (i, [vp, i, I3, I3, i, ctypes.POINTER(M)]),
            "acNodeStoreVertexBuffer": (i, [vp, i, i, ctypes.POINTER(M)]),
            "acNodeStoreMesh": (i, [vp, i, ctypes.POINTER(M)]),
            "acNodeIntegrateSubstep": (i, [vp, i, i, I3, I3, R]),
            "acNodeIntegrate": (i, [vp, R]),
            "acNodeIntegrateGBC": (i, [vp, MI, R]),
            "acNodePeriodicBoundcondStep": (i, [vp, i, i]),
            "acNodePeriodicBoundconds": (i, [vp, i]),
            "acNodeGeneralBoundcondStep": (i, [vp, i, i, MI]),
            "acNodeGeneralBoundconds": (i, [vp, i, MI]),
            "acNodeReduceScal": (i, [vp, i, i, i, ctypes.POINTER(R)]),
            "acNodeReduceVec": (i, [vp, i, i, i, i, i, ctypes.POINTER(R)]),
            "acNodeReduceVecScal": (i, [vp, i, i, i, i, i, i, ctypes.POINTER(R)]),
            "acGetNode": (vp, []),
            "acInit": (i, [MI]),
            "acQuit": (i, []),
            "acLoad": (i, [M]),
            "acReduceScal": (R, [i, i]),
            "acDeviceIntegrateStep": (i, [vp, i, R]),
            "acDeviceIntegrateSubstepPeriodic": (i, [vp, i, i, R]),
            "acDeviceSubstepReadsPeriodicImages": (i, [vp]),
            "acDeviceSetReadPeriodicImages": (i, [vp, i]),
            "acDeviceIntegrateStepGraph": (i, [vp, i, R]),
            "acKernelConfigCount": (i, []),
            "acKernelConfigGet": (i, [i, ctypes.POINTER(abi.KernelConfig)]),
            "acDeviceGetKernelConfig": (i, [vp, ctypes.POINTER(abi.KernelConfig)]),
            "acDeviceSetKernelConfig": (i, [vp, abi.KernelConfig]),
            "acDeviceTimeKernelConfig": (ctypes.c_float, [vp, abi.KernelConfig, i]),
            "acDeviceRandomizeUniform": (i, [vp, i, i, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong]),
            "acDeviceGetStreamHandle": (vp, [vp, i]),
            "acDeviceCopyVertexBuffer": (i, [vp, i, i]),
            "acDeviceRandomizeUniformSinZ": (i, [vp, i, i, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double]),
            "acEnsembleCreate": (i, [i, i, ctypes.POINTER(MI), ctypes.POINTER(vp)]),
            "acEnsembleDestroy": (i, [vp]),
            "acEnsembleBoxCount": (i, [vp]),
            "acEnsembleSetKernelConfig": (i, [vp, abi.KernelConfig]),
            "acEnsembleGetKernelConfig": (i, [vp, ctypes.POINTER(abi.KernelConfig)]),
            "acEnsembleAutoOptimize": (i, [vp]),
            "acEnsembleBoxZDepth": (i, [vp, i]),
            "acEnsembleRandomizeUniform": (i, [vp, i, i, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong]),
            "acEnsembleSetVertexBuffer": (i, [vp, i, i, R]),
            "acEnsembleLoadMesh": (i, [vp, i, M]),
            "acEnsembleStoreMesh": (i, [vp, i, ctypes.POINTER(M)]),
            "acEnsemblePeriodicBoundconds": (i, [vp]),
            "acEnsembleIntegrateStep": (i, [vp, R]),
            "acEnsembleIntegrateStepGraph": (i, [vp, R]),
            "acEnsembleIntegrateStepLooped": (i, [vp, R]),
            "acEnsembleDiagnostics": (i, [vp, ctypes.POINTER(ctypes.c_double)]),
            "acEnsembleSynchronize": (i, [vp]),
            "acEnsembleGetStreamHandle": (vp, [vp]),
            "acDeviceSetCompensatedIntegration": (i, [vp, i]),
            "acDeviceTimestepCFL": (i, [vp, i, ctypes.POINTER(ctypes.c_double)]),
            "acEnsembleTimestepCFL": (i, [vp, ctypes.POINTER(ctypes.c_double)]),
            "acDevicePackData": (i, [vp, i, I3, I3, vp]),
            "acDeviceUnpackData": (i, [vp, i, vp, I3, I3]),
            "acDeviceHaloPackedSize": (ctypes.c_size_t, [vp]),
            "acDevicePackHalo": (i, [vp, i, vp]),
            "acDeviceUnpackHalo": (i, [vp, i, vp]),
            "acDeviceHaloExchangeSelf": (i, [vp, i]),
            "acHostUpdateBuiltinParams": (i, [ctypes.POINTER(MI)]),
            "acHostMeshCreate": (i, [MI, ctypes.POINTER(M)]),
            "acHostMeshRandomize": (i, [ctypes.POINTER(M)]),
            "acHostMeshDestroy": (i, [ctypes.POINTER(M)]),
            "acLoadConfig": (i, [ctypes.c_char_p, ctypes.POINTER(MI)]),
            "acDeviceGetMeshInfo": (i, [vp, ctypes.POINTER(MI)]),
            "acDeviceDiagnostics": (i, [vp, i, ctypes.POINTER(ctypes.c_double)]),
            "acDeviceIntegrateStepGBC": (i, [vp, i, MI, R]),
            "acHostWriteMeshInfo": (i, [MI, ctypes.c_char_p]),
            "acHostSaveMeshSnapshot": (i, [M, i, R, ctypes.c_char_p]),
            "acHostReadMeshSnapshot": (i, [M, i, ctypes.POINTER(R), ctypes.c_char_p]),
            "acTimeseriesBegin": (i, [ctypes.c_char_p, i]),
            "acDeviceTimeseriesAppend": (i, [vp, i, i, R, R, ctypes.c_char_p, ctypes.POINTER(i)]),
            "acDeviceSaveMeshSnapshot": (i, [vp, i, i, R, ctypes.c_char_p]),
            "acDeviceLoadMeshSnapshot": (i, [vp, i, i, ctypes.POINTER(R), ctypes.c_char_p]),
            "acDeviceRunSimulation": (i, [vp, ctypes.c_char_p, ctypes.POINTER(self.SimulationResult)]),
        }
        self.symbols = sorted(decl)
        for name, (res, args) in decl.items():
            fn = getattr(d, name)
            fn.restype = res
            fn.argtypes = args

    # ---- configuration helpers -----------------------------------------------------------
    def update_builtin_params(self, info):
        self.dll.acHostUpdateBuiltinParams(ctypes.byref(info))
        return info

    def load_config(self, path: str):
        info = self.MeshInfo()
        res = self.dll.acLoadConfig(path.encode(), ctypes.byref(info))
        return info, res

    def config(self, n=64, ny=None, nz=None, **overrides):
        """AcMeshInfo of the 8-field MHD configuration used by the tests and the benchmark.

        Grid spacing 2 pi / n, nu = eta = 5e-3, cs = 1, gamma = 5/3, zeta = 0.01, mu0 = 1.4,
        cp = 1, lnT0 = 1.2, lnrho0 = 1.3 (remaining values from the reference's
        config/astaroth.conf:42-67). `overrides` set any real parameter by name.
        """
        info = self.MeshInfo()
        ctypes.memset(ctypes.byref(info), 0, ctypes.sizeof(info))
        nx = int(n)
        ny = int(ny if ny is not None else n)
        nz = int(nz if nz is not None else n)
        info.int_params[IP["AC_nx"]] = nx
        info.int_params[IP["AC_ny"]] = ny
        info.int_params[IP["AC_nz"]] = nz
        vals = dict(
            AC_dsx=2 * math.pi / nx, AC_dsy=2 * math.pi / ny, AC_dsz=2 * math.pi / nz,
            AC_nu_visc=5e-3, AC_eta=5e-3, AC_cs_sound=1.0, AC_gamma=5.0 / 3.0, AC_zeta=0.01,
            AC_mu0=1.4, AC_cp_sound=1.0, AC_lnT0=1.2, AC_lnrho0=1.3, AC_cdt=0.4, AC_cdtv=0.3,
            AC_cdts=1.0,
        )
        vals.update(overrides)
        for k, v in vals.items():
            info.real_params[RP[k]] = v
        info.real_params[RP["AC_cs2_sound"]] = info.real_params[RP["AC_cs_sound"]] ** 2
        self.update_builtin_params(info)
        return info

    # ---- snapshot files (host side; no GPU needed) ----------------------------------------
    def write_mesh_info(self, info, dirpath: str):
        if self.dll.acHostWriteMeshInfo(info, dirpath.encode()) != abi.AC_SUCCESS:
            raise AstarothError("acHostWriteMeshInfo(%s) failed" % dirpath)

    def save_mesh_snapshot(self, mesh: "HostMesh", step: int, t: float, dirpath: str):
        if self.dll.acHostSaveMeshSnapshot(mesh.struct, int(step), self.real(t), dirpath.encode()) != abi.AC_SUCCESS:
            raise AstarothError("acHostSaveMeshSnapshot(step %d, %s) failed" % (step, dirpath))

    def read_mesh_snapshot(self, mesh: "HostMesh", step: int, dirpath: str) -> float:
        t = self.real()
        if self.dll.acHostReadMeshSnapshot(mesh.struct, int(step), ctypes.byref(t), dirpath.encode()) != abi.AC_SUCCESS:
            raise AstarothError("acHostReadMeshSnapshot(step %d, %s) failed" % (step, dirpath))
        return t.value

    def timeseries_begin(self, path: str, start_step: int = 0):
        if self.dll.acTimeseriesBegin(path.encode(), int(start_step)) != abi.AC_SUCCESS:
            raise AstarothError("acTimeseriesBegin(%s) failed" % path)

    def device_create(self, device_id: int, info, autotune: bool | None = None) -> "Device":
        return Device(self, device_id, info, autotune)

    def node_create(self, info, node_id: int = 0, autotune: bool | None = None) -> "Node":
        return Node(self, info, node_id, autotune)

    def ensemble_create(self, infos, device_id: int = 0, autotune: bool | None = None) -> "Ensemble":
        return Ensemble(self, infos, device_id, autotune)

    def kernel_configs(self):
        out = []
        for k in range(self.dll.acKernelConfigCount()):
            c = abi.KernelConfig()
            self.dll.acKernelConfigGet(k, ctypes.byref(c))
            out.append(c)
        return out


def dims_of(info):
    return tuple(int(info.int_params[IP[k]]) for k in ("AC_nx", "AC_ny", "AC_nz"))


def real_params_double(info) -> np.ndarray:
    return np.array([float(info.real_params[k]) for k in range(abi.NUM_REAL_PARAMS)], dtype=np.float64)


class HostMesh:
    """AcMesh whose eight buffers are numpy arrays of shape (mz, my, mx) (x fastest)."""

    def __init__(self, lib: Library, info, fields=None):
        self.lib = lib
        self.info = info
        nx, ny, nz = dims_of(info)
        self.shape = (nz + 6, ny + 6, nx + 6)
        if fields is None:
            self.fields = [np.zeros(self.shape, dtype=lib.dtype) for _ in range(abi.NUM_VTXBUF_HANDLES)]
        else:
            self.fields = [np.ascontiguousarray(f, dtype=lib.dtype) for f in fields]
            for f in self.fields:
                if f.shape != self.shape:
                    raise ValueError("field shape %s != %s" % (f.shape, self.shape))
        self.struct = lib.Mesh()
        self.struct.info = info
        self._sync_pointers()

    def _sync_pointers(self):
        P = ctypes.POINTER(self.lib.real)
        for q, f in enumerate(self.fields):
            self.struct.vertex_buffer[q] = f.ctypes.data_as(P)

    def copy(self) -> "HostMesh":
        return HostMesh(self.lib, self.info, [f.copy() for f in self.fields])

    def interior(self, q):
        return self.fields[q][3:-3, 3:-3, 3:-3]

    def __getitem__(self, q):
        return self.fields[q]


class Device:
    """Handle returned by acDeviceCreate (include/astaroth.h:507)."""

    def __init__(self, lib: Library, device_id: int, info, autotune=None):
        self.lib = lib
        self.info = info
        self.dims = dims_of(info)
        self.handle = ctypes.c_void_p()
        old = os.environ.get("AC_AUTOTUNE")
        if autotune is not None:
            os.environ["AC_AUTOTUNE"] = "1" if autotune else "0"
        try:
            res = lib.dll.acDeviceCreate(int(device_id), info, ctypes.byref(self.handle))
        finally:
            if autotune is not None:
                if old is None:
                    del os.environ["AC_AUTOTUNE"]
                else:
                    os.environ["AC_AUTOTUNE"] = old
        if res != abi.AC_SUCCESS:
            raise AstarothError("acDeviceCreate failed")

    @classmethod
    def view(cls, lib: Library, info, handle) -> "Device":
        """A Device object over an existing handle (a Node's Device, say), which it does not own:
        do not destroy it."""
        self = cls.__new__(cls)
        self.lib, self.info, self.dims = lib, info, dims_of(info)
        self.handle = ctypes.c_void_p(handle)
        return self

    # context manager
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def _check(self, res, what):
        if res != abi.AC_SUCCESS:
            raise AstarothError("%s returned AC_FAILURE" % what)
        return res

    def destroy(self):
        if self.handle:
            self.lib.dll.acDeviceDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def synchronize(self, stream=abi.STREAM_ALL):
        return self._check(self.lib.dll.acDeviceSynchronizeStream(self.handle, stream), "acDeviceSynchronizeStream")

    def swap_buffers(self):
        return self.lib.dll.acDeviceSwapBuffers(self.handle)

    def load_mesh(self, mesh: HostMesh, stream=abi.STREAM_DEFAULT):
        self._check(self.lib.dll.acDeviceLoadMesh(self.handle, stream, mesh.struct), "acDeviceLoadMesh")
        self.synchronize(stream)

    def store_mesh(self, mesh: HostMesh, stream=abi.STREAM_DEFAULT):
        self._check(self.lib.dll.acDeviceStoreMesh(self.handle, stream, ctypes.byref(mesh.struct)), "acDeviceStoreMesh")
        self.synchronize(stream)
        return mesh

    def load_uniform(self, name: str, value, stream=abi.STREAM_DEFAULT):
        return self.lib.dll.acDeviceLoadScalarUniform(self.handle, stream, RP[name], value)

    def integrate_substep(self, step, start=None, end=None, dt=1e-3, stream=abi.STREAM_DEFAULT):
        nx, ny, nz = self.dims
        start = start or (3, 3, 3)
        end = end or (nx + 3, ny + 3, nz + 3)
        return self.lib.dll.acDeviceIntegrateSubstep(self.handle, stream, step, int3(*start), int3(*end), dt)

    def periodic_boundconds(self, start=None, end=None, stream=abi.STREAM_DEFAULT):
        nx, ny, nz = self.dims
        start = start or (0, 0, 0)
        end = end or (nx + 6, ny + 6, nz + 6)
        return self._check(self.lib.dll.acDevicePeriodicBoundconds(self.handle, stream, int3(*start), int3(*end)),
                           "acDevicePeriodicBoundconds")

    def periodic_boundcond_step(self, handle, start=None, end=None, stream=abi.STREAM_DEFAULT):
        nx, ny, nz = self.dims
        start = start or (0, 0, 0)
        end = end or (nx + 6, ny + 6, nz + 6)
        return self.lib.dll.acDevicePeriodicBoundcondStep(self.handle, stream, handle, int3(*start), int3(*end))

    def general_boundconds(self, config, start=None, end=None, bindex=(-1, -1, -1), handle=None,
                           stream=abi.STREAM_DEFAULT):
        """acDeviceGeneralBoundconds (all fields) or acDeviceGeneralBoundcondStep (one field) with
        the face types of `config` (AC_bc_type_{bot,top}_{x,y,z})."""
        nx, ny, nz = self.dims
        start = int3(*(start or (0, 0, 0)))
        end = int3(*(end or (nx + 6, ny + 6, nz + 6)))
        if handle is None:
            r = self.lib.dll.acDeviceGeneralBoundconds(self.handle, stream, start, end, config, int3(*bindex))
        else:
            r = self.lib.dll.acDeviceGeneralBoundcondStep(self.handle, stream, _vh(handle), start, end, config,
                                                          int3(*bindex))
        return self._check(r, "acDeviceGeneralBoundconds")

    def integrate_step_gbc(self, config, dt, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceIntegrateStepGBC(self.handle, stream, config, dt),
                           "acDeviceIntegrateStepGBC")

    def integrate_step(self, dt, stream=abi.STREAM_DEFAULT, graph=False):
        fn = self.lib.dll.acDeviceIntegrateStepGraph if graph else self.lib.dll.acDeviceIntegrateStep
        return self._check(fn(self.handle, stream, dt), "acDeviceIntegrateStep")

    def reduce_scal(self, rtype, handle, stream=abi.STREAM_DEFAULT, double=False):
        if double:
            out = ctypes.c_double()
            res = self.lib.dll.acDeviceReduceScalDouble(self.handle, stream, _rt(rtype), _vh(handle), ctypes.byref(out))
        else:
            out = self.lib.real()
            res = self.lib.dll.acDeviceReduceScal(self.handle, stream, _rt(rtype), _vh(handle), ctypes.byref(out))
        self._check(res, "acDeviceReduceScal")
        return out.value

    def reduce_vec(self, rtype, h0, h1, h2, stream=abi.STREAM_DEFAULT, double=False):
        args = (self.handle, stream, _rt(rtype), _vh(h0), _vh(h1), _vh(h2))
        if double:
            out = ctypes.c_double()
            res = self.lib.dll.acDeviceReduceVecDouble(*args, ctypes.byref(out))
        else:
            out = self.lib.real()
            res = self.lib.dll.acDeviceReduceVec(*args, ctypes.byref(out))
        self._check(res, "acDeviceReduceVec")
        return out.value

    def reduce_vec_scal(self, rtype, h0, h1, h2, h3, stream=abi.STREAM_DEFAULT, double=False):
        args = (self.handle, stream, _rt(rtype), _vh(h0), _vh(h1), _vh(h2), _vh(h3))
        if double:
            out = ctypes.c_double()
            res = self.lib.dll.acDeviceReduceVecScalDouble(*args, ctypes.byref(out))
        else:
            out = self.lib.real()
            res = self.lib.dll.acDeviceReduceVecScal(*args, ctypes.byref(out))
        self._check(res, "acDeviceReduceVecScal")
        return out.value

    def set_vertex_buffer(self, handle, value, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceSetVertexBuffer(self.handle, stream, _vh(handle), value),
                           "acDeviceSetVertexBuffer")

    def randomize_uniform(self, handle, lo, hi, seed, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceRandomizeUniform(self.handle, stream, _vh(handle), lo, hi, seed),
                           "acDeviceRandomizeUniform")

    def timestep_cfl(self, stream=abi.STREAM_DEFAULT) -> float:
        out = ctypes.c_double()
        self._check(self.lib.dll.acDeviceTimestepCFL(self.handle, stream, ctypes.byref(out)), "acDeviceTimestepCFL")
        return out.value

    def diagnostics(self, stream=abi.STREAM_DEFAULT) -> dict:
        """acDeviceDiagnostics: {'uu': [max, min, rms], 'aa': [...], 'lnrho': [...]} (FP64)."""
        out = (ctypes.c_double * 9)()
        self._check(self.lib.dll.acDeviceDiagnostics(self.handle, stream, out), "acDeviceDiagnostics")
        v = list(out)
        return {"uu": v[0:3], "aa": v[3:6], "lnrho": v[6:9]}

    def mesh_info(self):
        info = self.lib.MeshInfo()
        self._check(self.lib.dll.acDeviceGetMeshInfo(self.handle, ctypes.byref(info)), "acDeviceGetMeshInfo")
        return info

    def timeseries_append(self, step: int, dt: float, t: float, path: str, stream=abi.STREAM_DEFAULT) -> bool:
        """Appends the diagnostics row of `step`; returns True when a statistic was NaN."""
        nan = ctypes.c_int(0)
        self._check(self.lib.dll.acDeviceTimeseriesAppend(self.handle, stream, int(step), self.lib.real(dt),
                                                          self.lib.real(t), path.encode(), ctypes.byref(nan)),
                    "acDeviceTimeseriesAppend")
        return bool(nan.value)

    def save_mesh_snapshot(self, step: int, t: float, dirpath: str, stream=abi.STREAM_DEFAULT):
        self._check(self.lib.dll.acDeviceSaveMeshSnapshot(self.handle, stream, int(step), self.lib.real(t),
                                                          dirpath.encode()), "acDeviceSaveMeshSnapshot")

    def load_mesh_snapshot(self, step: int, dirpath: str, stream=abi.STREAM_DEFAULT) -> float:
        t = self.lib.real()
        self._check(self.lib.dll.acDeviceLoadMeshSnapshot(self.handle, stream, int(step), ctypes.byref(t),
                                                          dirpath.encode()), "acDeviceLoadMeshSnapshot")
        return t.value

    def run_simulation(self, dirpath: str):
        """acDeviceRunSimulation: returns (last step, t, stop reason)."""
        r = self.lib.SimulationResult()
        self._check(self.lib.dll.acDeviceRunSimulation(self.handle, dirpath.encode(), ctypes.byref(r)),
                    "acDeviceRunSimulation")
        return r.last_step, r.t, r.stop

    def set_compensated_integration(self, enable=True):
        return self._check(self.lib.dll.acDeviceSetCompensatedIntegration(self.handle, int(bool(enable))),
                           "acDeviceSetCompensatedIntegration")

    def randomize_uniform_sinz(self, handle, lo, hi, seed, amplitude, phase_per_cell, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceRandomizeUniformSinZ(self.handle, stream, _vh(handle), lo, hi,
                                                                     seed, amplitude, phase_per_cell),
                           "acDeviceRandomizeUniformSinZ")

    def kernel_config(self) -> abi.KernelConfig:
        c = abi.KernelConfig()
        self.lib.dll.acDeviceGetKernelConfig(self.handle, ctypes.byref(c))
        return c

    def set_kernel_config(self, variant, tile_x, tile_y, z_depth=0):
        c = abi.KernelConfig(variant, tile_x, tile_y, z_depth)
        return self._check(self.lib.dll.acDeviceSetKernelConfig(self.handle, c), "acDeviceSetKernelConfig")

    def time_kernel_config(self, config: abi.KernelConfig, iters=5) -> float:
        return float(self.lib.dll.acDeviceTimeKernelConfig(self.handle, config, iters))

    def autotune(self):
        return self._check(self.lib.dll.acDeviceAutoOptimize(self.handle), "acDeviceAutoOptimize")

    def copy_vertex_buffer(self, handle, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceCopyVertexBuffer(self.handle, stream, _vh(handle)),
                           "acDeviceCopyVertexBuffer")

    def stream_handle(self, stream=abi.STREAM_DEFAULT):
        return self.lib.dll.acDeviceGetStreamHandle(self.handle, stream)

    # ---- halo pack / unpack; `ptr` is a device address (int) of a caller-owned buffer ----
    def pack_data(self, start, dims, ptr, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDevicePackData(self.handle, stream, abi.int3(*start),
                                                         abi.int3(*dims), ctypes.c_void_p(ptr)),
                           "acDevicePackData")

    def unpack_data(self, ptr, start, dims, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceUnpackData(self.handle, stream, ctypes.c_void_p(ptr),
                                                           abi.int3(*start), abi.int3(*dims)),
                           "acDeviceUnpackData")

    def halo_packed_size(self) -> int:
        return int(self.lib.dll.acDeviceHaloPackedSize(self.handle))

    def pack_halo(self, ptr, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDevicePackHalo(self.handle, stream, ctypes.c_void_p(ptr)),
                           "acDevicePackHalo")

    def unpack_halo(self, ptr, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceUnpackHalo(self.handle, stream, ctypes.c_void_p(ptr)),
                           "acDeviceUnpackHalo")

    def halo_exchange_self(self, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acDeviceHaloExchangeSelf(self.handle, stream),
                           "acDeviceHaloExchangeSelf")


class Ensemble:
    """Handle of acEnsembleCreate: independent periodic boxes advanced together
    (include/astaroth_ext.h)."""

    DIAGNOSTICS = ("max_u", "rms_u", "max_a", "rms_a", "max_lnrho", "min_lnrho", "rms_lnrho")

    def __init__(self, lib: Library, infos, device_id: int = 0, autotune=None):
        self.lib = lib
        self.infos = list(infos)
        arr = (abi.structs(lib.double)[0] * len(self.infos))(*self.infos)
        self.handle = ctypes.c_void_p()
        old = os.environ.get("AC_AUTOTUNE")
        if autotune is not None:
            os.environ["AC_AUTOTUNE"] = "1" if autotune else "0"
        try:
            res = lib.dll.acEnsembleCreate(device_id, len(self.infos), arr, ctypes.byref(self.handle))
        finally:
            if autotune is not None:
                if old is None:
                    del os.environ["AC_AUTOTUNE"]
                else:
                    os.environ["AC_AUTOTUNE"] = old
        if res != abi.AC_SUCCESS:
            raise AstarothError("acEnsembleCreate failed")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def destroy(self):
        if self.handle:
            self.lib.dll.acEnsembleDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def _check(self, res, what):
        if res != abi.AC_SUCCESS:
            raise AstarothError("%s returned AC_FAILURE" % what)
        return res

    @property
    def nboxes(self):
        return len(self.infos)

    def set_kernel_config(self, tile_x, tile_y, z_depth=0):
        c = abi.KernelConfig(3, tile_x, tile_y, z_depth)
        return self._check(self.lib.dll.acEnsembleSetKernelConfig(self.handle, c), "acEnsembleSetKernelConfig")

    def kernel_config(self) -> abi.KernelConfig:
        c = abi.KernelConfig()
        self._check(self.lib.dll.acEnsembleGetKernelConfig(self.handle, ctypes.byref(c)), "acEnsembleGetKernelConfig")
        return c

    def autotune(self):
        return self._check(self.lib.dll.acEnsembleAutoOptimize(self.handle), "acEnsembleAutoOptimize")

    def box_z_depths(self):
        """Output planes per work item of every box (0: whole columns)."""
        return [self.lib.dll.acEnsembleBoxZDepth(self.handle, b) for b in range(self.nboxes)]

    def randomize_uniform(self, box, handle, lo, hi, seed):
        return self._check(self.lib.dll.acEnsembleRandomizeUniform(self.handle, box, _vh(handle), lo, hi, seed),
                           "acEnsembleRandomizeUniform")

    def set_vertex_buffer(self, box, handle, value):
        return self._check(self.lib.dll.acEnsembleSetVertexBuffer(self.handle, box, _vh(handle), value),
                           "acEnsembleSetVertexBuffer")

    def load_mesh(self, box, mesh: HostMesh):
        return self._check(self.lib.dll.acEnsembleLoadMesh(self.handle, box, mesh.struct), "acEnsembleLoadMesh")

    def store_mesh(self, box, mesh: HostMesh):
        return self._check(self.lib.dll.acEnsembleStoreMesh(self.handle, box, ctypes.byref(mesh.struct)),
                           "acEnsembleStoreMesh")

    def periodic_boundconds(self):
        return self._check(self.lib.dll.acEnsemblePeriodicBoundconds(self.handle), "acEnsemblePeriodicBoundconds")

    def integrate_step(self, dt, mode="batched"):
        fn = {"batched": self.lib.dll.acEnsembleIntegrateStep, "graph": self.lib.dll.acEnsembleIntegrateStepGraph,
              "looped": self.lib.dll.acEnsembleIntegrateStepLooped}[mode]
        return self._check(fn(self.handle, self.lib.real(dt)), "acEnsembleIntegrateStep(%s)" % mode)

    def diagnostics(self) -> np.ndarray:
        out = np.zeros((self.nboxes, 7), dtype=np.float64)
        self._check(self.lib.dll.acEnsembleDiagnostics(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "acEnsembleDiagnostics")
        return out

    def synchronize(self):
        return self._check(self.lib.dll.acEnsembleSynchronize(self.handle), "acEnsembleSynchronize")

    def timestep_cfl(self) -> np.ndarray:
        out = np.zeros(self.nboxes, dtype=np.float64)
        self._check(self.lib.dll.acEnsembleTimestepCFL(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "acEnsembleTimestepCFL")
        return out

    def stream_handle(self):
        return self.lib.dll.acEnsembleGetStreamHandle(self.handle)


def _rt(r):
    return RT[r] if isinstance(r, str) else int(r)


def _vh(h):
    return VTX[h] if isinstance(h, str) else int(h)


class Node:
    """The Node interface (acNode*) of include/astaroth.h: one Device on GPU 0 for the whole grid."""

    def __init__(self, lib: Library, info, node_id: int = 0, autotune=None):
        self.lib = lib
        self.info = info
        self.dims = dims_of(info)
        self.handle = ctypes.c_void_p()
        old = os.environ.get("AC_AUTOTUNE")
        if autotune is not None:
            os.environ["AC_AUTOTUNE"] = "1" if autotune else "0"
        try:
            res = lib.dll.acNodeCreate(int(node_id), info, ctypes.byref(self.handle))
        finally:
            if autotune is not None:
                if old is None:
                    del os.environ["AC_AUTOTUNE"]
                else:
                    os.environ["AC_AUTOTUNE"] = old
        if res != abi.AC_SUCCESS:
            raise AstarothError("acNodeCreate failed")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def _check(self, res, what):
        if res != abi.AC_SUCCESS:
            raise AstarothError("%s returned AC_FAILURE" % what)
        return res

    def destroy(self):
        if self.handle:
            self.lib.dll.acNodeDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def synchronize(self, stream=abi.STREAM_ALL):
        return self._check(self.lib.dll.acNodeSynchronizeStream(self.handle, stream), "acNodeSynchronizeStream")

    def query_device_configuration(self) -> abi.DeviceConfiguration:
        c = abi.DeviceConfiguration()
        self._check(self.lib.dll.acNodeQueryDeviceConfiguration(self.handle, ctypes.byref(c)),
                    "acNodeQueryDeviceConfiguration")
        return c

    def load_mesh(self, mesh: HostMesh, stream=abi.STREAM_DEFAULT):
        self._check(self.lib.dll.acNodeLoadMesh(self.handle, stream, mesh.struct), "acNodeLoadMesh")
        self.synchronize(stream)

    def store_mesh(self, mesh: HostMesh, stream=abi.STREAM_DEFAULT):
        self._check(self.lib.dll.acNodeStoreMesh(self.handle, stream, ctypes.byref(mesh.struct)), "acNodeStoreMesh")
        self.synchronize(stream)
        return mesh

    def integrate(self, dt):
        return self._check(self.lib.dll.acNodeIntegrate(self.handle, self.lib.real(dt)), "acNodeIntegrate")

    def integrate_gbc(self, config, dt):
        return self._check(self.lib.dll.acNodeIntegrateGBC(self.handle, config, self.lib.real(dt)),
                           "acNodeIntegrateGBC")

    def integrate_substep(self, step, start, end, dt, stream=abi.STREAM_DEFAULT):
        return self._check(self.lib.dll.acNodeIntegrateSubstep(self.handle, stream, step, int3(*start),
                                                               int3(*end), self.lib.real(dt)),
                           "acNodeIntegrateSubstep")

    def swap_buffers(self):
        return self._check(self.lib.dll.acNodeSwapBuffers(self.handle), "acNodeSwapBuffers")

    def periodic_boundconds(self, stream=abi.STREAM_DEFAULT):
        self._check(self.lib.dll.acNodePeriodicBoundconds(self.handle, stream), "acNodePeriodicBoundconds")
        self.synchronize(stream)

    def reduce_scal(self, rtype, handle, stream=abi.STREAM_DEFAULT):
        out = self.lib.real()
        self._check(self.lib.dll.acNodeReduceScal(self.handle, stream, _rt(rtype), _vh(handle), ctypes.byref(out)),
                    "acNodeReduceScal")
        return out.value

    def reduce_vec(self, rtype, a, b, c, stream=abi.STREAM_DEFAULT):
        out = self.lib.real()
        self._check(self.lib.dll.acNodeReduceVec(self.handle, stream, _rt(rtype), _vh(a), _vh(b), _vh(c),
                                                 ctypes.byref(out)), "acNodeReduceVec")
        return out.value
