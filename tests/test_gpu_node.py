"""The Node interface (reference include/astaroth.h:355-500, src/core/node.cc) on one GPU,
through the C ABI.

* acNodeIntegrate / acNodeIntegrateGBC are the Device's step on one GPU (the reference's
  inner/outer split hides the halo exchange between GPUs): their interior, and with general
  boundaries every cell, equal acDeviceIntegrateStep / acDeviceIntegrateStepGBC bit for bit,
  whether the Device fills the ghost cells or reads their periodic images, on boxes with and
  without partial tiles and z ranges.
* acNodeIntegrateSubstep clips its region to the interior.
* acNodeReduceScal / acNodeReduceVec: the Device reductions, RMS types finalised as
  sqrt(sum / (n_x n_y n_z)) (node.cc:932-957); MAX unchanged.
* acNodeQueryDeviceConfiguration: one device, subgrid = grid; offset stores copy src -> dst;
  acGetNode is the legacy interface's view of its Device.
"""
import ctypes

import numpy as np
import pytest

from astaroth_amd import abi
from astaroth_amd.abi import IP, int3
from astaroth_amd.api import HostMesh, Library
from astaroth_amd.initial import mhd_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


def host_mesh(lib, info, seed_config=1):
    nx, ny, nz = (info.int_params[IP["AC_" + a]] for a in ("nx", "ny", "nz"))
    return HostMesh(lib, info, mhd_fields(nx, ny, nz, np.float64, config_id=seed_config))


@pytest.mark.parametrize("dims,images", [((24, 20, 18), 0), ((24, 20, 18), 1), ((24, 5, 18), 0),
                                         ((48, 40, 18), 0), ((40, 36, 64), 0), ((34, 70, 9), 0)],
                         ids=["fill", "images", "thin-y", "wide", "deep", "odd"])
def test_node_integrate_equals_device_step(lib, dims, images):
    info = lib.config(n=dims[0], ny=dims[1], nz=dims[2])
    mesh = host_mesh(lib, info)
    with lib.node_create(info, autotune=False) as node:
        # the Node's Device fills the ghost cells or reads their images
        dev0 = node.query_device_configuration().devices[0]
        assert lib.dll.acDeviceSetReadPeriodicImages(ctypes.c_void_p(dev0), images) == abi.AC_SUCCESS
        node.load_mesh(mesh)
        for _ in range(2):
            node.integrate(1e-3)
        got = node.store_mesh(mesh.copy())
    with lib.device_create(0, info, autotune=False) as dev:
        assert lib.dll.acDeviceSetReadPeriodicImages(dev.handle, images) == abi.AC_SUCCESS
        dev.load_mesh(mesh)
        for _ in range(2):
            dev.integrate_step(1e-3)
        ref = dev.store_mesh(mesh.copy())
    for q in range(8):
        assert np.array_equal(got.interior(q), ref.interior(q)), abi.VTXBUF_NAMES[q]


@pytest.mark.parametrize("dims", [(20, 16, 14), (40, 36, 20)], ids=["small", "wide"])
def test_node_integrate_gbc_equals_device(lib, dims):
    info = lib.config(n=dims[0], ny=dims[1], nz=dims[2])
    info.int_params[IP["AC_bc_type_bot_x"]] = abi.AC_BOUNDCOND_SYMMETRIC
    info.int_params[IP["AC_bc_type_top_x"]] = abi.AC_BOUNDCOND_ANTISYMMETRIC
    info.int_params[IP["AC_bc_type_bot_z"]] = abi.AC_BOUNDCOND_SYMMETRIC
    mesh = host_mesh(lib, info, seed_config=2)
    with lib.node_create(info, autotune=False) as node:
        node.load_mesh(mesh)
        node.integrate_gbc(info, 1e-3)
        got = node.store_mesh(mesh.copy())
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(mesh)
        dev.integrate_step_gbc(info, lib.real(1e-3))
        ref = dev.store_mesh(mesh.copy())
    for q in range(8):
        assert np.array_equal(got.fields[q], ref.fields[q]), abi.VTXBUF_NAMES[q]


def test_node_substep_clips_to_interior(lib):
    n = 16
    info = lib.config(n=n)
    mesh = host_mesh(lib, info)
    m = n + 6
    with lib.node_create(info, autotune=False) as node:
        node.load_mesh(mesh)
        node.periodic_boundconds()
        node.integrate_substep(0, (0, 0, 0), (m, m, m), 1e-3)  # reaches into the ghost zone
        node.synchronize()
        node.swap_buffers()
        got = node.store_mesh(mesh.copy())
        node.integrate_substep(0, (0, 0, 0), (3, 3, 3), 1e-3)  # no interior cell: nothing to do
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(mesh)
        dev.periodic_boundconds()
        assert dev.integrate_substep(0, (3, 3, 3), (n + 3,) * 3, dt=1e-3) == abi.AC_SUCCESS
        dev.synchronize()
        dev.swap_buffers()
        ref = dev.store_mesh(mesh.copy())
    for q in range(8):
        assert np.array_equal(got.interior(q), ref.interior(q)), abi.VTXBUF_NAMES[q]


def test_node_reductions_and_configuration(lib):
    nx, ny, nz = 20, 18, 12
    info = lib.config(n=nx, ny=ny, nz=nz)
    mesh = host_mesh(lib, info)
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(mesh)
        dev_max = dev.reduce_scal("RTYPE_MAX", "VTXBUF_LNRHO")
        dev_rms = dev.reduce_scal("RTYPE_RMS", "VTXBUF_UUX")
        dev_vrms = dev.reduce_vec("RTYPE_RMS", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ")
    with lib.node_create(info, autotune=False) as node:
        node.load_mesh(mesh)
        assert node.reduce_scal("RTYPE_MAX", "VTXBUF_LNRHO") == dev_max
        inv_n = 1.0 / (nx * ny * nz)
        assert node.reduce_scal("RTYPE_RMS", "VTXBUF_UUX") == pytest.approx(
            np.sqrt(inv_n * dev_rms), rel=1e-6)
        assert node.reduce_vec("RTYPE_RMS", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ") == pytest.approx(
            np.sqrt(inv_n * dev_vrms), rel=1e-6)
        c = node.query_device_configuration()
        assert c.num_devices == 1 and bool(c.devices[0])
        assert tuple(c.grid.n) == (nx, ny, nz) and tuple(c.grid.m) == (nx + 6, ny + 6, nz + 6)
        assert tuple(c.subgrid.n) == tuple(c.grid.n) and tuple(c.subgrid.m) == tuple(c.grid.m)
        # offset store: plane 4 of the device -> plane 7 of the host buffer
        out = HostMesh(lib, info)
        mx, my = nx + 6, ny + 6
        r = lib.dll.acNodeStoreMeshWithOffset(node.handle, abi.STREAM_DEFAULT, int3(0, 0, 4),
                                              int3(0, 0, 7), mx * my, ctypes.byref(out.struct))
        assert r == abi.AC_SUCCESS
        node.synchronize()
        for q in range(8):
            assert np.array_equal(out.fields[q][7], mesh.fields[q][4])
            assert not out.fields[q][4].any()


def test_get_node_views_the_legacy_device(lib):
    info = lib.config(n=16)
    mesh = host_mesh(lib, info)
    assert lib.dll.acInit(info) == abi.AC_SUCCESS
    try:
        assert lib.dll.acLoad(mesh.struct) == abi.AC_SUCCESS
        h = lib.dll.acGetNode()
        assert h
        out = lib.real()
        assert lib.dll.acNodeReduceScal(ctypes.c_void_p(h), abi.STREAM_DEFAULT, abi.RT["RTYPE_MAX"],
                                        abi.VTX["VTXBUF_ENTROPY"], ctypes.byref(out)) == abi.AC_SUCCESS
        assert out.value == lib.dll.acReduceScal(abi.RT["RTYPE_MAX"], abi.VTX["VTXBUF_ENTROPY"])
    finally:
        assert lib.dll.acQuit() == abi.AC_SUCCESS


def test_node_load_with_offset_copies_src_to_dst(lib):
    """acNodeLoadVertexBufferWithOffset copies num_vertices cells from `src` of the host mesh to
    `dst` of the device, as the Device form does (the reference Node passes `src` for both)."""
    n = 12
    info = lib.config(n=n)
    mesh = host_mesh(lib, info)
    m = n + 6
    with lib.node_create(info, autotune=False) as node:
        zero = HostMesh(lib, info)
        node.load_mesh(zero)
        q = abi.VTX["VTXBUF_UUY"]
        r = lib.dll.acNodeLoadVertexBufferWithOffset(node.handle, abi.STREAM_DEFAULT, mesh.struct, q,
                                                     int3(0, 0, 2), int3(0, 0, 9), 2 * m * m)
        assert r == abi.AC_SUCCESS
        node.synchronize()
        out = node.store_mesh(HostMesh(lib, info))
    assert np.array_equal(out.fields[q][9:11], mesh.fields[q][2:4])
    assert not out.fields[q][:9].any() and not out.fields[q][11:].any()
    for other in range(8):
        if other != q:
            assert not out.fields[other].any()

