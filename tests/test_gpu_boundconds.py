"""General boundary conditions (acDeviceGeneralBoundconds / acDeviceGeneralBoundcondStep,
reference device.cc:457-475 over boundconds.cuh:3-128,192-242) and the steps that use them
(acDeviceIntegrateStepGBC, the legacy acIntegrateGBC / acBoundcondStepGBC).

The reference kernel writes 1.0 into every ghost cell and reads only the x-top type, so it
cannot be the oracle; the fill its comments describe is pinned against numpy.pad, applied axis
after axis with one side at a time: mode "wrap" for a periodic face, "reflect" (mirror about the
edge cell without repeating it: ghost 2, 1, 0 <- 4, 5, 6) for a symmetric face, and the negated
"reflect" for an antisymmetric face. Properties of the 6th-order stencils on the filled field
pin the mirror point independently: across a symmetric face the first derivative at the
boundary cell is exactly 0; across an antisymmetric face the second derivative there is
b0 f(b) / h^2 (the neighbour pairs cancel).
"""
import ctypes
import itertools

import numpy as np
import pytest

from astaroth_amd import abi
from astaroth_amd.api import AstarothError, HostMesh, Library
from astaroth_amd.initial import mhd_fields

pytestmark = pytest.mark.gpu

P, S, A = abi.AC_BOUNDCOND_PERIODIC, abi.AC_BOUNDCOND_SYMMETRIC, abi.AC_BOUNDCOND_ANTISYMMETRIC


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


def with_faces(lib, info, bot, top):
    for a, b, t in zip("xyz", bot, top):
        info.int_params[abi.IP["AC_bc_type_bot_" + a]] = b
        info.int_params[abi.IP["AC_bc_type_top_" + a]] = t
    return info


def pad_side(arr, axis, side, mode):
    width = [(0, 0)] * 3
    width[axis] = (3, 0) if side == 0 else (0, 3)
    if mode == P:
        out = np.pad(arr, width, mode="wrap")
    else:
        out = np.pad(arr, width, mode="reflect")
        if mode == A:
            sl = [slice(None)] * 3
            sl[axis] = slice(0, 3) if side == 0 else slice(-3, None)
            out[tuple(sl)] *= -1
    return out


def numpy_fill(field, bot, top):
    """Ghost cells of a (mz, my, mx) array from its interior, x then y then z."""
    out = field[3:-3, 3:-3, 3:-3]
    for axis_xyz in range(3):
        ax = 2 - axis_xyz  # array axis of x is 2
        if bot[axis_xyz] == P and top[axis_xyz] == P:
            w = [(0, 0)] * 3
            w[ax] = (3, 3)
            out = np.pad(out, w, mode="wrap")
        else:
            # one side at a time; the second side pads the already padded array, whose own
            # ghosts lie outside the mirror's reach (mirrored sources are interior cells)
            lo = pad_side(out, ax, 0, bot[axis_xyz])
            hi = pad_side(out, ax, 1, top[axis_xyz])
            sl_lo = [slice(None)] * 3
            sl_lo[ax] = slice(0, 3)
            out = np.concatenate([lo[tuple(sl_lo)], hi], axis=ax)
    return out


def random_fields(lib, n, seed):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((nz + 6, ny + 6, nx + 6)) for _ in range(8)]


FACES = [((P, P, P), (P, P, P)), ((S, S, S), (S, S, S)), ((A, A, A), (A, A, A)),
         ((S, A, P), (A, S, P)), ((A, P, S), (S, P, A)), ((P, S, A), (P, A, S))]


@pytest.mark.parametrize("bot,top", FACES)
def test_general_fill_matches_numpy_pad(lib, bot, top):
    n = (9, 7, 6)
    info = with_faces(lib, lib.config(n=n[0], ny=n[1], nz=n[2]), bot, top)
    fields = random_fields(lib, n, 5)
    mesh = HostMesh(lib, info, fields)
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(mesh)
        dev.general_boundconds(info)
        out = HostMesh(lib, info)
        dev.store_mesh(out)
    for q in range(8):
        exp = numpy_fill(mesh.fields[q], bot, top)
        assert np.array_equal(out.fields[q], exp), (q, bot, top)
        assert np.array_equal(out.fields[q][3:-3, 3:-3, 3:-3], mesh.fields[q][3:-3, 3:-3, 3:-3])


def test_all_periodic_equals_periodic_fill(lib):
    n = (12, 10, 8)
    info = lib.config(n=n[0], ny=n[1], nz=n[2])  # AC_bc_type_* = 0: periodic
    mesh = HostMesh(lib, info, random_fields(lib, n, 9))
    outs = []
    for general in (False, True):
        with lib.device_create(0, info, autotune=False) as dev:
            dev.load_mesh(mesh)
            if general:
                dev.general_boundconds(info)
            else:
                dev.periodic_boundconds()
            o = HostMesh(lib, info)
            dev.store_mesh(o)
            outs.append(o)
    for q in range(8):
        assert np.array_equal(outs[0].fields[q], outs[1].fields[q])


def test_mirror_point_by_stencil_properties():
    lib = Library(True)
    n = (10, 8, 8)
    info = with_faces(lib, lib.config(n=n[0], ny=n[1], nz=n[2]), (S, A, S), (A, S, P))
    fields = [np.asarray(f) for f in mhd_fields(*n, np.float64, config_id=5)]
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(HostMesh(lib, info, fields))
        dev.general_boundconds(info)
        out = HostMesh(lib, info)
        dev.store_mesh(out)
    a = (3 / 4, -3 / 20, 1 / 60)
    b0 = -49 / 18
    bk = (3 / 2, -3 / 20, 1 / 90)
    f = out.fields[abi.VTX["VTXBUF_LNRHO"]]
    nx = n[0]
    # x: bottom symmetric about i = 3, top antisymmetric about i = nx + 2
    d1 = sum(a[k - 1] * (f[:, :, 3 + k] - f[:, :, 3 - k]) for k in (1, 2, 3))
    assert np.all(d1[3:-3, 3:-3] == 0)
    c = nx + 2
    d2 = b0 * f[:, :, c] + sum(bk[k - 1] * (f[:, :, c + k] + f[:, :, c - k]) for k in (1, 2, 3))
    assert np.allclose(d2[3:-3, 3:-3], b0 * f[3:-3, 3:-3, c], rtol=0, atol=1e-15)
    # y: bottom antisymmetric about j = 3
    d2y = b0 * f[:, 3, :] + sum(bk[k - 1] * (f[:, 3 + k, :] + f[:, 3 - k, :]) for k in (1, 2, 3))
    assert np.allclose(d2y[3:-3, 3:-3], b0 * f[3:-3, 3, 3:-3], rtol=0, atol=1e-15)
    # the mirror is not about a half cell: the first ghost cell differs from the boundary cell
    assert not np.array_equal(f[3:-3, 3:-3, 2], f[3:-3, 3:-3, 3])


def test_single_field_and_region(lib):
    n = (8, 8, 8)
    info = with_faces(lib, lib.config(n=8), (S, S, A), (A, P, S))
    mesh = HostMesh(lib, info, random_fields(lib, n, 2))
    start, end = (0, 2, 1), (12, 14, 10)
    q0 = abi.VTX["VTXBUF_AY"]
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(mesh)
        dev.general_boundconds(info, start=start, end=end, handle="VTXBUF_AY")
        out = HostMesh(lib, info)
        dev.store_mesh(out)
    full = numpy_fill(mesh.fields[q0], (S, S, A), (A, P, S))
    region = np.zeros(mesh.shape, dtype=bool)
    region[start[2]:end[2], start[1]:end[1], start[0]:end[0]] = True
    for q in range(8):
        exp = mesh.fields[q] if q != q0 else np.where(region, full, mesh.fields[q])
        assert np.array_equal(out.fields[q], exp), q


def test_general_fill_rejects(lib):
    info = lib.config(n=8)
    with lib.device_create(0, info, autotune=False) as dev:
        with pytest.raises(AstarothError):  # multi-process bindex forms are not provided
            dev.general_boundconds(info, bindex=(0, -1, -1))
        bad = with_faces(lib, lib.config(n=8), (P, 7, P), (P, P, P))
        with pytest.raises(AstarothError):
            dev.general_boundconds(bad)
    small = with_faces(lib, lib.config(n=8, ny=3, nz=8), (P, S, P), (P, S, P))
    with lib.device_create(0, small, autotune=False) as dev:
        with pytest.raises(AstarothError):  # a 3-cell axis cannot be mirrored
            dev.general_boundconds(small)


def test_step_gbc(lib):
    n = (16, 12, 10)
    info = lib.config(n=n[0], ny=n[1], nz=n[2])
    fields = mhd_fields(*n, np.float64, config_id=1)
    outs = []
    for gbc in (False, True):
        with lib.device_create(0, info, autotune=False) as dev:
            dev.load_mesh(HostMesh(lib, info, fields))
            for _ in range(2):
                if gbc:
                    dev.integrate_step_gbc(info, lib.real(1e-3))
                else:
                    dev.integrate_step(lib.real(1e-3))
            o = HostMesh(lib, info)
            dev.store_mesh(o)
            outs.append(o)
    for q in range(8):  # periodic faces: the general step is the periodic step
        assert np.array_equal(outs[0].fields[q], outs[1].fields[q])
    # symmetric faces on x: stays finite and the interior changes
    info_s = with_faces(lib, lib.config(n=n[0], ny=n[1], nz=n[2]), (S, P, P), (S, P, P))
    with lib.device_create(0, info_s, autotune=False) as dev:
        dev.load_mesh(HostMesh(lib, info_s, fields))
        dev.integrate_step_gbc(info_s, lib.real(1e-3))
        o = HostMesh(lib, info_s)
        dev.store_mesh(o)
    assert all(np.all(np.isfinite(o.interior(q))) for q in range(8))
    assert not np.array_equal(o.interior(0), outs[1].interior(0))


def test_legacy_gbc_calls():
    lib = Library(False)
    d = lib.dll
    d.acInit.argtypes = [lib.MeshInfo]
    d.acLoad.argtypes = [lib.Mesh]
    d.acStore.argtypes = [ctypes.POINTER(lib.Mesh)]
    d.acBoundcondStepGBC.argtypes = [lib.MeshInfo]
    d.acIntegrateGBC.argtypes = [lib.MeshInfo, lib.real]
    n = (8, 8, 8)
    info = with_faces(lib, lib.config(n=8), (A, S, P), (S, A, P))
    mesh = HostMesh(lib, info, random_fields(lib, n, 4))
    assert d.acInit(info) == 0
    try:
        assert d.acLoad(mesh.struct) == 0
        assert d.acBoundcondStepGBC(info) == 0
        out = HostMesh(lib, info)
        assert d.acStore(ctypes.byref(out.struct)) == 0
        for q in range(8):
            assert np.array_equal(out.fields[q], numpy_fill(mesh.fields[q], (A, S, P), (S, A, P)))
        assert d.acIntegrateGBC(info, lib.real(1e-4)) == 0
    finally:
        d.acQuit()
