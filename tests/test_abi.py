"""CPU checks of the C-ABI libraries: they load, export every symbol include/*.h declares, agree
with the reference's struct layout (sizeof(AcMeshInfo) = 424 / 728 bytes, include/astaroth.h:
109-114) and enum order, and the host-side helpers behave as the reference's
(src/core/astaroth.cc:190-266, src/utils/config_loader.c:49-112). No GPU call is made."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from astaroth_amd import abi
from astaroth_amd.api import Library, HostMesh, lib_path

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:AcResult|AcReal|int|float|void\s*\*|void|Error|bool)\s+(ac\w+)\s*\(", text, re.M)))


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    if not os.path.exists(lib_path(request.param)):
        pytest.fail("library not built: " + lib_path(request.param))
    return Library(request.param)


def test_exports_every_declared_symbol(lib):
    for header in ("astaroth.h", "astaroth_ext.h", "astaroth_fortran.h"):
        names = declared_functions(header)
        assert names, header
        for name in names:
            assert hasattr(lib.dll, name), (header, name)
    # acLoadConfig is the one utils function the library provides
    assert hasattr(lib.dll, "acLoadConfig")


def test_oracle_exports_utils_api():
    for prec in ("fp32", "fp64"):
        dll = ctypes.CDLL(os.path.join(REPO, "oracle", "liboracle_%s.so" % prec))
        for name in declared_functions("astaroth_utils.h"):
            if name == "acLoadConfig":
                continue
            assert hasattr(dll, name), name


def test_struct_layout(lib):
    MI, M = abi.structs(lib.double)
    assert ctypes.sizeof(MI) == (728 if lib.double else 424)
    size_mi, size_m = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.dll.acAbiRealSize(ctypes.byref(size_mi), ctypes.byref(size_m)) == (8 if lib.double else 4)
    assert size_mi.value == ctypes.sizeof(MI) and size_m.value == ctypes.sizeof(M)


def test_enum_order_matches_reference_generator():
    # builtin int params first (acc/src/code_generator.c:491-500), then the module's
    assert abi.INT_PARAMS[:3] == ["AC_nx", "AC_ny", "AC_nz"] and abi.IP["AC_nxyz"] == 14
    assert abi.IP["AC_max_steps"] == 15 and abi.IP["AC_bc_type_bot_z"] == 24
    assert abi.NUM_INT_PARAMS == 25 and abi.NUM_INT3_PARAMS == 2 and abi.NUM_REAL_PARAMS == 75
    assert abi.RP["AC_dt"] == 6 and abi.RP["AC_nu_visc"] == 35 and abi.RP["AC_cs2_sound"] == 74
    assert abi.VTXBUF_NAMES == ["VTXBUF_LNRHO", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ",
                                "VTXBUF_AX", "VTXBUF_AY", "VTXBUF_AZ", "VTXBUF_ENTROPY"]
    assert abi.RT["RTYPE_SUM"] == 7 and abi.RT["RTYPE_ALFVEN_RMS"] == 6


def test_builtin_params(lib):
    info = lib.config(n=32, ny=16, nz=8)
    ip = info.int_params
    assert (ip[abi.IP["AC_mx"]], ip[abi.IP["AC_my"]], ip[abi.IP["AC_mz"]]) == (38, 22, 14)
    assert (ip[abi.IP["AC_nx_min"]], ip[abi.IP["AC_nx_max"]]) == (3, 35)
    assert (ip[abi.IP["AC_nz_min"]], ip[abi.IP["AC_nz_max"]]) == (3, 11)
    assert ip[abi.IP["AC_mxy"]] == 38 * 22 and ip[abi.IP["AC_nxyz"]] == 32 * 16 * 8
    inv = info.real_params[abi.RP["AC_inv_dsx"]]
    assert inv == pytest.approx(32 / (2 * math.pi), rel=1e-6)


def test_load_config(lib, tmp_path):
    cfg = tmp_path / "astaroth.conf"
    cfg.write_text("/* comment */\nAC_nx = 24\nAC_ny = 20\nAC_nz = 16\n\nAC_dsx = 0.25\n"
                   "AC_dsy = 0.5\nAC_dsz = 0.125\n// Hydro\nAC_nu_visc = 5e-3\nAC_bin_save_t = 1e666\n")
    info, res = lib.load_config(str(cfg))
    assert res == abi.AC_FAILURE  # most parameters are unset, exactly as with the reference
    ip = info.int_params
    assert (ip[abi.IP["AC_nx"]], ip[abi.IP["AC_mx"]], ip[abi.IP["AC_nxyz"]]) == (24, 30, 24 * 20 * 16)
    assert info.real_params[abi.RP["AC_inv_dsz"]] == pytest.approx(8.0)
    assert info.real_params[abi.RP["AC_nu_visc"]] == pytest.approx(5e-3)
    assert math.isinf(info.real_params[abi.RP["AC_bin_save_t"]])
    assert math.isnan(info.real_params[abi.RP["AC_eta"]])  # unset reals stay 0xFF..FF
    assert ip[abi.IP["AC_max_steps"]] == -1


def test_host_mesh_helpers(lib):
    info = lib.config(n=8)
    mesh = lib.Mesh()
    assert lib.dll.acHostMeshCreate(info, ctypes.byref(mesh)) == 0
    assert lib.dll.acHostMeshRandomize(ctypes.byref(mesh)) == 0
    m = 14**3
    vals = np.ctypeslib.as_array(mesh.vertex_buffer[3], shape=(m,)).copy()
    assert vals.min() >= 0 and vals.max() < 1 and vals.std() > 0.2
    assert lib.dll.acHostMeshDestroy(ctypes.byref(mesh)) == 0
    hm = HostMesh(lib, info)
    assert hm.fields[0].shape == (14, 14, 14)


def test_device_create_rejects_stale_builtin_sizes(lib, capfd):
    """acDeviceCreate refuses a configuration whose AC_mx/my/mz (which size every load, store and
    fill) disagree with the grid, before it touches a GPU: unset, and stale from another grid."""
    info = lib.config(n=16, ny=12, nz=8)
    handle = ctypes.c_void_p()
    unset = lib.MeshInfo.from_buffer_copy(info)
    for k in ("AC_mx", "AC_my", "AC_mz", "AC_mxy"):
        unset.int_params[abi.IP[k]] = 0
    stale = lib.MeshInfo.from_buffer_copy(info)
    stale.int_params[abi.IP["AC_nx"]] = 10  # grid changed after the sizes were derived
    for bad in (unset, stale):
        assert lib.dll.acDeviceCreate(0, bad, ctypes.byref(handle)) == abi.AC_FAILURE
        assert not handle.value
        assert "call acHostUpdateBuiltinParams first" in capfd.readouterr().err
