"""Snapshot and time-series files (host side, no GPU): the reference standalone driver's formats
(samples/standalone/simulation.cc:62-147, 251-362) written through the C ABI and read back by the
library and by astaroth_amd.snapshot (the reference analysis reader's interface,
analysis/python/astar/data/read.py).

The byte layout is pinned by its definition rather than by the writer's own reader: a .mesh file
must equal, byte for byte, numpy's serialisation of [t, field values x fastest] in the
library's AcReal, and mesh_info.list lines must match the reference's printf formats.
"""
import os

import numpy as np
import pytest

from astaroth_amd import abi, snapshot
from astaroth_amd.api import AstarothError, HostMesh, Library


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


def random_mesh(lib, info, seed):
    nx, ny, nz = (info.int_params[abi.IP["AC_" + a]] for a in ("nx", "ny", "nz"))
    rng = np.random.default_rng(seed)
    return HostMesh(lib, info, [rng.standard_normal((nz + 6, ny + 6, nx + 6)) for _ in range(8)])


def test_mesh_info_list_format(lib, tmp_path):
    info = lib.config(n=16, ny=12, nz=10, AC_max_time=2.5)
    lib.write_mesh_info(info, str(tmp_path))
    text = (tmp_path / "mesh_info.list").read_text().splitlines(keepends=True)
    s = 8 if lib.double else 4
    # the reference's printf formats (simulation.cc:81-108)
    assert text[0] == "size_t AcRealSize %d \n" % s
    assert text[1] == "int endian 1 \n"  # little-endian host
    assert "int AC_nx 16\n" in text and "int AC_mx 22\n" in text and "int AC_nz 10\n" in text
    assert "real AC_dsx %g\n" % info.real_params[abi.RP["AC_dsx"]] in text
    assert "real AC_max_time 2.5\n" in text
    assert len(text) == 2 + abi.NUM_INT_PARAMS + abi.NUM_INT3_PARAMS + abi.NUM_REAL_PARAMS
    c = snapshot.MeshInfo(str(tmp_path) + "/").contents
    assert c["AcRealSize"] == s and c["endian"] == 1
    for name, k in abi.IP.items():
        assert c[name] == info.int_params[k], name
    for name, k in abi.RP.items():
        assert c[name] == float("%g" % info.real_params[k]), name
    for name, k in abi.I3P.items():
        v = info.int3_params[k]
        assert c[name] == [v.x, v.y, v.z], name


def test_snapshot_bytes_and_round_trip(lib, tmp_path):
    info = lib.config(n=9, ny=7, nz=5)
    mesh = random_mesh(lib, info, 3)
    t = 0.1234567
    d = str(tmp_path)
    lib.save_mesh_snapshot(mesh, 42, t, d)
    lib.write_mesh_info(info, d)
    dt = np.dtype(lib.dtype).newbyteorder("<")
    for q, name in enumerate(snapshot.FIELDS):
        raw = (tmp_path / ("%s_42.mesh" % name)).read_bytes()
        expect = np.concatenate([np.array([t], dtype=lib.dtype), mesh.fields[q].ravel()]).astype(dt).tobytes()
        assert raw == expect, name
    back = HostMesh(lib, info)
    t2 = lib.read_mesh_snapshot(back, 42, d)
    assert t2 == lib.dtype(t)
    for q in range(8):
        assert np.array_equal(back.fields[q], mesh.fields[q])
    # the analysis reader: array[x, y, z], the time stamp, every field
    m = snapshot.Mesh(42, d + "/")
    assert m.ok and m.timestamp == lib.dtype(t)
    assert np.array_equal(m.lnrho, mesh.fields[abi.VTX["VTXBUF_LNRHO"]].transpose(2, 1, 0))
    assert np.array_equal(m.uu[2], mesh.fields[abi.VTX["VTXBUF_UUZ"]].transpose(2, 1, 0))
    assert np.array_equal(m.aa[0], mesh.fields[abi.VTX["VTXBUF_AX"]].transpose(2, 1, 0))
    assert np.array_equal(m.ss, mesh.fields[abi.VTX["VTXBUF_ENTROPY"]].transpose(2, 1, 0))
    assert m.lnrho.shape == (15, 13, 11)
    assert np.allclose(m.xx[1], info.real_params[abi.RP["AC_dsx"]])
    assert snapshot.parse_directory(d) == [42]
    assert not snapshot.Mesh(7, d + "/").ok


def test_snapshot_read_errors(tmp_path):
    lib32, lib64 = Library(False), Library(True)
    info32 = lib32.config(n=6)
    d = str(tmp_path)
    lib32.save_mesh_snapshot(random_mesh(lib32, info32, 1), 1, 0.5, d)
    with pytest.raises(AstarothError):  # missing step
        lib32.read_mesh_snapshot(HostMesh(lib32, info32), 2, d)
    with pytest.raises(AstarothError):  # FP32 files are half the length an FP64 build expects
        lib64.read_mesh_snapshot(HostMesh(lib64, lib64.config(n=6)), 1, d)
    with pytest.raises(AstarothError):  # a different grid
        lib32.read_mesh_snapshot(HostMesh(lib32, lib32.config(n=7)), 1, d)
    p = tmp_path / "VTXBUF_AZ_1.mesh"
    p.write_bytes(p.read_bytes()[:-4])  # truncated
    with pytest.raises(AstarothError):
        lib32.read_mesh_snapshot(HostMesh(lib32, info32), 1, d)
    with pytest.raises(AstarothError):  # unwritable directory
        lib32.save_mesh_snapshot(random_mesh(lib32, info32, 1), 1, 0.5, d + "/no/such/dir")


def test_timeseries_title_and_resume_line(tmp_path):
    lib = Library(False)
    path = str(tmp_path / "timeseries.ts")
    lib.timeseries_begin(path, 0)
    lib.timeseries_begin(path, 5)  # a resumed run appends a line break only
    text = open(path).read()
    names = ["VTXBUF_LNRHO", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ", "VTXBUF_AX", "VTXBUF_AY",
             "VTXBUF_AZ", "VTXBUF_ENTROPY"]
    title = "step  t_step  dt  uu_total_min  uu_total_rms  uu_total_max  " + "".join(
        "%s_min  %s_rms  %s_max  " % (n, n, n) for n in names)
    assert text == title + "\n\n"
    # a row in the reference's format parses under the reference reader's column names
    with open(path, "a") as f:
        f.write("%d %e %e %e %e %e " % (3, 0.5, 0.01, 1.0, 2.0, 3.0)
                + "".join("%e %e %e " % (q, q + 0.5, q + 1) for q in range(8)) + "\n")
    ts = snapshot.parse_ts(str(tmp_path) + "/")
    assert ts["exist"] and list(ts["step"]) == [3]
    assert ts["t_step"][0] == 0.5 and ts["dt"][0] == 0.01
    assert ts["uutot_min"][0] == 1.0 and ts["uutot_max"][0] == 3.0
    assert ts["lnrho_min"][0] == 0.0 and ts["uux_rms"][0] == 1.5 and ts["aaz_max"][0] == 7.0
    assert ts["ss_max"][0] == 8.0
    assert not snapshot.parse_ts(str(tmp_path) + "/", "missing.ts")["exist"]
    assert os.path.exists(path)
