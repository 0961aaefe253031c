"""The reference standalone driver's loop, snapshots and diagnostics on the GPU
(samples/standalone/simulation.cc:251-560), through the C ABI: acDeviceTimeseriesAppend,
acDeviceSaveMeshSnapshot / acDeviceLoadMeshSnapshot and acDeviceRunSimulation.

Checks:
* a diagnostics row equals numpy statistics of the stored state (min, sqrt(sum/N), max of |u|
  and of every field over the interior), to the 6 digits the row prints;
* a device snapshot is the periodic-filled `in` state, and loading it restores that state;
* the driver writes rows and snapshots at the steps the reference would, stops for the
  reference's reasons, and a run resumed from a snapshot ends bitwise where an uninterrupted
  run ends (the driver's dt comes from the state, so the two runs take the same steps).
"""
import numpy as np
import pytest

from astaroth_amd import abi, snapshot
from astaroth_amd.api import HostMesh, Library
from astaroth_amd.initial import mhd_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


def set_int(info, name, v):
    info.int_params[abi.IP[name]] = v


def set_real(info, name, v):
    info.real_params[abi.RP[name]] = v


def loaded_device(lib, info, seed_config=1):
    nx, ny, nz = (info.int_params[abi.IP["AC_" + a]] for a in ("nx", "ny", "nz"))
    dev = lib.device_create(0, info, autotune=False)
    dev.load_mesh(HostMesh(lib, info, mhd_fields(nx, ny, nz, np.float64, config_id=seed_config)))
    dev.synchronize()
    return dev


def stats(a):
    a = a.astype(np.float64)
    return a.min(), np.sqrt(np.mean(a * a)), a.max()


def test_timeseries_row_matches_numpy(lib, tmp_path):
    info = lib.config(n=20, ny=16, nz=12)
    with loaded_device(lib, info) as dev:
        dev.integrate_step(1e-3)
        path = str(tmp_path / "timeseries.ts")
        lib.timeseries_begin(path, 0)
        assert dev.timeseries_append(1, 1e-3, 1e-3, path) is False
        out = HostMesh(lib, info)
        dev.store_mesh(out)
    ts = snapshot.parse_ts(str(tmp_path) + "/")
    assert list(ts["step"]) == [1]
    inner = [out.interior(q) for q in range(8)]
    u = np.sqrt(sum(inner[abi.VTX[n]].astype(np.float64) ** 2
                    for n in ("VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ")))
    expect = {"uutot": stats(u)}
    for name, q in abi.VTX.items():
        expect[snapshot._ts_key(name)] = stats(inner[q])
    for key, (mn, rms, mx) in expect.items():
        for suffix, v in (("_min", mn), ("_rms", rms), ("_max", mx)):
            got = ts[key + suffix][0]
            assert got == pytest.approx(v, rel=6e-6, abs=1e-300), (key + suffix, got, v)


def test_device_snapshot_round_trip(lib, tmp_path):
    info = lib.config(n=16, ny=12, nz=10)
    d = str(tmp_path)
    with loaded_device(lib, info) as dev:
        dev.integrate_step(1e-3)
        dev.save_mesh_snapshot(5, 0.25, d)
        dev.periodic_boundconds()
        ref = HostMesh(lib, info)
        dev.store_mesh(ref)
    host = HostMesh(lib, info)
    assert lib.read_mesh_snapshot(host, 5, d) == lib.dtype(0.25)
    for q in range(8):
        assert np.array_equal(host.fields[q], ref.fields[q])
    with lib.device_create(0, info, autotune=False) as dev2:
        assert dev2.load_mesh_snapshot(5, d) == lib.dtype(0.25)
        back = HostMesh(lib, info)
        dev2.store_mesh(back)
    for q in range(8):
        assert np.array_equal(back.fields[q], ref.fields[q])


def driver_info(lib, max_steps, save_steps=2, bin_steps=3, start_step=0, max_time=0.0,
                bin_save_t=1e30):
    info = lib.config(n=16, ny=16, nz=12)
    set_int(info, "AC_max_steps", max_steps)
    set_int(info, "AC_save_steps", save_steps)
    set_int(info, "AC_bin_steps", bin_steps)
    set_int(info, "AC_start_step", start_step)
    set_real(info, "AC_max_time", max_time)
    set_real(info, "AC_bin_save_t", bin_save_t)
    return info


def test_run_simulation_outputs_and_resume(lib, tmp_path):
    full, part = tmp_path / "full", tmp_path / "part"
    full.mkdir()
    part.mkdir()
    # uninterrupted: steps 1..6
    with loaded_device(lib, driver_info(lib, 7)) as dev:
        last, t, stop = dev.run_simulation(str(full))
    assert (last, abi.STOP_REASONS[stop]) == (6, "max_steps")
    assert snapshot.parse_directory(str(full)) == [0, 3, 6]
    ts = snapshot.parse_ts(str(full) + "/")
    assert list(ts["step"]) == [0, 2, 4, 6]
    assert ts["dt"][0] == 0 and ts["t_step"][0] == 0
    assert np.all(ts["dt"][1:] > 0) and ts["t_step"][-1] == pytest.approx(float(t), rel=1e-6)
    m6 = snapshot.Mesh(6, str(full) + "/")
    assert m6.ok and m6.timestamp == t
    # the same run stopped after step 3, then resumed from snapshot 3 by a new Device
    with loaded_device(lib, driver_info(lib, 4)) as dev:
        last, _, _ = dev.run_simulation(str(part))
    assert last == 3 and snapshot.parse_directory(str(part)) == [0, 3]
    info = driver_info(lib, 7, start_step=3)
    with lib.device_create(0, info, autotune=False) as dev:
        last, t2, stop = dev.run_simulation(str(part))
    assert (last, abi.STOP_REASONS[stop]) == (6, "max_steps") and t2 == t
    r6 = snapshot.Mesh(6, str(part) + "/")
    assert r6.timestamp == m6.timestamp
    for a, b in ((m6.lnrho, r6.lnrho), (m6.ss, r6.ss)) + tuple(zip(m6.uu, r6.uu)) + tuple(zip(m6.aa, r6.aa)):
        assert np.array_equal(a, b)
    ts2 = snapshot.parse_ts(str(part) + "/")  # rows 0, 2 from the first run, 4, 6 after the resume
    assert list(ts2["step"]) == [0, 2, 4, 6]
    assert np.array_equal(ts2["uutot_max"], ts["uutot_max"])


def test_run_simulation_stops(lib, tmp_path):
    # STOP file: snapshot of the step it was seen at
    d = tmp_path / "stopfile"
    d.mkdir()
    (d / "STOP").write_text("")
    with loaded_device(lib, driver_info(lib, 50)) as dev:
        last, _, stop = dev.run_simulation(str(d))
    assert (last, abi.STOP_REASONS[stop]) == (1, "stop_file")
    assert snapshot.parse_directory(str(d)) == [0, 1]
    # max_time: stops on the step that reaches it, without another snapshot
    d = tmp_path / "maxtime"
    d.mkdir()
    with loaded_device(lib, driver_info(lib, 50, max_time=1e-9)) as dev:
        last, t, stop = dev.run_simulation(str(d))
    assert (last, abi.STOP_REASONS[stop]) == (1, "max_time") and t >= 1e-9
    assert snapshot.parse_directory(str(d)) == [0]
    # bin_save_t: a snapshot each time t passes the next multiple
    d = tmp_path / "bint"
    d.mkdir()
    info = driver_info(lib, 6, bin_steps=1000)
    with loaded_device(lib, info) as dev:
        dt = dev.timestep_cfl()
    info = driver_info(lib, 6, bin_steps=1000, bin_save_t=1.5 * dt)
    with loaded_device(lib, info) as dev:
        dev.run_simulation(str(d))
    steps = snapshot.parse_directory(str(d))
    assert steps[0] == 0 and len(steps) >= 3


def test_run_simulation_nan_stop(tmp_path):
    lib = Library(False)
    info = driver_info(lib, 50, save_steps=1)
    nx, ny, nz = 16, 16, 12
    fields = mhd_fields(nx, ny, nz, np.float64, config_id=1)
    fields[abi.VTX["VTXBUF_ENTROPY"]][6, 7, 8] = np.nan
    with lib.device_create(0, info, autotune=False) as dev:
        dev.load_mesh(HostMesh(lib, info, fields))
        dev.synchronize()
        last, _, stop = dev.run_simulation(str(tmp_path))
    assert abi.STOP_REASONS[stop] == "nan" and last == 1
    assert snapshot.parse_directory(str(tmp_path)) == [0, 1]
