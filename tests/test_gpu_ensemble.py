"""Batched ensembles (include/astaroth_ext.h acEnsemble*) through the C ABI.

The batched step runs the lean z-march kernel over a (box, tile, z range) work list. Every box
must match the same box advanced alone by a Device: the per-cell operations of the z-march do
not depend on the tile shape or z-depth, but the compiler may fuse multiply-adds differently in
different kernel instantiations, so across instantiations the bar is a few ulps of the field's
magnitude (stated below). The graph step replays the eager step's kernels and must match it
bitwise. The per-box FP64 diagnostics must match the stored fields.
"""
import numpy as np
import pytest

from astaroth_amd.api import Library, HostMesh
from astaroth_amd.abi import VTX, RP
from astaroth_amd.initial import device_init_mhd, ensemble_init_mhd

pytestmark = pytest.mark.gpu

BOXES = [(32, 32, 32), (48, 32, 16), (24, 40, 20), (64, 16, 28), (13, 9, 11)]
NUS = [5e-3, 1e-3, 8e-3, 2.5e-3, 6e-3]


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


def box_infos(lib):
    infos = []
    for (nx, ny, nz), nu in zip(BOXES, NUS):
        info = lib.config(n=nx, ny=ny, nz=nz, AC_nu_visc=nu)
        infos.append(info)
    return infos


ULP_BAR = 8  # across kernel instantiations: max |a - b| <= 8 ulp of max |b| per field


def run_ensemble(lib, steps, mode, config=None):
    infos = box_infos(lib)
    with lib.ensemble_create(infos) as ens:
        if config is not None:
            ens.set_kernel_config(*config)
        for b in range(len(infos)):
            ensemble_init_mhd(ens, b, config_id=4, box_id=b)
        ens.periodic_boundconds()
        for _ in range(steps):
            ens.integrate_step(1e-3, mode=mode)
        diag = ens.diagnostics()
        out = []
        for b, info in enumerate(infos):
            m = HostMesh(lib, info)
            ens.store_mesh(b, m)
            out.append(m)
    return out, diag


def test_ensemble_matches_single_box_devices(lib):
    steps = 2
    batched, _ = run_ensemble(lib, steps, "batched")
    for b, info in enumerate(box_infos(lib)):
        with lib.device_create(0, info, autotune=False) as dev:
            dev.set_kernel_config(3, 16, 16, 0)
            device_init_mhd(dev, config_id=4, box_id=b)
            for _ in range(steps):
                dev.integrate_step(1e-3)
            m = HostMesh(lib, info)
            dev.store_mesh(m)
        for q in range(8):
            ulp = np.spacing(np.abs(m.interior(q)).max())
            err = np.abs(batched[b].interior(q).astype(np.float64) - m.interior(q)).max()
            assert err <= ULP_BAR * ulp, (BOXES[b], q, err / ulp)


@pytest.mark.parametrize("config", [(32, 4, 0), (16, 8, 14), (16, 16, 28)])
def test_ensemble_modes_and_configs_agree(lib, config):
    ref, dref = run_ensemble(lib, 2, "batched", config)
    for mode in ("graph", "looped"):
        got, dgot = run_ensemble(lib, 2, mode, config)
        for b in range(len(BOXES)):
            for q in range(8):
                if mode == "graph":  # the same kernels replayed
                    assert np.array_equal(ref[b].interior(q), got[b].interior(q)), (mode, config, b, q)
                else:
                    ulp = np.spacing(np.abs(ref[b].interior(q)).max())
                    err = np.abs(got[b].interior(q).astype(np.float64) - ref[b].interior(q)).max()
                    assert err <= ULP_BAR * ulp, (mode, config, b, q, err / ulp)
        assert np.allclose(dref, dgot, rtol=1e-6 if lib.dtype == np.float32 else 1e-13, atol=0)


def test_ensemble_graph_replays(lib):
    """Several replays of the recorded step (both buffer parities), with eager steps between."""
    a, _ = run_ensemble(lib, 4, "batched")
    infos = box_infos(lib)
    with lib.ensemble_create(infos) as ens:
        for b in range(len(infos)):
            ensemble_init_mhd(ens, b, config_id=4, box_id=b)
        ens.periodic_boundconds()
        ens.integrate_step(1e-3, mode="graph")
        ens.integrate_step(1e-3, mode="batched")
        ens.integrate_step(1e-3, mode="graph")
        ens.integrate_step(1e-3, mode="graph")
        for b, info in enumerate(infos):
            m = HostMesh(lib, info)
            ens.store_mesh(b, m)
            for q in range(8):
                assert np.array_equal(a[b].interior(q), m.interior(q)), (b, q)


def test_ensemble_diagnostics(lib):
    states, diag = run_ensemble(lib, 1, "batched")
    for b, m in enumerate(states):
        u = [m.interior(VTX[k]).astype(np.float64) for k in ("VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ")]
        a = [m.interior(VTX[k]).astype(np.float64) for k in ("VTXBUF_AX", "VTXBUF_AY", "VTXBUF_AZ")]
        lr = m.interior(VTX["VTXBUF_LNRHO"]).astype(np.float64)
        u2 = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
        a2 = a[0] ** 2 + a[1] ** 2 + a[2] ** 2
        exp = [np.sqrt(u2.max()), np.sqrt(u2.mean()), np.sqrt(a2.max()), np.sqrt(a2.mean()),
               lr.max(), lr.min(), np.sqrt((lr ** 2).mean())]
        assert np.allclose(diag[b], exp, rtol=1e-12, atol=0), (b, diag[b], exp)


def test_ensemble_rejects_bad_arguments(lib):
    infos = box_infos(lib)[:2]
    with lib.ensemble_create(infos) as ens:
        d = lib.dll
        assert d.acEnsembleRandomizeUniform(ens.handle, 2, 0, 0.0, 1.0, 1) != 0
        assert d.acEnsembleSetVertexBuffer(ens.handle, -1, 0, lib.real(1.0)) != 0
        assert d.acEnsembleIntegrateStep(ens.handle, lib.real(float("nan"))) != 0
        with pytest.raises(Exception):
            ens.set_kernel_config(64, 4, 0)


def test_ensemble_cfl_timestep(lib):
    """Per-box CFL dt from the last step's diagnostics: the single-box formula with each box's
    own spacing and viscosity."""
    import math
    infos = box_infos(lib)
    with lib.ensemble_create(infos) as ens:
        for b in range(len(infos)):
            ensemble_init_mhd(ens, b, config_id=4, box_id=b)
        ens.periodic_boundconds()
        ens.integrate_step(1e-3)
        d = ens.diagnostics()
        dts = ens.timestep_cfl()
    for b, info in enumerate(infos):
        rp = lambda k: float(info.real_params[RP[k]])
        dsmin = min(rp("AC_dsx"), rp("AC_dsy"), rp("AC_dsz"))
        uu_dt = rp("AC_cdt") * dsmin / (d[b, 0] + math.sqrt(rp("AC_cs2_sound")))
        visc_dt = rp("AC_cdtv") * dsmin ** 2 / max(max(rp("AC_nu_visc"), rp("AC_eta")), rp("AC_gamma"))
        assert dts[b] == pytest.approx(min(uu_dt, visc_dt), rel=1e-12)


@pytest.mark.parametrize("config", [(16, 16, 0), (32, 4, 14), (16, 8, 0)])
def test_batched_step_reads_no_ghost_cell(lib, config):
    """The batched solve reads periodic images, not ghost cells: a state whose ghosts hold NaN
    steps bitwise like the same state with filled ghosts, and after the step every ghost cell
    holds its periodic image (the step's final fill)."""
    infos = box_infos(lib)
    rng = np.random.default_rng(12)
    meshes = []
    for info, (nx, ny, nz) in zip(infos, BOXES):
        interior = [rng.uniform(-1e-3, 1e-3, (nz, ny, nx)) for _ in range(8)]
        meshes.append([np.pad(f, 3, mode="wrap") for f in interior])
    outs = []
    for poison in (False, True):
        with lib.ensemble_create(infos) as ens:
            ens.set_kernel_config(*config)
            for b, info in enumerate(infos):
                fields = [f.copy() for f in meshes[b]]
                if poison:
                    for f in fields:
                        f[:3] = np.nan
                        f[-3:] = np.nan
                        f[:, :3] = np.nan
                        f[:, -3:] = np.nan
                        f[:, :, :3] = np.nan
                        f[:, :, -3:] = np.nan
                ens.load_mesh(b, HostMesh(lib, info, fields))
            ens.integrate_step(1e-3, mode="batched")
            res = []
            for b, info in enumerate(infos):
                m = HostMesh(lib, info)
                ens.store_mesh(b, m)
                res.append(m)
        outs.append(res)
    for b in range(len(infos)):
        for q in range(8):
            a, p = outs[0][b].fields[q], outs[1][b].fields[q]
            assert np.array_equal(a, p), (BOXES[b], q)
            assert np.array_equal(a, np.pad(a[3:-3, 3:-3, 3:-3], 3, mode="wrap")), (BOXES[b], q)
