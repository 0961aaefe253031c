"""The 96-box ensemble workload against the oracle: all 96 boxes of the workload (edges from
{32, 48, 64, 96, 128}, cubic and anisotropic, nu ~ U[1e-3, 1e-2] per box, on-device seeded
uniform-field states) advanced 2 RK3 steps at dt = 1e-3 by acEnsembleIntegrateStepGraph (one
persistent batched launch per substep, the step recorded in a HIP graph), in FP32 and in FP64,
with the library's own kernel configuration. Every box is compared with the oracle, as the
reference verifies a caller's steps against the host model (samples/mpitest/main.cc:343-405).

Oracle: the CPU restatement of the host model solver in the two-buffer RK3 form of the GPU path
(integration.cuh:14-37), FP64, each box on its own with its own nu, started from the FP64
library's initial state of that box (the on-device generator equals the host one bit for bit:
tests/test_gpu_parity.py). Per-box diagnostics are compared with the host reductions
(modelreduce.c:76-209).

Bars: FP64, acEvalError (verification.c:78-102) on every field of every box or, where a field
misses its 5 ulps after the two steps, within 5 ulps of the long-double oracle around the worst
cell (or within twice the FP64 oracle's own distance from it; oracle_slices.fp64_verdict),
diagnostics within
1e-12 relative; FP32, range-normalised error <= 1e-5 per field of every box, diagnostics within
1e-5 relative.
"""
import numpy as np
import pytest

import pyoracle
from astaroth_amd import abi
from astaroth_amd.api import HostMesh, Library, real_params_double
from astaroth_amd.initial import ensemble96_boxes, ensemble_init_mhd
from oracle_slices import fp64_verdict

pytestmark = pytest.mark.gpu

STEPS, DT, CONFIG_ID = 2, 1e-3, 4
BOXES, NUS = ensemble96_boxes()
NAMES = abi.VTXBUF_NAMES


def box_infos(lib):
    return [lib.config(n=b[0], ny=b[1], nz=b[2], AC_nu_visc=nu) for b, nu in zip(BOXES, NUS)]


def init_all(ens):
    for b in range(ens.nboxes):
        ensemble_init_mhd(ens, b, config_id=CONFIG_ID, box_id=b)
    ens.periodic_boundconds()


def diagnostics_of(f):
    """max|u|, rms|u|, max|A|, rms|A|, max lnrho, min lnrho, rms lnrho by the host reductions."""
    u, a, lr = [f[1], f[2], f[3]], [f[4], f[5], f[6]], [f[0]]
    r = pyoracle.reduce
    return np.array([r(u, "RTYPE_MAX"), r(u, "RTYPE_RMS"), r(a, "RTYPE_MAX"), r(a, "RTYPE_RMS"),
                     r(lr, "RTYPE_MAX"), r(lr, "RTYPE_MIN"), r(lr, "RTYPE_RMS")], dtype=np.float64)


@pytest.fixture(scope="module")
def oracle_boxes():
    pyoracle.set_threads(16)
    lib = Library(True)
    infos = box_infos(lib)
    ref, diag = [], []
    with lib.ensemble_create(infos) as ens:
        init_all(ens)
        for b, info in enumerate(infos):
            m = HostMesh(lib, info)
            ens.store_mesh(b, m)
            f = m.fields
            pyoracle.integrate_steps_gpu_form(f, real_params_double(info), DT, STEPS)
            diag.append(diagnostics_of(f))
            ref.append([x[3:-3, 3:-3, 3:-3].copy() for x in f])
    print("oracle: %d boxes, %.3g cells" % (len(ref), sum(x * y * z for x, y, z in BOXES)), flush=True)
    return ref, np.array(diag)


def run_library(lib):
    infos = box_infos(lib)
    out = []
    with lib.ensemble_create(infos) as ens:
        print("%s ensemble kernel configuration %s" % ("fp64" if lib.double else "fp32",
                                                       ens.kernel_config().as_tuple()), flush=True)
        init_all(ens)
        for _ in range(STEPS):
            ens.integrate_step(DT, mode="graph")
        diag = ens.diagnostics()
        for b, info in enumerate(infos):
            m = HostMesh(lib, info)
            ens.store_mesh(b, m)
            out.append([m.interior(q) for q in range(8)])
    return out, diag


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_ensemble96_against_oracle(oracle_boxes, double):
    ref, ref_diag = oracle_boxes
    lib = Library(double)
    infos = box_infos(lib)
    got, diag = run_library(lib)
    rel = 1e-12 if double else 1e-5
    worst = {}
    for b, shape in enumerate(BOXES):
        if b % 16 == 0:
            print("checking boxes %d-%d" % (b, min(b + 15, len(BOXES) - 1)), flush=True)
        for q in range(8):
            if double:
                ok, v = fp64_verdict(ref[b][q], got[b][q], q, STEPS, DT, CONFIG_ID, shape, box_id=b,
                                     real_params=real_params_double(infos[b]))
                if v["rule"] != "acEvalError":
                    print("box %d %s %s: %s" % (b, shape, NAMES[q], v), flush=True)
                assert ok, (b, shape, NAMES[q], v)
                e = v["ulp"]
            else:
                rng = float(np.ptp(ref[b][q]))
                e = float(np.max(np.abs(got[b][q].astype(np.float64) - ref[b][q]))) / (rng if rng > 0 else 1.0)
                assert e <= 1e-5, (b, shape, NAMES[q], e)
            worst[NAMES[q]] = max(worst.get(NAMES[q], 0.0), e)
        scale = np.array([ref_diag[b][0], ref_diag[b][0], ref_diag[b][2], ref_diag[b][2],
                          abs(ref_diag[b][4]), abs(ref_diag[b][5]), abs(ref_diag[b][4])])
        d = np.abs(diag[b] - ref_diag[b]) / np.maximum(np.abs(ref_diag[b]), scale)
        assert np.all(np.isfinite(diag[b])) and d.max() <= rel, (b, shape, diag[b], ref_diag[b], d)
    print("worst per field over 96 boxes (%s):" % ("ulp at the max-abs point" if double else "range-normalised"),
          {k: "%.3g" % v for k, v in worst.items()}, flush=True)
