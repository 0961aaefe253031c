"""The 256^3 workload against the oracle: the uniform-field box advanced 10 RK3 steps at dt = 1e-3
by the FP32 library (with its autotuner at device creation, as the workload runs it) and by the
FP64 library, with the FP64 diagnostics (max, min and rms of |u|, |A| and lnrho, acDeviceDiagnostics)
after every step. This is the reference's own acceptance check of a caller's run -- integrate, then
acVerifyMesh and the reductions against the host model (samples/mpitest/main.cc:343-405) -- at the
workload's full size.

Oracle: the CPU restatement of the host model solver in the two-buffer RK3 form of the GPU path
(integration.cuh:14-37), FP64, over the whole periodic box; its reductions are the host
reductions (modelreduce.c:76-209).

Bars:
* FP64 library: after one step and after 10, acEvalError of the reference (verification.c:78-102)
  on every field, whole box, or, where a field misses its 5 ulps, no farther than 5 ulps from the
  long-double oracle around the worst cell (or within twice the FP64 oracle's own distance from
  it; oracle_slices.fp64_verdict); every diagnostic of every step within 1e-12 of the oracle's,
  relative to the larger of its magnitude and that quantity's maximum magnitude (the minimum of
  |u| lies near 0).
* FP32 library: range-normalised error <= 1e-5 per field of the final state (the precision
  target of the FP32 build against FP64); every diagnostic of every step within 1e-5, relative as
  above.
"""
import numpy as np
import pytest

import pyoracle
from astaroth_amd import abi
from astaroth_amd.api import HostMesh, Library, real_params_double
from astaroth_amd.initial import device_init_mhd, mhd_fields
from oracle_slices import fp64_verdict

pytestmark = pytest.mark.gpu

N, STEPS, DT, CONFIG_ID = 256, 10, 1e-3, 2
NAMES = abi.VTXBUF_NAMES
KEYS = ("uu", "aa", "lnrho")


def oracle_diagnostics(f):
    """max, min, rms (sqrt(sum / N)) of |u|, |A| and lnrho by the host reductions."""
    groups = {"uu": [f[1], f[2], f[3]], "aa": [f[4], f[5], f[6]], "lnrho": [f[0]]}
    return {k: [float(pyoracle.reduce(groups[k], r)) for r in ("RTYPE_MAX", "RTYPE_MIN", "RTYPE_RMS")]
            for k in KEYS}


@pytest.fixture(scope="module")
def oracle_run():
    pyoracle.set_threads(16)
    f = mhd_fields(N, N, N, np.float64, config_id=CONFIG_ID)
    rp = real_params_double(Library(True).config(n=N))
    diags = []
    for s in range(STEPS):
        pyoracle.integrate_steps_gpu_form(f, rp, DT, 1)
        diags.append(oracle_diagnostics(f))
        if s == 0:
            first = [x[3:-3, 3:-3, 3:-3].copy() for x in f]
        print("oracle step %d: max|u| %.17g" % (s, diags[-1]["uu"][0]), flush=True)
    return [x[3:-3, 3:-3, 3:-3].copy() for x in f], diags, first


def run_library(double, steps=STEPS):
    lib = Library(double)
    info = lib.config(n=N)
    with lib.device_create(0, info, autotune=not double) as dev:
        cfg = dev.kernel_config()
        print("%s kernel configuration %s" % ("fp64" if double else "fp32", cfg.as_tuple()), flush=True)
        device_init_mhd(dev, config_id=CONFIG_ID)
        diags = []
        for _ in range(steps):
            dev.integrate_step(DT)
            diags.append(dev.diagnostics())
        out = HostMesh(lib, info)
        dev.store_mesh(out)
        dev.synchronize()
    return [out.interior(q) for q in range(8)], diags


def check_diagnostics(ref, got, rel, label):
    worst = 0.0
    for s in range(STEPS):
        for k in KEYS:
            for i, what in enumerate(("max", "min", "rms")):
                a, b = ref[s][k][i], got[s][k][i]
                # relative to the larger of the value and the group's maximum magnitude: the
                # minimum of |u| lies near zero
                d = abs(a - b) / max(abs(a), abs(ref[s][k][0]), abs(ref[s][k][1]))
                worst = max(worst, d)
                assert np.isfinite(b) and d <= rel, (label, s, k, what, a, b, d)
    print("%s: worst diagnostic relative difference over %d steps %.3g" % (label, STEPS, worst), flush=True)


def test_fp64_256_one_step_against_oracle(oracle_run):
    """One step, the reference's own check of an integration. Over 16.7 million cells a cell at 5
    ulps is likely even after one step, so the verdict is acEvalError or the long-double
    tie-break of fp64_verdict, with the rule used printed per field."""
    ref = oracle_run[2]
    got, _ = run_library(True, steps=1)
    for q in range(8):
        ok, v = fp64_verdict(ref[q], got[q], q, 1, DT, CONFIG_ID, (N, N, N))
        print("  %-16s %6.3g ulp %s %s" % (NAMES[q], v["ulp"], v["rule"],
                                          {k: v[k] for k in ("cell", "gpu_to_exact_ulp", "oracle_to_exact_ulp")
                                           if k in v}), flush=True)
        assert ok, (NAMES[q], v)


def test_fp64_256_against_oracle(oracle_run):
    ref, ref_diag = oracle_run[0], oracle_run[1]
    got, diag = run_library(True)
    check_diagnostics(ref_diag, diag, 1e-12, "fp64")
    for q in range(8):
        ok, v = fp64_verdict(ref[q], got[q], q, STEPS, DT, CONFIG_ID, (N, N, N))
        print("  %-16s %6.3g ulp %s %s" % (NAMES[q], v["ulp"], v["rule"],
                                          {k: v[k] for k in ("cell", "gpu_to_exact_ulp", "oracle_to_exact_ulp")
                                           if k in v}), flush=True)
        assert ok, (NAMES[q], v)


def test_fp32_256_against_oracle(oracle_run):
    ref, ref_diag = oracle_run[0], oracle_run[1]
    got, diag = run_library(False)
    check_diagnostics(ref_diag, diag, 1e-5, "fp32")
    for q in range(8):
        rng = float(np.ptp(ref[q]))
        err = float(np.max(np.abs(got[q].astype(np.float64) - ref[q]))) / (rng if rng > 0 else 1.0)
        print("  %-16s range-normalised %.3g" % (NAMES[q], err), flush=True)
        assert err <= 1e-5, (NAMES[q], err)
