"""The multi-step FP64 verdict of the GPU tests (tests/oracle_slices.fp64_verdict) on the CPU.

* A legitimate second FP64 computation passes: the host model solver's own form
  (acHostIntegrateStep, modelsolver.c:987-1041: w stored explicitly) judged against the two-buffer
  form of the GPU path (integration.cuh:14-37) after 3 steps, every field.
* A planted error does not: the same result with 40 ulps added at the largest cell of a field
  fails acEvalError and the long-double tie-break.
* The long-double cone slice the tie-break uses equals the long-double oracle on the whole
  periodic box (an anisotropic box, a slice wrapping a corner).
"""
import numpy as np

import pyoracle
from astaroth_amd.api import Library, real_params_double
from astaroth_amd.initial import mhd_fields
from oracle_slices import cone_slice, fp64_verdict, window

DIMS, STEPS, DT, CONFIG_ID = (20, 16, 12), 3, 1e-3, 2


def run(form, dtype=np.float64):
    nx, ny, nz = DIMS
    f = mhd_fields(nx, ny, nz, dtype, config_id=CONFIG_ID)
    rp = real_params_double(Library(True).config(n=nx, ny=ny, nz=nz))
    (pyoracle.integrate_steps_gpu_form if form == "gpu" else pyoracle.integrate_steps)(f, rp, DT, STEPS)
    return [x[3:-3, 3:-3, 3:-3].copy() for x in f]


def test_host_form_passes_and_planted_error_fails():
    pyoracle.set_threads(8)
    ref, host = run("gpu"), run("host")
    for q in range(8):
        ok, v = fp64_verdict(ref[q], host[q], q, STEPS, DT, CONFIG_ID, DIMS)
        assert ok, (q, v)
    q = 1
    bad = host[q].copy()
    k = np.unravel_index(int(np.argmax(np.abs(bad))), bad.shape)
    bad[k] += 40 * np.spacing(abs(bad[k]))
    ok, v = fp64_verdict(ref[q], bad, q, STEPS, DT, CONFIG_ID, DIMS)
    assert not ok and v["rule"] == "long-double tie-break", v


def test_long_double_cone_slice_equals_whole_box():
    pyoracle.set_threads(8)
    whole = run("gpu", np.longdouble)
    nx, ny, nz = DIMS
    centre = (nx - 1, 0, nz - 2)  # the 3^3 window wraps in x, y and z
    sl = cone_slice(None, STEPS, DT, tuple(c - 1 for c in centre), 3, np.longdouble, CONFIG_ID, dims=DIMS)
    for q in range(8):
        w = whole[q]
        nzw, nyw, nxw = w.shape
        ix = (np.arange(3) + centre[0] - 1) % nxw
        iy = (np.arange(3) + centre[1] - 1) % nyw
        iz = (np.arange(3) + centre[2] - 1) % nzw
        assert np.array_equal(sl[q], w[np.ix_(iz, iy, ix)]), q
        assert np.array_equal(window(whole[q].astype(np.float64), centre, 3), w[np.ix_(iz, iy, ix)].astype(np.float64))


def test_verify_field_fast_path_equals_long_double_path():
    """verify_field finds the max-abs-error point on FP64 differences and re-checks the near
    maxima in long double; the result equals the all-long-double evaluation."""
    rng = np.random.default_rng(3)
    model = rng.uniform(-1e-3, 1e-3, 20000)
    for cand in (model + rng.integers(-6, 7, model.size) * np.spacing(np.abs(model)),
                 model.astype(np.float32).astype(np.float64),
                 model * (1 + 1e-9 * rng.standard_normal(model.size))):
        for c in (cand, cand.astype(np.float32)) if cand.dtype == np.float64 else (cand,):
            fast = pyoracle.verify_field(model, c, 53, np.finfo(np.float64).eps)
            slow = pyoracle.verify_field(model.astype(np.longdouble), c.astype(np.longdouble), 53,
                                         np.finfo(np.float64).eps)
            assert fast == slow, (fast, slow)
