"""The 512^3 precision study against the oracle in long double, on a small slice.

The periodic 512^3 box in the perturbed state (uniform fields, u_x += 1e-2 sin 2z) is advanced
10 RK3 steps at dt = 1e-3 by the FP64 and by the FP32 library (acDeviceIntegrateStep). The oracle
cannot step the whole box in long double within a test, but a cell after 10 steps depends only on
the cells within 3 x 30 = 90 of it (one 3-cell stencil reach per substep). So a 4^3 slice at the
box centre is computed in long double from the 190^3 block around it, read back from the device
before the first step: each of the 30 substeps runs the oracle's two-buffer GPU form
(integration.cuh:14-37) over the region the later substeps still need, 3 cells smaller on every
side than the one before, and the block's outer cells stand in for the ghost cells.

Bars (range-normalised over the whole box, at the slice's worst cell):
* FP64 library: <= 1e-12 (rounding of 30 substeps in FP64), and the reference's acEvalError
  (verification.c:78-102: under 5 ulps at the point of maximum absolute error, or under
  min |value| x eps) with the long-double slice as the model. The reference applies it after one
  integration; where 30 substeps of FP64 rounding take a field past 5 ulps, the field passes if it
  lies within twice the distance of the FP64 oracle's own slice (the same block, 30 substeps in
  double) from the long-double one, plus one ulp (as oracle_slices.fp64_verdict decides);
* FP32 library: <= 1e-5 (the precision study's target), plain and with compensated
  integration (acDeviceSetCompensatedIntegration: FP32 state plus its rounding-error term).
"""
import ctypes
import sys
import os

import numpy as np
import pytest

from astaroth_amd import abi
from astaroth_amd.abi import int3
from astaroth_amd.api import HostMesh, Library, real_params_double
from astaroth_amd.initial import device_init_mhd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import pyoracle  # noqa: E402

pytestmark = pytest.mark.gpu

N, STEPS, DT = 512, 10, 1e-3
REACH = 3 * 3 * STEPS            # cells a substep's stencil reaches, times 30 substeps
S0, SW = 254, 4                  # slice: interior cells [S0, S0 + SW) on every axis
B0 = S0 + 3 - REACH - 3          # block start (buffer index) including 3 stand-in ghost layers
BW = SW + 2 * REACH + 6          # block width
assert B0 >= 0 and B0 + BW <= N + 6  # the block lies inside the box (no periodic wrap needed)


def planes(lib, dev, info, z0, nz):
    """Buffer planes [z0, z0 + nz) of the device's `in` set, as (8, nz, my, mx) numpy arrays."""
    mesh = HostMesh(lib, info)  # untouched pages stay unallocated
    mxy = (N + 6) * (N + 6)
    r = lib.dll.acDeviceStoreMeshWithOffset(dev.handle, abi.STREAM_DEFAULT, int3(0, 0, z0),
                                            int3(0, 0, z0), nz * mxy, ctypes.byref(mesh.struct))
    assert r == abi.AC_SUCCESS
    dev.synchronize()
    return [f[z0:z0 + nz].copy() for f in mesh.fields]


def run_library(double, compensated=False):
    lib = Library(double)
    info = lib.config(n=N)
    with lib.device_create(0, info, autotune=False) as dev:
        if compensated:
            dev.set_compensated_integration(True)
        device_init_mhd(dev, config_id=3, perturb_sin2z=True)
        block = [p[:, B0:B0 + BW, B0:B0 + BW] for p in planes(lib, dev, info, B0, BW)] if double else None
        for _ in range(STEPS):
            dev.integrate_step(DT)
        dev.synchronize()
        sl = [p[:, S0 + 3:S0 + 3 + SW, S0 + 3:S0 + 3 + SW] for p in planes(lib, dev, info, S0 + 3, SW)]
        rng = [dev.reduce_scal("RTYPE_MAX", q, double=True) - dev.reduce_scal("RTYPE_MIN", q, double=True)
               for q in abi.VTXBUF_NAMES]
    return info, block, sl, rng


@pytest.fixture(scope="module")
def results():
    info, block, sl64, rng = run_library(True)
    print("FP64 library done", flush=True)
    _, _, sl32, _ = run_library(False)
    _, _, sl32c, _ = run_library(False, compensated=True)
    print("FP32 libraries done, long double oracle on the block", flush=True)
    # the oracle on the block, in long double and in double: 30 substeps over shrinking regions
    pyoracle.set_threads(16)
    rp = real_params_double(info)
    lo = 3 + REACH  # local index of the slice start

    def oracle_slice(dtype):
        a = [np.ascontiguousarray(f, dtype=dtype) for f in block]
        b = [np.zeros_like(f) for f in a]
        for j in range(3 * STEPS):
            m = 3 * (3 * STEPS - 1 - j)  # how far beyond the slice this substep must still reach
            start, end = (lo - m,) * 3, (lo + SW + m,) * 3
            pyoracle.substep_gpu_form(a, b, rp, j % 3, DT, start, end)
            a, b = b, a
        return [np.ascontiguousarray(f[lo:lo + SW, lo:lo + SW, lo:lo + SW]) for f in a]

    ref_ld = oracle_slice(np.longdouble)
    ref64 = oracle_slice(np.float64)
    ref = [f.astype(np.float64) for f in ref_ld]
    return ref, sl64, sl32, sl32c, rng, (ref_ld, ref64)


def errors(ref, got, rng):
    out = {}
    for q, name in enumerate(abi.VTXBUF_NAMES):
        d = np.abs(got[q].astype(np.float64) - ref[q])
        out[name] = float(d.max()) / (rng[q] if rng[q] > 0 else 1.0)
    return out


def test_fp64_slice_against_long_double(results):
    ref, sl64, _, _, rng, _ = results
    err = errors(ref, sl64, rng)
    print("FP64 vs long double, range-normalised:", {k: "%.2e" % v for k, v in err.items()})
    assert all(np.isfinite(v) for v in err.values())
    assert max(err.values()) <= 1e-12, err


def test_fp64_slice_acevalerror_against_long_double(results):
    _, sl64, _, _, _, (ref_ld, ref64) = results
    bad = {}
    for q, name in enumerate(abi.VTXBUF_NAMES):
        v = pyoracle.verify_field(ref_ld[q], sl64[q], 53, np.finfo(np.float64).eps)
        ok, rule = v["ok"], "acEvalError"
        if not ok:  # distances from the long-double slice, in ulps of its largest value
            ulp = float(np.spacing(np.float64(np.max(np.abs(ref_ld[q])))))
            gpu_err = float(np.max(np.abs(sl64[q].astype(np.longdouble) - ref_ld[q])))
            ora_err = float(np.max(np.abs(ref64[q].astype(np.longdouble) - ref_ld[q])))
            ok = gpu_err <= max(5 * ulp, 2 * ora_err + ulp)
            rule = "tie-break: GPU %.2f ulp, FP64 oracle %.2f ulp from long double" % (gpu_err / ulp, ora_err / ulp)
        print("FP64 %-16s acEvalError vs long double: %.2f ulp, abs %.2e, %s %s"
              % (name, v["ulp"], v["abs"], rule, "OK" if ok else "FAIL"), flush=True)
        if not ok:
            bad[name] = dict(v, rule=rule)
    assert not bad, bad


def test_fp32_slice_against_long_double(results):
    ref, _, sl32, _, rng, _ = results
    err = errors(ref, sl32, rng)
    print("FP32 vs long double, range-normalised:", {k: "%.2e" % v for k, v in err.items()})
    assert max(err.values()) <= 1e-5, err


def test_fp32_compensated_slice_against_long_double(results):
    ref, _, _, sl32c, rng, _ = results
    err = errors(ref, sl32c, rng)
    print("FP32 compensated vs long double, range-normalised:", {k: "%.2e" % v for k, v in err.items()})
    assert all(np.isfinite(v) for v in err.values())
    assert max(err.values()) <= 1e-5, err
