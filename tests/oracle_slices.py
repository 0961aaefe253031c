"""The oracle on a slice of a periodic box (test helper, CPU).

A cell after S RK3 steps depends only on the cells within 3 x 3 x S of it (one 3-cell stencil
reach per substep). cone_slice recomputes a slice of the uniform-field box from the block of
initial cells around it, generated on the host with its periodic images, over 3 S substeps of the
oracle's two-buffer form (integration.cuh:14-37) whose regions shrink by 3 cells per side; the
block's outer 3 layers stand in for ghost cells. The block may be wider than the box: its cells
are then periodic images of one another, which the substeps do not mind.

fp64_verdict applies the reference's acEvalError (verification.c:78-102) to an FP64 result after
several steps and, where a field misses it, decides on the exact solution instead.
"""
import numpy as np

import pyoracle
from astaroth_amd.api import Library, real_params_double
from astaroth_amd.initial import mhd_block


def cone_slice(n, steps, dt, origin, sw, dtype, config_id, round_to=None, dims=None, box_id=0,
               real_params=None):
    """The oracle's sw^3 slice of interior cells [origin, origin + sw) (wrapped periodically) of the
    uniform-field box (n^3, or `dims` = (nx, ny, nz)) after `steps` steps, in `dtype`; the initial
    cells are optionally rounded to `round_to` first, as the library's initial state is.
    `real_params` defaults to the default configuration of the box. Returns 8 arrays (z, y, x),
    long double when `dtype` is, else FP64."""
    nx, ny, nz = dims if dims is not None else (n, n, n)
    reach = 9 * steps
    rp = real_params if real_params is not None else real_params_double(Library(True).config(n=nx, ny=ny, nz=nz))
    bo = tuple(o - reach - 3 for o in origin)
    bw = sw + 2 * reach + 6
    blk = mhd_block(nx, ny, nz, bo, (bw, bw, bw), np.float64, config_id=config_id, box_id=box_id)
    if round_to is not None:
        blk = [f.astype(round_to) for f in blk]
    a = [np.ascontiguousarray(f, dtype=dtype) for f in blk]
    b = [np.zeros_like(f) for f in a]
    lo = 3 + reach
    for j in range(3 * steps):
        m = 3 * (3 * steps - 1 - j)
        pyoracle.substep_gpu_form(a, b, rp, j % 3, dt, (lo - m,) * 3, (lo + sw + m,) * 3)
        a, b = b, a
    out = np.longdouble if np.dtype(dtype) == np.dtype(np.longdouble) else np.float64
    return [f[lo:lo + sw, lo:lo + sw, lo:lo + sw].astype(out) for f in a]


def window(arr, centre, sw):
    """The sw^3 cells of a (z, y, x) interior array starting sw // 2 before `centre` = (i, j, k),
    wrapped periodically."""
    nz, ny, nx = arr.shape
    i, j, k = centre
    ix = (np.arange(sw) + i - sw // 2) % nx
    iy = (np.arange(sw) + j - sw // 2) % ny
    iz = (np.arange(sw) + k - sw // 2) % nz
    return arr[np.ix_(iz, iy, ix)].astype(np.float64)


def fp64_verdict(ref, got, q, steps, dt, config_id, dims, box_id=0, real_params=None, sw=3):
    """Verdict on field q of an FP64 result `got` (interior, (z, y, x)) against the FP64 oracle
    `ref` of the same form after `steps` steps.

    acEvalError as the reference applies it (5 ulps at the point of maximum absolute error, or an
    error below min |value| x eps). The reference checks one integration; over many substeps the
    two FP64 computations (fused multiply-adds and another summation order on the GPU) drift
    apart by a few more ulps. Where acEvalError fails, the distance is measured from the exact
    solution instead of from a second FP64 result: the oracle in long double on the sw^3 cells
    around the worst cell. The field passes if the result lies within acEvalError's 5 ulps of it
    (ulps of the largest value there), or within twice the FP64 oracle's own distance from it plus
    one ulp. Returns (ok, details)."""
    v = pyoracle.verify_field(ref, got, 53, np.finfo(np.float64).eps)
    if v["ok"]:
        return True, dict(v, rule="acEvalError")
    diff = np.abs(got.astype(np.float64) - ref)
    k, j, i = np.unravel_index(int(np.argmax(diff)), diff.shape)
    origin = (i - sw // 2, j - sw // 2, k - sw // 2)
    exact = cone_slice(None, steps, dt, origin, sw, np.longdouble, config_id, dims=dims, box_id=box_id,
                       real_params=real_params)[q]
    g, r = window(got, (i, j, k), sw), window(ref, (i, j, k), sw)
    ulp = float(np.spacing(np.float64(np.max(np.abs(exact)))))
    gpu_err = float(np.max(np.abs(g.astype(np.longdouble) - exact)))
    ref_err = float(np.max(np.abs(r.astype(np.longdouble) - exact)))
    ok = gpu_err <= max(5 * ulp, 2 * ref_err + ulp)
    return ok, dict(v, rule="long-double tie-break", cell=(int(i), int(j), int(k)), gpu_to_exact_ulp=gpu_err / ulp,
                    oracle_to_exact_ulp=ref_err / ulp)
