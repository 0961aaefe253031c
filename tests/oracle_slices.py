"""The oracle on a slice of a periodic box (test helper, CPU).

A cell after S RK3 steps depends only on the cells within 3 x 3 x S of it (one 3-cell stencil
reach per substep). cone_slice recomputes a slice of the uniform-field box from the block of
initial cells around it, generated on the host with its periodic images, over 3 S substeps of the
oracle's two-buffer form (integration.cuh:14-37) whose regions shrink by 3 cells per side; the
block's outer 3 layers stand in for ghost cells.
"""
import numpy as np

import pyoracle
from astaroth_amd.api import Library, real_params_double
from astaroth_amd.initial import mhd_block


def cone_slice(n, steps, dt, origin, sw, dtype, config_id, round_to=None):
    """The oracle's sw^3 slice of interior cells [origin, origin + sw) (wrapped periodically) of the
    uniform-field n^3 box after `steps` steps, in `dtype`; the initial cells are optionally rounded
    to `round_to` first, as the library's initial state is. Returns 8 FP64 arrays (z, y, x)."""
    reach = 9 * steps
    rp = real_params_double(Library(True).config(n=n))
    bo = tuple(o - reach - 3 for o in origin)
    bw = sw + 2 * reach + 6
    blk = mhd_block(n, n, n, bo, (bw, bw, bw), np.float64, config_id=config_id)
    if round_to is not None:
        blk = [f.astype(round_to) for f in blk]
    a = [np.ascontiguousarray(f, dtype=dtype) for f in blk]
    b = [np.zeros_like(f) for f in a]
    lo = 3 + reach
    for j in range(3 * steps):
        m = 3 * (3 * steps - 1 - j)
        pyoracle.substep_gpu_form(a, b, rp, j % 3, dt, (lo - m,) * 3, (lo + sw + m,) * 3)
        a, b = b, a
    return [f[lo:lo + sw, lo:lo + sw, lo:lo + sw].astype(np.float64) for f in a]
