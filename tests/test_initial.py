"""Host generators of the synthetic initial states (CPU): a block cut from a periodic box by
mhd_block equals the same cells of mhd_fields, including blocks that wrap around the box; the
96-box ensemble set is fixed by its seed."""
import numpy as np

from astaroth_amd.initial import ensemble96_boxes, mhd_block, mhd_fields


def test_block_equals_whole_box_cut():
    nx, ny, nz = 11, 7, 9
    whole = mhd_fields(nx, ny, nz, np.float64, config_id=5, box_id=2)
    tiled = [np.tile(f[3:-3, 3:-3, 3:-3], (3, 3, 3)) for f in whole]  # periodic images
    for origin, width in (((0, 0, 0), (nx, ny, nz)), ((-4, -6, -2), (20, 15, 13)),
                          ((8, 5, 7), (9, 4, 6))):
        blk = mhd_block(nx, ny, nz, origin, width, np.float64, config_id=5, box_id=2)
        o = [origin[a] + (nx, ny, nz)[a] for a in range(3)]
        for q in range(8):
            exp = tiled[q][o[2]:o[2] + width[2], o[1]:o[1] + width[1], o[0]:o[0] + width[0]]
            assert np.array_equal(blk[q], exp), (origin, q)


def test_ensemble96_box_set():
    boxes, nus = ensemble96_boxes()
    assert len(boxes) == 96 and len(nus) == 96
    assert all(min(b) >= 32 and max(b) <= 128 for b in boxes)
    assert all(1e-3 <= nu <= 1e-2 for nu in nus)
    assert all(b[0] == b[1] == b[2] for b in boxes[::2])
    assert sum(x * y * z for x, y, z in boxes) == 61505536


def test_cone_slice_wrapping_a_corner_equals_whole_box():
    """The slice method of the full-size GPU tests, on a slice that wraps around a box corner:
    bit for bit the cells of the whole periodic box stepped by the oracle."""
    import pyoracle
    from oracle_slices import cone_slice
    from astaroth_amd.api import Library, real_params_double

    n, steps, dt, sw = 24, 2, 1e-3, 4
    whole = mhd_fields(n, n, n, np.float64, config_id=2)
    pyoracle.integrate_steps_gpu_form(whole, real_params_double(Library(True).config(n=n)), dt, steps)
    for origin in ((n - 2, n - 2, n - 2), (5, n - 3, 0)):
        got = cone_slice(n, steps, dt, origin, sw, np.float64, config_id=2)
        idx = [np.array([(o + t) % n for t in range(sw)]) + 3 for o in origin]
        for q in range(8):
            exp = whole[q][np.ix_(idx[2], idx[1], idx[0])]
            assert np.array_equal(got[q], exp), (origin, q)
