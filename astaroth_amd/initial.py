"""Synthetic initial conditions of the benchmark configurations.

The uniform generator is the counter-based splitmix64 stream that the library's
acDeviceRandomizeUniform evaluates on the GPU (csrc/init.hip); `uniform_interior` reproduces it
bit for bit on the host so small cases can be generated on either side. Seeds are derived from
(config id, box id, field id) and are recorded with every result.

The uniform-field distribution: lnrho ~ U(-1e-2, 1e-2); u and A components ~ U(-1e-3, 1e-3);
entropy = 0 (the perturbed state adds u_x += 1e-2 sin(2 z)).
"""
from __future__ import annotations

import numpy as np

from .abi import IP, VTX

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

FIELD_RANGES = {
    "VTXBUF_LNRHO": (-1e-2, 1e-2),
    "VTXBUF_UUX": (-1e-3, 1e-3),
    "VTXBUF_UUY": (-1e-3, 1e-3),
    "VTXBUF_UUZ": (-1e-3, 1e-3),
    "VTXBUF_AX": (-1e-3, 1e-3),
    "VTXBUF_AY": (-1e-3, 1e-3),
    "VTXBUF_AZ": (-1e-3, 1e-3),
    "VTXBUF_ENTROPY": (0.0, 0.0),
}


def seed_for(config_id: int, box_id: int, field_id: int) -> int:
    return (config_id * 1000003 + box_id * 1009 + field_id) & 0xFFFFFFFFFFFFFFFF


def uniform_interior(nx, ny, nz, lo, hi, seed) -> np.ndarray:
    """lo + (hi - lo) u over the interior, shape (nz, ny, nx), FP64."""
    with np.errstate(over="ignore"):
        t = np.arange(1, nx * ny * nz + 1, dtype=np.uint64)
        z = np.uint64(seed) + t * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(nz, ny, nx)


def uniform_cells(nx, ny, nz, lo, hi, seed, i, j, k) -> np.ndarray:
    """The generator's value at interior cells (i, j, k) (broadcast integer arrays, taken modulo
    the box: the periodic image of any cell), FP64."""
    i, j, k = (np.asarray(v, dtype=np.int64) % n for v, n in ((i, nx), (j, ny), (k, nz)))
    with np.errstate(over="ignore"):
        t = (i + nx * (j + ny * k)).astype(np.uint64) + np.uint64(1)
        z = np.uint64(seed) + t * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def mhd_block(nx, ny, nz, origin, width, dtype=np.float64, config_id=1, box_id=0):
    """Eight arrays of shape (width_z, width_y, width_x) holding the uniform-field state of an
    nx x ny x nz periodic box on the block of interior cells [origin, origin + width) (any
    integers: cells outside the interior are their periodic images). Equal, cell for cell, to
    the same block cut from mhd_fields, without forming the whole box."""
    (x0, y0, z0), (wx, wy, wz) = origin, width
    k = np.arange(z0, z0 + wz)[:, None, None]
    j = np.arange(y0, y0 + wy)[None, :, None]
    i = np.arange(x0, x0 + wx)[None, None, :]
    out = []
    for name, (lo, hi) in FIELD_RANGES.items():
        q = VTX[name]
        if hi == lo:
            f = np.full((wz, wy, wx), lo, dtype=np.float64)
        else:
            f = uniform_cells(nx, ny, nz, lo, hi, seed_for(config_id, box_id, q), i, j, k)
        out.append(f.astype(dtype))
    return out


def ensemble96_boxes(count=96, seed=4):
    """Box edges and viscosities of the 96-box ensemble: edges drawn from {32, 48, 64, 96, 128},
    every even box cubic and every odd box anisotropic, nu ~ U[1e-3, 1e-2] per box."""
    rng = np.random.default_rng(seed)
    edges = np.array([32, 48, 64, 96, 128])
    boxes, nus = [], []
    for b in range(count):
        if b % 2 == 0:  # cubic
            e = int(rng.choice(edges))
            boxes.append((e, e, e))
        else:  # anisotropic
            boxes.append(tuple(int(v) for v in rng.choice(edges, 3)))
        nus.append(float(rng.uniform(1e-3, 1e-2)))
    return boxes, nus


def periodic_fill(f: np.ndarray) -> np.ndarray:
    """Ghost cells of one (mz, my, mx) array from the wrapped interior (in place)."""
    f[...] = np.pad(f[3:-3, 3:-3, 3:-3], 3, mode="wrap")
    return f


def mhd_fields(nx, ny, nz, dtype=np.float64, config_id=1, box_id=0, perturb_sin2z=False):
    """Eight (mz, my, mx) arrays of the uniform-field distribution, ghosts filled periodically."""
    out = []
    for name, (lo, hi) in FIELD_RANGES.items():
        q = VTX[name]
        f = np.zeros((nz + 6, ny + 6, nx + 6), dtype=np.float64)
        if hi != lo:
            f[3:-3, 3:-3, 3:-3] = uniform_interior(nx, ny, nz, lo, hi, seed_for(config_id, box_id, q))
        if perturb_sin2z and name == "VTXBUF_UUX":
            dsz = 2 * np.pi / nz
            zc = (np.arange(nz) * dsz)[:, None, None]
            f[3:-3, 3:-3, 3:-3] += 1e-2 * np.sin(2 * zc)
        out.append(periodic_fill(f).astype(dtype))
    return out


def device_init_mhd(dev, config_id=1, box_id=0, perturb_sin2z=False):
    """The same distribution generated on the GPU (acDeviceRandomizeUniform + periodic fill);
    with perturb_sin2z, u_x also gets + 1e-2 sin(2 z) (acDeviceRandomizeUniformSinZ)."""
    for name, (lo, hi) in FIELD_RANGES.items():
        q = VTX[name]
        if perturb_sin2z and name == "VTXBUF_UUX":
            nz = dev.info.int_params[IP["AC_nz"]]
            dsz = 2 * np.pi / nz  # the grid spacing of mhd_fields, in FP64
            dev.randomize_uniform_sinz(q, lo, hi, seed_for(config_id, box_id, q), 1e-2, 2 * dsz)
        elif hi == lo:
            dev.set_vertex_buffer(q, lo)
        else:
            dev.randomize_uniform(q, lo, hi, seed_for(config_id, box_id, q))
    dev.periodic_boundconds()
    dev.synchronize()


def ensemble_init_mhd(ens, box, config_id=4, box_id=None):
    """device_init_mhd for one box of an Ensemble (seeds from (config_id, box_id, field));
    call ens.periodic_boundconds() once all boxes are set."""
    box_id = box if box_id is None else box_id
    for name, (lo, hi) in FIELD_RANGES.items():
        q = VTX[name]
        if hi == lo:
            ens.set_vertex_buffer(box, q, lo)
        else:
            ens.randomize_uniform(box, q, lo, hi, seed_for(config_id, box_id, q))
