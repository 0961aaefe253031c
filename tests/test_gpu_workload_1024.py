"""The 1024^3 FP32 workload on the GPU: 5 RK3 steps of the uniform-field box in which every substep
takes its ghost cells from the 26-block pack + unpack (acDeviceHaloExchangeSelf: the Grid layer's
halo blocks, src/core/device.cc:1375-1381, 1758-1818, exchanged by a box that is its own periodic
neighbour) instead of the periodic fill. Each field buffer holds 1030^3 cells (4.37 GB in FP32),
so buffer planes beyond z = 1012 sit past 4 GiB; the box is 64 tiles wide, so the solve orders its tiles
in strips.

Checks:
* after the last step the ghost cells are stale; a further exchange makes them equal, bit for bit,
  to what the periodic fill writes, on 16 whole planes (the six z-ghost planes, the interior
  planes they mirror, and planes through the box centre);
* on those planes the exchanged ghost cells are the periodic images of the interior
  (numpy.pad(mode="wrap") of each plane's interior and the z planes they copy);
* a 4^3 slice at the box centre and one at the box corner (wrapping across every face, and at
  the far end of every buffer) equal the oracle's, computed in FP64 from the block of initial
  cells each depends on (one 3-cell stencil reach per substep: 45 cells), range-normalised
  error <= 1e-5 per field (the FP32 build's precision target against FP64).

dt is 1e-3 scaled by (512 / 1024)^2: the explicit viscous terms limit dt ~ dx^2.
"""
import ctypes

import numpy as np
import pytest

from astaroth_amd import abi
from astaroth_amd.abi import int3
from astaroth_amd.api import HostMesh, Library
from astaroth_amd.initial import device_init_mhd
from oracle_slices import cone_slice

pytestmark = pytest.mark.gpu

N, STEPS, CONFIG_ID, SW = 1024, 5, 5, 4
DT = 1e-3 * (512.0 / N) ** 2
ORIGINS = {"centre": (510, 510, 510), "corner": (N - 2, N - 2, N - 2)}
NAMES = abi.VTXBUF_NAMES


def slice_planes():
    """Buffer z planes holding the slices' cells, plus the ghost planes and their images."""
    zs = set(range(0, 6)) | set(range(N, N + 6))
    for o in ORIGINS.values():
        zs |= {(o[2] + k) % N + 3 for k in range(SW)}
    return sorted(zs)


def read_planes(lib, dev, info, zs):
    """{z: 8 arrays (my, mx)} of the device's `in` set (a host mesh whose untouched pages stay
    unallocated)."""
    mesh = HostMesh(lib, info)
    mxy = (N + 6) * (N + 6)
    for z in zs:
        r = lib.dll.acDeviceStoreMeshWithOffset(dev.handle, abi.STREAM_DEFAULT, int3(0, 0, z), int3(0, 0, z),
                                                mxy, ctypes.byref(mesh.struct))
        assert r == abi.AC_SUCCESS
    dev.synchronize()
    out = {z: [f[z].copy() for f in mesh.fields] for z in zs}
    del mesh
    return out


def slice_of(planes, origin):
    idx = [(np.arange(SW) + o) % N + 3 for o in origin]
    return [np.stack([planes[z][q][np.ix_(idx[1], idx[0])] for z in idx[2]]).astype(np.float64)
            for q in range(8)]


@pytest.fixture(scope="module")
def run():
    lib = Library(False)
    info = lib.config(n=N)
    zs = slice_planes()
    with lib.device_create(0, info, autotune=False) as dev:
        device_init_mhd(dev, config_id=CONFIG_ID)
        for step in range(STEPS):
            for s in range(3):
                dev.halo_exchange_self()
                dev.integrate_substep(s, dt=DT)
                dev.swap_buffers()
            dev.synchronize()
            print("1024^3 step %d done" % step, flush=True)
        stale = read_planes(lib, dev, info, zs)
        dev.halo_exchange_self()
        exchanged = read_planes(lib, dev, info, zs)
        dev.periodic_boundconds()
        filled = read_planes(lib, dev, info, zs)
        rng = [dev.reduce_scal("RTYPE_MAX", q, double=True) - dev.reduce_scal("RTYPE_MIN", q, double=True)
               for q in NAMES]
    return dict(zs=zs, stale=stale, exchanged=exchanged, filled=filled, rng=rng)


def test_exchange_equals_fill_at_1024(run):
    for z in run["zs"]:
        for q in range(8):
            assert np.array_equal(run["exchanged"][z][q], run["filled"][z][q]), (z, NAMES[q])


def test_exchange_writes_periodic_images_at_1024(run):
    ex, st = run["exchanged"], run["stale"]
    for q in range(8):
        for k in range(3):  # z ghost planes copy the interior planes one box length away
            assert np.array_equal(ex[k][q][3:-3, 3:-3], ex[N + k][q][3:-3, 3:-3]), (k, NAMES[q])
            assert np.array_equal(ex[N + 3 + k][q][3:-3, 3:-3], ex[3 + k][q][3:-3, 3:-3]), (k, NAMES[q])
        for z in run["zs"]:
            p = ex[z][q]
            src = ex[(z - 3) % N + 3][q] if (z < 3 or z >= N + 3) else p
            assert np.array_equal(p, np.pad(src[3:-3, 3:-3], 3, mode="wrap")), (z, NAMES[q])
    # the exchange had work to do: the ghost cells it overwrote were stale
    changed = sum(int(np.count_nonzero(ex[0][q] != st[0][q])) for q in range(8))
    assert changed > 0


@pytest.mark.parametrize("where", list(ORIGINS))
def test_slice_at_1024_against_oracle(run, where):
    origin = ORIGINS[where]
    got = slice_of(run["filled"], origin)
    ref = cone_slice(N, STEPS, DT, origin, SW, np.float64, CONFIG_ID, round_to=np.float32)
    worst = {}
    for q in range(8):
        r = run["rng"][q]
        err = float(np.max(np.abs(got[q] - ref[q]))) / (r if r > 0 else 1.0)
        worst[NAMES[q]] = err
        assert np.all(np.isfinite(got[q])) and err <= 1e-5, (where, NAMES[q], err)
    print("%s slice, range-normalised:" % where, {k: "%.2e" % v for k, v in worst.items()}, flush=True)
