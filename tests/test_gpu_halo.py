"""Halo pack / unpack through the C ABI (reference src/core/kernels/packing.cuh:3-92 and the Grid
layer's 26 halo blocks, src/core/device.cc:1375-1381, 1758-1818).

All checks are bitwise: packing is a permutation of stored values.
* acDevicePackData of a block equals numpy slicing of the loaded mesh, laid out [f][z][y][x].
* acDeviceUnpackData writes exactly the block and nothing else.
* acDevicePackHalo + acDeviceUnpackHalo (and acDeviceHaloExchangeSelf) fill every ghost cell
  exactly as the periodic fill does, i.e. as np.pad(interior, 3, mode="wrap").
* The packed halo buffer holds the 26 blocks in the reference's CommData order.
"""
import numpy as np
import pytest

from astaroth_amd.api import Library, HostMesh
from astaroth_amd.hiprt import Hip

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[False, True], ids=["fp32", "fp64"])
def lib(request):
    return Library(request.param)


@pytest.fixture(scope="module")
def hip():
    return Hip()


def random_mesh(lib, info, n, seed):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    fields = [rng.standard_normal((nz + 6, ny + 6, nx + 6)).astype(lib.dtype) for _ in range(8)]
    return HostMesh(lib, info, fields)


def reference_halo_blocks(n):
    """The Grid layer's ghost-side origins b0 and extents, in its CommData order."""
    nx, ny, nz = n
    G = 3
    c = (G, G, G)
    blocks = [((0, 0, 0), c), ((G + nx, 0, 0), c), ((0, G + ny, 0), c), ((0, 0, G + nz), c),
              ((G + nx, G + ny, 0), c), ((G + nx, 0, G + nz), c), ((0, G + ny, G + nz), c),
              ((G + nx, G + ny, G + nz), c)]
    blocks += [((G, 0, 0), (nx, G, G)), ((G, G + ny, 0), (nx, G, G)), ((G, 0, G + nz), (nx, G, G)),
               ((G, G + ny, G + nz), (nx, G, G))]
    blocks += [((0, G, 0), (G, ny, G)), ((G + nx, G, 0), (G, ny, G)), ((0, G, G + nz), (G, ny, G)),
               ((G + nx, G, G + nz), (G, ny, G))]
    blocks += [((0, 0, G), (G, G, nz)), ((G + nx, 0, G), (G, G, nz)), ((0, G + ny, G), (G, G, nz)),
               ((G + nx, G + ny, G), (G, G, nz))]
    blocks += [((G, G, 0), (nx, ny, G)), ((G, G, G + nz), (nx, ny, G))]
    blocks += [((G, 0, G), (nx, G, nz)), ((G, G + ny, G), (nx, G, nz))]
    blocks += [((0, G, G), (G, ny, nz)), ((G + nx, G, G), (G, ny, nz))]
    return blocks


def block_of(fields, start, dims):
    x0, y0, z0 = start
    dx, dy, dz = dims
    return np.concatenate([f[z0:z0 + dz, y0:y0 + dy, x0:x0 + dx].ravel() for f in fields])


@pytest.mark.parametrize("n,start,dims", [((24, 16, 8), (3, 3, 3), (24, 16, 8)),
                                          ((24, 16, 8), (0, 5, 2), (7, 3, 9)),
                                          ((13, 11, 9), (16, 0, 0), (3, 17, 15))])
def test_pack_unpack_block(lib, hip, n, start, dims):
    info = lib.config(n=n[0], ny=n[1], nz=n[2])
    mesh = random_mesh(lib, info, n, 7)
    count = 8 * dims[0] * dims[1] * dims[2]
    nbytes = count * np.dtype(lib.dtype).itemsize
    buf = hip.malloc(nbytes)
    try:
        with lib.device_create(0, info, autotune=False) as dev:
            dev.load_mesh(mesh)
            dev.pack_data(start, dims, buf)
            dev.synchronize()
            packed = hip.to_host(buf, count, lib.dtype)
            assert np.array_equal(packed, block_of(mesh.fields, start, dims))
            # unpack different values into the same block: only the block changes
            new = (np.arange(count, dtype=np.float64) * 0.25 - 7).astype(lib.dtype)
            hip.to_device(buf, new)
            dev.unpack_data(buf, start, dims)
            out = HostMesh(lib, info)
            dev.store_mesh(out)
        x0, y0, z0 = start
        dx, dy, dz = dims
        cells = dx * dy * dz
        for q in range(8):
            exp = mesh.fields[q].copy()
            exp[z0:z0 + dz, y0:y0 + dy, x0:x0 + dx] = new[q * cells:(q + 1) * cells].reshape(dz, dy, dx)
            assert np.array_equal(out.fields[q], exp)
    finally:
        hip.free(buf)


def test_pack_rejects_blocks_outside_the_box(lib, hip):
    info = lib.config(n=8)
    buf = hip.malloc(8 * 4 * 4 * 4 * 8)
    try:
        with lib.device_create(0, info, autotune=False) as dev:
            assert lib.dll.acDevicePackData(dev.handle, 0, _i3(12, 0, 0), _i3(4, 4, 4), _vp(buf)) != 0
            assert lib.dll.acDeviceUnpackData(dev.handle, 0, _vp(buf), _i3(-1, 0, 0), _i3(4, 4, 4)) != 0
            assert lib.dll.acDevicePackData(dev.handle, 0, _i3(0, 0, 0), _i3(4, 4, 4), _vp(buf)) == 0
            dev.synchronize()
    finally:
        hip.free(buf)


def _i3(x, y, z):
    from astaroth_amd.abi import int3
    return int3(x, y, z)


def _vp(p):
    import ctypes
    return ctypes.c_void_p(p)


@pytest.mark.parametrize("n", [(16, 16, 16), (20, 12, 7), (9, 13, 11)])
def test_halo_pack_layout_and_self_exchange(lib, hip, n):
    info = lib.config(n=n[0], ny=n[1], nz=n[2])
    mesh = random_mesh(lib, info, n, 11)
    blocks = reference_halo_blocks(n)
    total = sum(8 * d[0] * d[1] * d[2] for _, d in blocks)
    buf = hip.malloc(total * np.dtype(lib.dtype).itemsize)
    try:
        with lib.device_create(0, info, autotune=False) as dev:
            assert dev.halo_packed_size() == total
            dev.load_mesh(mesh)
            dev.pack_halo(buf)
            dev.synchronize()
            packed = hip.to_host(buf, total, lib.dtype)
            # each block holds the periodic interior image of its ghost block
            wrapped = [np.pad(f[3:-3, 3:-3, 3:-3], 3, mode="wrap") for f in mesh.fields]
            off = 0
            for b0, d in blocks:
                cnt = 8 * d[0] * d[1] * d[2]
                assert np.array_equal(packed[off:off + cnt], block_of(wrapped, b0, d)), (b0, d)
                off += cnt
            dev.unpack_halo(buf)
            out = HostMesh(lib, info)
            dev.store_mesh(out)
            for q in range(8):
                assert np.array_equal(out.fields[q], wrapped[q])
            # the self exchange through the device-owned buffer, twice (buffer reuse)
            dev.load_mesh(mesh)
            dev.halo_exchange_self()
            dev.halo_exchange_self()
            out2 = HostMesh(lib, info)
            dev.store_mesh(out2)
            for q in range(8):
                assert np.array_equal(out2.fields[q], wrapped[q])
    finally:
        hip.free(buf)
