// Halo pack / unpack: copies between blocks of the eight `in` fields and a contiguous buffer.
//
// Same layout as the reference's kernel_pack_data / kernel_unpack_data
// (src/core/kernels/packing.cuh:3-92): a block of `dims` cells at `vba_start` is stored field
// after field, each field x fastest ([f][z][y][x]). The reference launches one 3-D grid per
// block; here one launch covers any number of blocks (the 26 halo blocks of the Grid layer,
// src/core/device.cc:1375-1381 and 1758-1818, in one go): a thread owns one cell and moves all
// eight fields, and threads are laid out block after block, x fastest, so reads and writes of
// the packed buffer are contiguous and the mesh side walks x rows.
#include "ac_internal.h"

namespace ac {

namespace {

constexpr int MAX_PACK_BLOCKS = 26;

struct PackBlock {
    int3 origin;  // first cell of the block in the mesh (local m-coordinates)
    int3 dims;    // block extent
    long first;   // index of the block's first cell among all blocks of the launch
    long offset;  // first element of the block in the packed buffer
    long cells;   // dims.x * dims.y * dims.z (the per-field stride inside the block)
};

struct PackArgs {
    AcReal* f[NF];
    AcReal* packed;
    int mx;
    long mxy;
    int nblocks;
    long total;
    PackBlock b[MAX_PACK_BLOCKS];
};

template <bool PACK>
__global__ void __launch_bounds__(256)
pack_kernel(const PackArgs a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.total)
        return;
    int r = 0;
    for (int q = 1; q < a.nblocks; ++q)
        r = t >= a.b[q].first ? q : r;
    const PackBlock& b = a.b[r];
    // 32-bit arithmetic within a block (fewer than 2^32 cells, checked at launch): a 64-bit
    // division costs several times a 32-bit one
    const unsigned local = (unsigned)(t - b.first);
    const unsigned dx = (unsigned)b.dims.x, dy = (unsigned)b.dims.y;
    const unsigned row = local / dx;
    const int x        = (int)(local - row * dx);
    const unsigned zz  = row / dy;
    const int y        = (int)(row - zz * dy);
    const int z        = (int)zz;
    const long m = (long)(b.origin.x + x) + (long)(b.origin.y + y) * a.mx + (long)(b.origin.z + z) * a.mxy;
    AcReal* p    = a.packed + b.offset + local;
    // All eight loads are issued before the first store: the buffer and the fields may alias
    // as far as the compiler knows, and in program order field by field each load would wait
    // for the store before it.
    AcReal v[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f)
        v[f] = PACK ? a.f[f][m] : p[f * b.cells];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        if (PACK)
            p[f * b.cells] = v[f];
        else
            a.f[f][m] = v[f];
    }
}

} // namespace

hipError_t
launch_pack(bool pack, AcReal* const* fields, int mx, int my, AcReal* packed, int nblocks,
            const int3* origins, const int3* dims, hipStream_t stream)
{
    if (nblocks < 1 || nblocks > MAX_PACK_BLOCKS)
        return hipErrorInvalidValue;
    PackArgs a;
    for (int f = 0; f < NF; ++f)
        a.f[f] = fields[f];
    a.packed  = packed;
    a.mx      = mx;
    a.mxy     = (long)mx * my;
    a.nblocks = nblocks;
    long first = 0, offset = 0;
    for (int i = 0; i < nblocks; ++i) {
        PackBlock& b = a.b[i];
        b.origin     = origins[i];
        b.dims       = dims[i];
        b.cells      = (long)dims[i].x * dims[i].y * dims[i].z;
        if (b.cells >= 0xFFFFFFFFL)
            return hipErrorInvalidValue;
        b.first      = first;
        b.offset     = offset;
        first += b.cells;
        offset += NF * b.cells;
    }
    for (int i = nblocks; i < MAX_PACK_BLOCKS; ++i)
        a.b[i] = a.b[nblocks - 1];
    a.total = first;
    if (first == 0)
        return hipSuccess;
    const long blocks = (first + 255) / 256;
    if (blocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    if (pack)
        pack_kernel<true><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    else
        pack_kernel<false><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

// The Grid layer's 26 halo blocks of a box of n interior cells (device.cc:1375-1381 for the
// extents, 1758-1818 for the ghost-side origins b0): 8 corners, 4 x-edges, 4 y-edges,
// 4 z-edges, 2 xy-sides, 2 xz-sides, 2 yz-sides, in that order. The interior block whose
// periodic image each ghost block is starts at a0 = ((b0 - 3) mod n) + 3 (device.cc:1102-1116).
int
halo_blocks(int3 n, int3* ghost_origin, int3* interior_origin, int3* dims)
{
    const int G = NGHOST;
    int k       = 0;
    auto add    = [&](int3 b0, int3 d) {
        ghost_origin[k] = b0;
        dims[k]         = d;
        auto wrap       = [](int b, int nn) { return ((b - NGHOST) % nn + nn) % nn + NGHOST; };
        interior_origin[k] = make_int3(wrap(b0.x, n.x), wrap(b0.y, n.y), wrap(b0.z, n.z));
        ++k;
    };
    const int3 corner = make_int3(G, G, G);
    add(make_int3(0, 0, 0), corner);
    add(make_int3(G + n.x, 0, 0), corner);
    add(make_int3(0, G + n.y, 0), corner);
    add(make_int3(0, 0, G + n.z), corner);
    add(make_int3(G + n.x, G + n.y, 0), corner);
    add(make_int3(G + n.x, 0, G + n.z), corner);
    add(make_int3(0, G + n.y, G + n.z), corner);
    add(make_int3(G + n.x, G + n.y, G + n.z), corner);
    const int3 ex = make_int3(n.x, G, G);
    add(make_int3(G, 0, 0), ex);
    add(make_int3(G, G + n.y, 0), ex);
    add(make_int3(G, 0, G + n.z), ex);
    add(make_int3(G, G + n.y, G + n.z), ex);
    const int3 ey = make_int3(G, n.y, G);
    add(make_int3(0, G, 0), ey);
    add(make_int3(G + n.x, G, 0), ey);
    add(make_int3(0, G, G + n.z), ey);
    add(make_int3(G + n.x, G, G + n.z), ey);
    const int3 ez = make_int3(G, G, n.z);
    add(make_int3(0, 0, G), ez);
    add(make_int3(G + n.x, 0, G), ez);
    add(make_int3(0, G + n.y, G), ez);
    add(make_int3(G + n.x, G + n.y, G), ez);
    const int3 sxy = make_int3(n.x, n.y, G);
    add(make_int3(G, G, 0), sxy);
    add(make_int3(G, G, G + n.z), sxy);
    const int3 sxz = make_int3(n.x, G, n.z);
    add(make_int3(G, 0, G), sxz);
    add(make_int3(G, G + n.y, G), sxz);
    const int3 syz = make_int3(G, n.y, n.z);
    add(make_int3(0, G, G), syz);
    add(make_int3(G + n.x, G, G), syz);
    return k;
}

} // namespace ac
