// Periodic ghost-cell fill.
//
// Same result as the reference's kernel_periodic_boundconds (src/core/kernels/boundconds.cuh:
// 131-190): every ghost cell inside [start, end) receives the interior value at its index
// wrapped modulo n on each axis. The reference launches one thread per cell of the whole
// [start, end) box, once per field; here one launch covers only the m^3 - n^3 ghost cells
// (enumerated by GhostShell, ghost.h) and copies every requested field per thread.
#include "ghost.h"

namespace ac {

namespace {

constexpr int MAX_FILL_FIELDS = NF;

struct FillArgs {
    AcReal* f[MAX_FILL_FIELDS];
    int nfields;
    GhostShell g;
    int3 start, end;
};

__global__ void __launch_bounds__(256)
periodic_fill_kernel(const FillArgs a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.g.cells())
        return;
    int i, j, k;
    a.g.cell(t, i, j, k);
    if (i < a.start.x || j < a.start.y || k < a.start.z || i >= a.end.x || j >= a.end.y ||
        k >= a.end.z)
        return;
    long dst, src;
    a.g.indices(i, j, k, dst, src);
#pragma unroll
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        if (q < a.nfields)
            a.f[q][dst] = a.f[q][src];
}

} // namespace

hipError_t
launch_periodic_fill(AcReal* const* fields, int nfields, int nx, int ny, int nz, int3 start,
                     int3 end, hipStream_t stream)
{
    if (nfields < 1 || nfields > MAX_FILL_FIELDS)
        return hipErrorInvalidValue;
    FillArgs a;
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        a.f[q] = q < nfields ? fields[q] : nullptr;
    a.nfields = nfields;
    a.g       = GhostShell::of(nx, ny, nz);
    a.start   = start;
    a.end     = end;
    const long total = a.g.cells();
    const long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    periodic_fill_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

} // namespace ac
