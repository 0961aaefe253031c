// Periodic ghost-cell fill.
//
// Same result as the reference's kernel_periodic_boundconds (src/core/kernels/boundconds.cuh:
// 131-190): every ghost cell inside [start, end) receives the interior value at its index
// wrapped modulo n on each axis. The reference launches one thread per cell of the whole
// [start, end) box, once per field; here one launch covers only the m^3 - n^3 ghost cells and
// copies every requested field per thread. The ghost shell is enumerated as
//   Z slabs: k in [0,3) u [n_z+3, m_z), all (i, j)            6 m_x m_y cells
//   Y slabs: k interior, j in [0,3) u [n_y+3, m_y), all i     6 n_z m_x cells
//   X slabs: k, j interior, i in [0,3) u [n_x+3, m_x)         6 n_z n_y cells
// so consecutive threads of the Z and Y slabs walk contiguous x rows.
#include "ac_internal.h"

namespace ac {

namespace {

constexpr int MAX_FILL_FIELDS = NF;

struct FillArgs {
    AcReal* f[MAX_FILL_FIELDS];
    int nfields;
    int nx, ny, nz, mx, my;
    int3 start, end;
    long cz, cy, cx; // cell counts of the three slab families
};

__global__ void __launch_bounds__(256)
periodic_fill_kernel(const FillArgs a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < a.cz) {
        const long plane = (long)a.mx * a.my;
        const int pz     = (int)(t / plane);
        const long rem   = t - (long)pz * plane;
        k                = pz < 3 ? pz : a.nz + pz;
        j                = (int)(rem / a.mx);
        i                = (int)(rem - (long)j * a.mx);
    }
    else if (t < a.cz + a.cy) {
        const long u    = t - a.cz;
        const long slab = 6L * a.mx;
        const int kk    = (int)(u / slab);
        const long rem  = u - (long)kk * slab;
        const int jr    = (int)(rem / a.mx);
        k               = 3 + kk;
        j               = jr < 3 ? jr : a.ny + jr;
        i               = (int)(rem - (long)jr * a.mx);
    }
    else if (t < a.cz + a.cy + a.cx) {
        const long u    = t - a.cz - a.cy;
        const long slab = 6L * a.ny;
        const int kk    = (int)(u / slab);
        const int rem   = (int)(u - (long)kk * slab);
        const int ir    = rem % 6;
        k               = 3 + kk;
        j               = 3 + rem / 6;
        i               = ir < 3 ? ir : a.nx + ir;
    }
    else {
        return;
    }
    if (i < a.start.x || j < a.start.y || k < a.start.z || i >= a.end.x || j >= a.end.y ||
        k >= a.end.z)
        return;
    const int is = (i - 3 + a.nx) % a.nx + 3;
    const int js = (j - 3 + a.ny) % a.ny + 3;
    const int ks = (k - 3 + a.nz) % a.nz + 3;
    const long dst = i + (long)j * a.mx + (long)k * a.mx * a.my;
    const long src = is + (long)js * a.mx + (long)ks * a.mx * a.my;
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
    a.nx      = nx;
    a.ny      = ny;
    a.nz      = nz;
    a.mx      = nx + 6;
    a.my      = ny + 6;
    a.start   = start;
    a.end     = end;
    a.cz      = 6L * a.mx * a.my;
    a.cy      = 6L * nz * a.mx;
    a.cx      = 6L * nz * ny;
    const long total = a.cz + a.cy + a.cx;
    const long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    periodic_fill_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

} // namespace ac
