// Ghost-cell fills: periodic, and per axis and side periodic, symmetric or antisymmetric.
//
// Same result as the reference's kernel_periodic_boundconds (src/core/kernels/boundconds.cuh:
// 131-190): every ghost cell inside [start, end) receives the interior value at its index
// wrapped modulo n on each axis. The reference launches one thread per cell of the whole
// [start, end) box, once per field; here one launch covers only the m^3 - n^3 ghost cells
// (enumerated by GhostShell, ghost.h) and copies every requested field per thread.
//
// The general fill is what the reference's kernel_symmetric_boundconds (boundconds.cuh:3-128,
// launched by acKernelGeneralBoundconds, :192-242) sets out to do, done as its comments describe
// rather than as it runs: that kernel stores sign * f * 0.0 + 1.0 (every ghost cell becomes 1),
// reads only the x-top boundary type for all six faces, and compares y against the x bound in
// two edge cases. Here each face takes its own type (AC_bc_type_{bot,top}_{x,y,z}): a periodic
// face wraps, a symmetric face mirrors about its boundary point (the first or last interior
// cell: ghost 2, 1, 0 take cells 4, 5, 6), an antisymmetric face mirrors with a sign change. An
// edge or corner cell mirrors or wraps on each of its ghost axes at once and takes the product
// of the signs, the same as filling the axes one after another.
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

// NFIX: the field count when fixed at compile time (all eight fields, the common call), else 0
template <int NFIX>
__global__ void __launch_bounds__(256)
periodic_fill_kernel(const FillArgs a)
{
    const int nf = NFIX ? NFIX : a.nfields;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.g.cells())
        return;
    int i, j, k;
    a.g.cell((unsigned)t, i, j, k);
    if (i < a.start.x || j < a.start.y || k < a.start.z || i >= a.end.x || j >= a.end.y ||
        k >= a.end.z)
        return;
    long dst, src;
    a.g.indices(i, j, k, dst, src);
    // Every field's load is issued before the first store. The field pointers may alias as far
    // as the compiler knows, so field by field each load would wait for the store before it:
    // eight memory round trips in a row per thread instead of one.
    AcReal v[MAX_FILL_FIELDS];
#pragma unroll
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        if (q < nf)
            v[q] = a.f[q][src];
#pragma unroll
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        if (q < nf)
            a.f[q][dst] = v[q];
}

// Per axis and side: AC_BOUNDCOND_PERIODIC (0), _SYMMETRIC (1) or _ANTISYMMETRIC (2).
struct GeneralArgs {
    AcReal* f[MAX_FILL_FIELDS];
    int nfields;
    GhostShell g;
    int3 start, end;
    int bot[3], top[3];
};

// Source coordinate on one axis of ghost coordinate c (box of n interior cells from 3): periodic
// images wrap by n; symmetric and antisymmetric ones mirror about the boundary point, the first
// (3) or last (n + 2) interior cell, the antisymmetric ones with a sign change.
__device__ __forceinline__ int
bc_source(int c, int n, int bot, int top, bool& negate)
{
    if (c < 3) {
        if (bot == AC_BOUNDCOND_PERIODIC)
            return c + n;
        negate ^= bot == AC_BOUNDCOND_ANTISYMMETRIC;
        return 6 - c;
    }
    if (c >= n + 3) {
        if (top == AC_BOUNDCOND_PERIODIC)
            return c - n;
        negate ^= top == AC_BOUNDCOND_ANTISYMMETRIC;
        return 2 * (n + 2) - c;
    }
    return c;
}

template <int NFIX>
__global__ void __launch_bounds__(256)
general_fill_kernel(const GeneralArgs a)
{
    const int nf = NFIX ? NFIX : a.nfields;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.g.cells())
        return;
    int i, j, k;
    a.g.cell((unsigned)t, i, j, k);
    if (i < a.start.x || j < a.start.y || k < a.start.z || i >= a.end.x || j >= a.end.y ||
        k >= a.end.z)
        return;
    bool negate  = false;
    const int is = bc_source(i, a.g.nx, a.bot[0], a.top[0], negate);
    const int js = bc_source(j, a.g.ny, a.bot[1], a.top[1], negate);
    const int ks = bc_source(k, a.g.nz, a.bot[2], a.top[2], negate);
    const long mxy = (long)a.g.mx * a.g.my;
    const long dst = i + (long)j * a.g.mx + (long)k * mxy;
    const long src = is + (long)js * a.g.mx + (long)ks * mxy;
    AcReal v[MAX_FILL_FIELDS]; // all loads before the first store, as in the periodic fill
#pragma unroll
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        if (q < nf)
            v[q] = a.f[q][src];
#pragma unroll
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        if (q < nf)
            a.f[q][dst] = negate ? -v[q] : v[q];
}

} // namespace

hipError_t
launch_general_fill(AcReal* const* fields, int nfields, int nx, int ny, int nz, int3 start,
                    int3 end, const int bot[3], const int top[3], hipStream_t stream)
{
    if (nfields < 1 || nfields > MAX_FILL_FIELDS)
        return hipErrorInvalidValue;
    GeneralArgs a;
    for (int q = 0; q < MAX_FILL_FIELDS; ++q)
        a.f[q] = q < nfields ? fields[q] : nullptr;
    a.nfields = nfields;
    a.g       = GhostShell::of(nx, ny, nz);
    if (!a.g.fits_32bit())
        return hipErrorInvalidValue;
    a.start   = start;
    a.end     = end;
    for (int d = 0; d < 3; ++d) {
        a.bot[d] = bot[d];
        a.top[d] = top[d];
    }
    const long total  = a.g.cells();
    const long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    if (nfields == MAX_FILL_FIELDS)
        general_fill_kernel<MAX_FILL_FIELDS><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    else
        general_fill_kernel<0><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

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
    if (!a.g.fits_32bit())
        return hipErrorInvalidValue;
    a.start   = start;
    a.end     = end;
    const long total = a.g.cells();
    const long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    if (nfields == MAX_FILL_FIELDS)
        periodic_fill_kernel<MAX_FILL_FIELDS><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    else
        periodic_fill_kernel<0><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

} // namespace ac
