// Device-side field initialisation: constant fill and a counter-based uniform generator.
//
// The generator is splitmix64 applied to (seed + (cell + 1) * golden gamma), where `cell` is the
// linear interior index (i - 3) + (j - 3) n_x + (k - 3) n_x n_y; the top 53 bits give
// u in [0, 1) and the value is lo + (hi - lo) u (+ amp sin(kdz (k - 3)) when amp != 0)
// evaluated in FP64 and rounded once to AcReal.
// Being counter based it needs no state, gives the same numbers for any launch shape, and is
// reproduced exactly by astaroth_amd/initial.py on the host.
#include "ac_internal.h"

namespace ac {

namespace {

__device__ __forceinline__ unsigned long long
splitmix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256)
fill_uniform_kernel(AcReal* f, int nx, int ny, int nz, int mx, long mxy, double lo, double hi,
                    unsigned long long seed, double amp, double kdz)
{
#pragma clang fp contract(off) // lo + (hi - lo) u (+ amp sin) rounds exactly as on the host
    const long t     = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)nx * ny * nz;
    if (t >= total)
        return;
    const int i = (int)(t % nx);
    const long r = t / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const unsigned long long z = splitmix64(seed + (unsigned long long)(t + 1) * 0x9E3779B97F4A7C15ULL);
    const double u             = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    double v                   = lo + (hi - lo) * u;
    if (amp != 0.0)
        v = v + amp * sin(kdz * (double)k);
    f[(i + 3) + (long)(j + 3) * mx + (long)(k + 3) * mxy] = (AcReal)v;
}

__global__ void __launch_bounds__(256)
fill_constant_kernel(AcReal* f, long n, AcReal value)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        f[t] = value;
}

__global__ void __launch_bounds__(256)
copy_kernel(AcReal* __restrict__ dst, const AcReal* __restrict__ src, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        dst[t] = src[t];
}

} // namespace

hipError_t
launch_copy(AcReal* dst, const AcReal* src, long n, hipStream_t stream)
{
    const long blocks = (n + 255) / 256;
    if (blocks <= 0)
        return hipSuccess;
    copy_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(dst, src, n);
    return hipGetLastError();
}

hipError_t
launch_fill_uniform(AcReal* field, const AcMeshInfo& info, double lo, double hi,
                    unsigned long long seed, hipStream_t stream, double amp, double kdz)
{
    const int nx = info.int_params[AC_nx], ny = info.int_params[AC_ny], nz = info.int_params[AC_nz];
    const int mx = nx + 6, my = ny + 6;
    const long total  = (long)nx * ny * nz;
    const long blocks = (total + 255) / 256;
    if (blocks <= 0)
        return hipSuccess;
    fill_uniform_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(
        field, nx, ny, nz, mx, (long)mx * my, lo, hi, seed, amp, kdz);
    return hipGetLastError();
}

hipError_t
launch_fill_constant(AcReal* field, long n, AcReal value, hipStream_t stream)
{
    const long blocks = (n + 255) / 256;
    if (blocks <= 0)
        return hipSuccess;
    fill_constant_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(field, n, value);
    return hipGetLastError();
}

} // namespace ac
