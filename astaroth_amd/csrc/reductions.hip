// Reductions over the interior of one, three or four fields.
//
// Replaces the reference's three-stage reduction (src/core/kernels/reductions.cuh:184-342:
// filter into an n^3 scratch buffer, 128-thread shared-memory tree, then a single thread
// looping over every partial) with two launches and no scratch field:
//   1. reduce_rows: each wave64 walks whole x rows (grid-stride over (j, k)), applies the filter
//      per cell, accumulates in FP64, reduces across the wave with DPP/shuffles and across the
//      block through LDS, and writes one FP64 partial per block;
//   2. reduce_partials: one block folds the partials in a fixed order.
// The tree shape depends only on the grid size, so results are deterministic run to run.
// Filters (reductions.cuh:18-50): value, square, exp(a)^2, vector length, squared length,
// Alfven speed |B| / sqrt(4 pi exp(lnrho)) and its square. Vector SUM sums lengths, as in the
// reference (reductions.cuh:280). MAX/MIN filters are evaluated in AcReal, like the reference,
// and compared exactly; sums accumulate in FP64.
#include "ac_internal.h"

namespace ac {

namespace {

constexpr int RED_THREADS  = 256;
constexpr int RED_WAVES    = RED_THREADS / 64;
constexpr int MAX_PARTIALS = 4096;

enum Op { OP_MAX, OP_MIN, OP_SUM };

struct RedArgs {
    const AcReal* v[4];
    int nvals;
    int rtype;
    int nx, ny, nz, mx, mxy;
};

__device__ __forceinline__ double
combine(int op, double a, double b)
{
    return op == OP_MAX ? (a > b ? a : b) : op == OP_MIN ? (a < b ? a : b) : a + b;
}

__device__ __forceinline__ double
identity(int op)
{
    return op == OP_MAX ? -INFINITY : op == OP_MIN ? INFINITY : 0.0;
}

__device__ __forceinline__ double
filter(const RedArgs& a, long idx)
{
    // No FMA contraction: MAX/MIN of vector lengths must round like the host's a*a + b*b + c*c.
#pragma clang fp contract(off)
    const double kFourPi = 12.566370614359172953850573533118;
    if (a.nvals == 1) {
        const AcReal x = a.v[0][idx];
        switch (a.rtype) {
        case RTYPE_MAX:
        case RTYPE_MIN:
        case RTYPE_SUM: return (double)x;
        case RTYPE_RMS: return (double)x * (double)x;
        case RTYPE_RMS_EXP: {
            const double e = exp((double)x);
            return e * e;
        }
        }
    }
    else if (a.nvals == 3) {
        const AcReal x = a.v[0][idx], y = a.v[1][idx], z = a.v[2][idx];
        switch (a.rtype) {
        case RTYPE_MAX:
        case RTYPE_MIN: return (double)sqrt(x * x + y * y + z * z);
        case RTYPE_SUM: return sqrt((double)x * x + (double)y * y + (double)z * z);
        case RTYPE_RMS: return (double)x * x + (double)y * y + (double)z * z;
        case RTYPE_RMS_EXP: {
            const double ex = exp((double)x), ey = exp((double)y), ez = exp((double)z);
            return ex * ex + ey * ey + ez * ez;
        }
        }
    }
    else {
        const AcReal x = a.v[0][idx], y = a.v[1][idx], z = a.v[2][idx], d = a.v[3][idx];
        switch (a.rtype) {
        case RTYPE_ALFVEN_MAX:
        case RTYPE_ALFVEN_MIN:
            return sqrt((double)x * x + (double)y * y + (double)z * z) /
                   sqrt(kFourPi * exp((double)d));
        case RTYPE_ALFVEN_RMS:
            return ((double)x * x + (double)y * y + (double)z * z) / (kFourPi * exp((double)d));
        }
    }
    return NAN;
}

__device__ __forceinline__ double
wave_reduce(int op, double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = combine(op, v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double
block_reduce(int op, double v, double* lds)
{
    v = wave_reduce(op, v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        lds[wave] = v;
    __syncthreads();
    double r = identity(op);
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            r = combine(op, r, lds[w]);
    }
    return r;
}

__global__ void __launch_bounds__(RED_THREADS)
reduce_rows(const RedArgs a, int op, double* partials)
{
    __shared__ double lds[RED_WAVES];
    const int lane  = threadIdx.x & 63;
    const int gwave = blockIdx.x * RED_WAVES + (threadIdx.x >> 6);
    const int nwave = gridDim.x * RED_WAVES;
    const int nrows = a.ny * a.nz;
    double acc = identity(op);
    for (int r = gwave; r < nrows; r += nwave) {
        const int j = 3 + r % a.ny, k = 3 + r / a.ny;
        const long row = (long)j * a.mx + (long)k * a.mxy + 3;
        for (int i = lane; i < a.nx; i += 64)
            acc = combine(op, acc, filter(a, row + i));
    }
    const double r = block_reduce(op, acc, lds);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = r;
}

__global__ void __launch_bounds__(RED_THREADS)
reduce_partials(const double* partials, int n, int op, double* result)
{
    __shared__ double lds[RED_WAVES];
    double acc = identity(op);
    for (int i = threadIdx.x; i < n; i += RED_THREADS)
        acc = combine(op, acc, partials[i]);
    const double r = block_reduce(op, acc, lds);
    if (threadIdx.x == 0)
        *result = r;
}

} // namespace

int
reduce_partials_needed()
{
    return MAX_PARTIALS;
}

hipError_t
launch_reduce(int rtype, int nvals, const AcReal* const* vals, const AcMeshInfo& info,
              double* d_partials, int max_partials, double* d_result, hipStream_t stream)
{
    RedArgs a;
    for (int q = 0; q < 4; ++q)
        a.v[q] = q < nvals ? vals[q] : nullptr;
    a.nvals = nvals;
    a.rtype = rtype;
    a.nx    = info.int_params[AC_nx];
    a.ny    = info.int_params[AC_ny];
    a.nz    = info.int_params[AC_nz];
    a.mx    = a.nx + 6;
    a.mxy   = a.mx * (a.ny + 6);
    int op;
    switch (rtype) {
    case RTYPE_MAX:
    case RTYPE_ALFVEN_MAX: op = OP_MAX; break;
    case RTYPE_MIN:
    case RTYPE_ALFVEN_MIN: op = OP_MIN; break;
    default: op = OP_SUM; break;
    }
    const int nrows  = a.ny * a.nz;
    int blocks       = (nrows + RED_WAVES * 4 - 1) / (RED_WAVES * 4);
    const int maxblk = max_partials < MAX_PARTIALS ? max_partials : MAX_PARTIALS;
    if (blocks > maxblk)
        blocks = maxblk;
    if (blocks < 1)
        blocks = 1;
    reduce_rows<<<dim3(blocks), dim3(RED_THREADS), 0, stream>>>(a, op, d_partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    reduce_partials<<<dim3(1), dim3(RED_THREADS), 0, stream>>>(d_partials, blocks, op, d_result);
    return hipGetLastError();
}

} // namespace ac
