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

#include <mutex>

// Per-block partials of the allocation-free reduction (acDeviceReduceNoAlloc): 4096 FP64 values
// (32 KiB), the most any reduction launch writes, since its block count is capped there. Each GPU
// holds its own copy in the library's code object until the process ends. In host code the name
// is a stand-in, not a device address: host code uses the address reduce_scratch_address looks
// up, never `scratch` itself.
static __device__ __align__(256) double scratch[4096];

namespace ac {

namespace {

constexpr int RED_THREADS  = 256;
constexpr int RED_WAVES    = RED_THREADS / 64;
constexpr int MAX_PARTIALS = 4096;
static_assert(sizeof(scratch) == MAX_PARTIALS * sizeof(double), "scratch holds one launch's partials");

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

// Filter of one cell given its NV values (NV = 1, 3 or 4 fields), for reduction type RT.
template <int NV, int RT>
__device__ __forceinline__ double
filter(const AcReal* x)
{
    // No FMA contraction: MAX/MIN of vector lengths must round like the host's a*a + b*b + c*c.
#pragma clang fp contract(off)
    const double kFourPi = 12.566370614359172953850573533118;
    if (NV == 1) {
        switch (RT) {
        case RTYPE_MAX:
        case RTYPE_MIN:
        case RTYPE_SUM: return (double)x[0];
        case RTYPE_RMS: return (double)x[0] * (double)x[0];
        case RTYPE_RMS_EXP: {
            const double e = exp((double)x[0]);
            return e * e;
        }
        }
    }
    else if (NV == 3) {
        switch (RT) {
        case RTYPE_MAX:
        case RTYPE_MIN: return (double)sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        case RTYPE_SUM:
            return sqrt((double)x[0] * x[0] + (double)x[1] * x[1] + (double)x[2] * x[2]);
        case RTYPE_RMS: return (double)x[0] * x[0] + (double)x[1] * x[1] + (double)x[2] * x[2];
        case RTYPE_RMS_EXP: {
            const double ex = exp((double)x[0]), ey = exp((double)x[1]), ez = exp((double)x[2]);
            return ex * ex + ey * ey + ez * ez;
        }
        }
    }
    else {
        switch (RT) {
        case RTYPE_ALFVEN_MAX:
        case RTYPE_ALFVEN_MIN:
            return sqrt((double)x[0] * x[0] + (double)x[1] * x[1] + (double)x[2] * x[2]) /
                   sqrt(kFourPi * exp((double)x[3]));
        case RTYPE_ALFVEN_RMS:
            return ((double)x[0] * x[0] + (double)x[1] * x[1] + (double)x[2] * x[2]) /
                   (kFourPi * exp((double)x[3]));
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

// Each wave walks whole x rows (grid-stride over (j, k)). A lane loads U cells of a row (U x NV
// values) before it folds any of them, so that every wave keeps U x NV loads in flight: with
// one load per fold the walk ran at 2.2 TB/s, bound by memory latency rather than bandwidth;
// now one 512^3 field takes 140 us in FP32 (3.8 TB/s) and 181 us in FP64 (5.9 TB/s), three
// fields 266 us (6.1 TB/s) and 553 us (5.8 TB/s).
template <int NV, int RT, int OP, int U>
__global__ void __launch_bounds__(RED_THREADS)
reduce_rows(const RedArgs a, double* partials)
{
    __shared__ double lds[RED_WAVES];
    const int lane  = threadIdx.x & 63;
    const int gwave = blockIdx.x * RED_WAVES + (threadIdx.x >> 6);
    const int nwave = gridDim.x * RED_WAVES;
    const int nrows = a.ny * a.nz;
    double acc      = identity(OP);
    for (int r = gwave; r < nrows; r += nwave) {
        const int j = 3 + r % a.ny, k = 3 + r / a.ny;
        const long row = (long)j * a.mx + (long)k * a.mxy + 3;
        for (int i0 = lane; i0 < a.nx; i0 += 64 * U) {
            AcReal x[U][NV];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + 64 * u;
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    x[u][q] = i < a.nx ? a.v[q][row + i] : AcReal(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (i0 + 64 * u < a.nx)
                    acc = combine(OP, acc, filter<NV, RT>(x[u]));
        }
    }
    const double res = block_reduce(OP, acc, lds);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = res;
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

// Diagnostics in one pass: max, min and sum of squares of |u|, |A| and lnrho over the interior,
// with the filters of the single reductions (MAX/MIN of a vector length rounded in AcReal as in
// filter<3, RTYPE_MAX>, sums of squares in FP64), so that the maxima and minima equal theirs
// bit for bit. Each lane loads U cells of its row (7 fields each) before folding them.
constexpr int NSTAT = 9; // u max, min, sum2; A max, min, sum2; lnrho max, min, sum2

struct DiagArgs {
    const AcReal* f[NF];
    int nx, ny, nz, mx, mxy;
};

template <int U>
__global__ void __launch_bounds__(RED_THREADS)
diagnostics_rows(const DiagArgs a, double* partials)
{
    __shared__ double lds[NSTAT][RED_WAVES];
    const int lane  = threadIdx.x & 63;
    const int gwave = blockIdx.x * RED_WAVES + (threadIdx.x >> 6);
    const int nwave = gridDim.x * RED_WAVES;
    const int nrows = a.ny * a.nz;
    double v[NSTAT] = {-INFINITY, INFINITY, 0.0, -INFINITY, INFINITY, 0.0, -INFINITY, INFINITY, 0.0};
    for (int r = gwave; r < nrows; r += nwave) {
        const int j = 3 + r % a.ny, k = 3 + r / a.ny;
        const long row = (long)j * a.mx + (long)k * a.mxy + 3;
        for (int i0 = lane; i0 < a.nx; i0 += 64 * U) {
            AcReal u[U][3], b[U][3], l[U];
#pragma unroll
            for (int w = 0; w < U; ++w) {
                const int i   = i0 + 64 * w;
                const bool ok = i < a.nx;
                u[w][0] = ok ? a.f[VTXBUF_UUX][row + i] : AcReal(0);
                u[w][1] = ok ? a.f[VTXBUF_UUY][row + i] : AcReal(0);
                u[w][2] = ok ? a.f[VTXBUF_UUZ][row + i] : AcReal(0);
                b[w][0] = ok ? a.f[VTXBUF_AX][row + i] : AcReal(0);
                b[w][1] = ok ? a.f[VTXBUF_AY][row + i] : AcReal(0);
                b[w][2] = ok ? a.f[VTXBUF_AZ][row + i] : AcReal(0);
                l[w]    = ok ? a.f[VTXBUF_LNRHO][row + i] : AcReal(0);
            }
#pragma unroll
            for (int w = 0; w < U; ++w) {
                if (i0 + 64 * w >= a.nx)
                    continue;
                const double un = filter<3, RTYPE_MAX>(u[w]), us = filter<3, RTYPE_RMS>(u[w]);
                const double an = filter<3, RTYPE_MAX>(b[w]), as = filter<3, RTYPE_RMS>(b[w]);
                const double ln = filter<1, RTYPE_MAX>(&l[w]), ls = filter<1, RTYPE_RMS>(&l[w]);
                v[0] = combine(OP_MAX, v[0], un);
                v[1] = combine(OP_MIN, v[1], un);
                v[2] += us;
                v[3] = combine(OP_MAX, v[3], an);
                v[4] = combine(OP_MIN, v[4], an);
                v[5] += as;
                v[6] = combine(OP_MAX, v[6], ln);
                v[7] = combine(OP_MIN, v[7], ln);
                v[8] += ls;
            }
        }
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NSTAT; ++q) {
        const int op = q % 3 == 0 ? OP_MAX : q % 3 == 1 ? OP_MIN : OP_SUM;
        const double x = wave_reduce(op, v[q]);
        if (lane == 0)
            lds[q][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < NSTAT) {
        const int q  = threadIdx.x;
        const int op = q % 3 == 0 ? OP_MAX : q % 3 == 1 ? OP_MIN : OP_SUM;
        double x     = lds[q][0];
        for (int i = 1; i < RED_WAVES; ++i)
            x = combine(op, x, lds[q][i]);
        partials[(long)blockIdx.x * NSTAT + q] = x;
    }
}

// Block q folds statistic q of every block's partials in a fixed order.
__global__ void __launch_bounds__(RED_THREADS)
diagnostics_finalize(const double* partials, int nblocks, double* result)
{
    __shared__ double lds[RED_WAVES];
    const int q  = blockIdx.x;
    const int op = q % 3 == 0 ? OP_MAX : q % 3 == 1 ? OP_MIN : OP_SUM;
    double acc   = identity(op);
    for (int i = threadIdx.x; i < nblocks; i += RED_THREADS)
        acc = combine(op, acc, partials[(long)i * NSTAT + q]);
    const double r = block_reduce(op, acc, lds);
    if (threadIdx.x == 0)
        result[q] = r;
}

} // namespace

int
diagnostics_partials_needed()
{
    return MAX_PARTIALS * NSTAT;
}

hipError_t
launch_diagnostics(const AcReal* const* fields, const AcMeshInfo& info, double* d_partials,
                   double* d_result, hipStream_t stream)
{
    DiagArgs a;
    for (int q = 0; q < NF; ++q)
        a.f[q] = fields[q];
    a.nx   = info.int_params[AC_nx];
    a.ny   = info.int_params[AC_ny];
    a.nz   = info.int_params[AC_nz];
    a.mx   = a.nx + 6;
    a.mxy  = a.mx * (a.ny + 6);
    int blocks = (a.ny * a.nz + RED_WAVES * 4 - 1) / (RED_WAVES * 4);
    blocks     = blocks < 1 ? 1 : blocks > MAX_PARTIALS ? MAX_PARTIALS : blocks;
    diagnostics_rows<4><<<dim3(blocks), dim3(RED_THREADS), 0, stream>>>(a, d_partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    diagnostics_finalize<<<dim3(NSTAT), dim3(RED_THREADS), 0, stream>>>(d_partials, blocks,
                                                                        d_result);
    return hipGetLastError();
}

int
reduce_partials_needed()
{
    return MAX_PARTIALS;
}

int
reduce_scratch_capacity()
{
    return MAX_PARTIALS;
}

hipError_t
reduce_scratch_address(int device_id, double** out)
{
    // one lookup per GPU: each holds its own copy of the array at its own address
    static std::mutex mu;
    static std::map<int, double*> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto it = cache.find(device_id);
    if (it != cache.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p            = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    cache[device_id] = static_cast<double*>(p);
    *out             = static_cast<double*>(p);
    return hipSuccess;
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
    const dim3 g(blocks), t(RED_THREADS);
#define AC_RED(NV, RT, OP, U)                                                                      \
    if (nvals == NV && rtype == RT) {                                                              \
        reduce_rows<NV, RT, OP, U><<<g, t, 0, stream>>>(a, d_partials);                            \
    } else
    // cells per lane and pass (measured at 512^3: 8 / 4; 16 / 8 were slower in FP32)
    constexpr int U1 = 8, U3 = 4;
    AC_RED(1, RTYPE_MAX, OP_MAX, U1)
    AC_RED(1, RTYPE_MIN, OP_MIN, U1)
    AC_RED(1, RTYPE_SUM, OP_SUM, U1)
    AC_RED(1, RTYPE_RMS, OP_SUM, U1)
    AC_RED(1, RTYPE_RMS_EXP, OP_SUM, U1)
    AC_RED(3, RTYPE_MAX, OP_MAX, U3)
    AC_RED(3, RTYPE_MIN, OP_MIN, U3)
    AC_RED(3, RTYPE_SUM, OP_SUM, U3)
    AC_RED(3, RTYPE_RMS, OP_SUM, U3)
    AC_RED(3, RTYPE_RMS_EXP, OP_SUM, U3)
    AC_RED(4, RTYPE_ALFVEN_MAX, OP_MAX, U3)
    AC_RED(4, RTYPE_ALFVEN_MIN, OP_MIN, U3)
    AC_RED(4, RTYPE_ALFVEN_RMS, OP_SUM, U3)
    return hipErrorInvalidValue;
#undef AC_RED
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    reduce_partials<<<dim3(1), dim3(RED_THREADS), 0, stream>>>(d_partials, blocks, op, d_result);
    return hipGetLastError();
}

} // namespace ac
