// Batched ensembles: many independent periodic boxes of uneven sizes advanced together.
//
// There is no counterpart in the reference, which runs one box per Device; this is the batching
// pattern of the single-device path. Each box keeps the reference's layout (8 `in` and 8 `out`
// fields of m_x m_y m_z cells, x fastest) and its own AcMeshInfo (grid spacing, viscosity, ...).
// One RK3 substep is
//   * one periodic-fill launch over the ghost shells of all boxes (a flat list of ghost cells),
//   * one persistent solve launch (solve_zlean_batch) over a (box, tile column, z range) work
//     list sorted by decreasing cost, dealt to the resident blocks in rounds of one item per
//     block, each XCD taking a contiguous range of a round's items,
// and a step ends with per-box FP64 diagnostics (max and sum of squares of |u| and |A|, max, min
// and sum of squares of lnrho) in two launches. acEnsembleIntegrateStepGraph records the whole
// step once per (dt, buffer parity) into a HIP graph and replays it.
#include "ghost.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

using namespace ac;

namespace {

constexpr int NDIAG     = 7;  // per-box diagnostics, see acEnsembleDiagnostics
constexpr int DIAG_CELLS  = 16384; // interior cells per block of the diagnostics launch
constexpr int DIAG_UNROLL = 4;     // cells per lane whose loads are in flight together
constexpr int FILL_BLOCK  = 256; // threads per block of the batched fill

// The field pointers of a box come from a device array, so the compiler cannot tell their
// address space and would use flat loads and stores; they are global memory.
typedef __attribute__((address_space(1))) AcReal GlobalReal;
__device__ __forceinline__ GlobalReal*
global(AcReal* p)
{
    return (GlobalReal*)p;
}
__device__ __forceinline__ const GlobalReal*
global(const AcReal* p)
{
    return (const GlobalReal*)p;
}

struct FillBox {
    AcReal* f[NF];
    GhostShell g;
    long first; // thread index of the box's first ghost cell, a multiple of FILL_BLOCK
};

// Each box's ghost cells start on a block boundary, so a block serves one box and finds it with
// a search keyed by its block index alone (uniform, scalar loads).
__global__ void __launch_bounds__(FILL_BLOCK)
periodic_fill_batch(const FillBox* __restrict__ boxes, const int nboxes)
{
    const long t0 = (long)blockIdx.x * FILL_BLOCK;
    int lo = 0, hi = nboxes - 1; // last box whose first cell is <= t0
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].first <= t0)
            lo = mid;
        else
            hi = mid - 1;
    }
    const FillBox& b = boxes[lo];
    const long u     = t0 - b.first + threadIdx.x;
    if (u >= b.g.cells()) // the padding up to the next box's first block
        return;
    int i, j, k;
    b.g.cell((unsigned)u, i, j, k);
    long dst, src;
    b.g.indices(i, j, k, dst, src);
    // all eight loads in flight before the first store (the field pointers may alias as far as
    // the compiler knows, which would otherwise make each load wait for the store before it)
    AcReal v[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q)
        v[q] = global(b.f[q])[src];
#pragma unroll
    for (int q = 0; q < NF; ++q)
        global(b.f[q])[dst] = v[q];
}

struct DiagBox {
    const AcReal* f[NF];
    int nx, ny, nz, mx;
    long mxy;
    int first_blk; // the box's first block of the diagnostics launch (ceil(cells / DIAG_CELLS) each)
};

__device__ __forceinline__ double
dmax(double a, double b)
{
    return a > b ? a : b;
}
__device__ __forceinline__ double
dmin(double a, double b)
{
    return a < b ? a : b;
}

// Quantity q of the diagnostics: 0 max |u|^2, 1 sum |u|^2, 2 max |A|^2, 3 sum |A|^2,
// 4 max lnrho, 5 min lnrho, 6 sum lnrho^2 (square roots taken when finalising).
__device__ __forceinline__ bool
is_max(int q)
{
    return q == 0 || q == 2 || q == 4;
}

// Every block takes DIAG_CELLS cells of one box's interior, walked as one flat list (x fastest,
// rows back to back), so that short rows leave no lane idle and every block does the same work
// whatever the box's size; the block finds its box by a search keyed by its index alone
// (uniform). Each lane has the loads of DIAG_UNROLL cells in flight before it accumulates them.
__global__ void __launch_bounds__(256)
diagnostics_batch(const DiagBox* __restrict__ boxes, const int nboxes, double* __restrict__ partials)
{
    __shared__ double lds[4][NDIAG];
    int lo = 0, hi = nboxes - 1; // last box whose first block is <= blockIdx.x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].first_blk <= (int)blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const DiagBox& b = boxes[lo];
    double v[NDIAG]  = {0.0, 0.0, 0.0, 0.0, -INFINITY, INFINITY, 0.0};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned nx = (unsigned)b.nx, nxy = (unsigned)b.nx * (unsigned)b.ny;
    const unsigned cbeg   = (blockIdx.x - (unsigned)b.first_blk) * (unsigned)DIAG_CELLS;
    const unsigned ncell  = min(nxy * (unsigned)b.nz, cbeg + (unsigned)DIAG_CELLS);
    const unsigned stride = 256u;
    for (unsigned c0 = cbeg + threadIdx.x; c0 < ncell; c0 += DIAG_UNROLL * stride) {
        AcReal f[DIAG_UNROLL][7];
        bool ok[DIAG_UNROLL];
#pragma unroll
        for (int uu = 0; uu < DIAG_UNROLL; ++uu) {
            const unsigned c = c0 + uu * stride;
            ok[uu]           = c < ncell;
            const unsigned k = (ok[uu] ? c : 0u) / nxy, rem = (ok[uu] ? c : 0u) - k * nxy;
            const unsigned j = rem / nx, i = rem - j * nx;
            const long m = (long)(i + 3) + (long)(j + 3) * b.mx + (long)(k + 3) * b.mxy;
            f[uu][0] = global(b.f[VTXBUF_UUX])[m];
            f[uu][1] = global(b.f[VTXBUF_UUY])[m];
            f[uu][2] = global(b.f[VTXBUF_UUZ])[m];
            f[uu][3] = global(b.f[VTXBUF_AX])[m];
            f[uu][4] = global(b.f[VTXBUF_AY])[m];
            f[uu][5] = global(b.f[VTXBUF_AZ])[m];
            f[uu][6] = global(b.f[VTXBUF_LNRHO])[m];
        }
#pragma unroll
        for (int uu = 0; uu < DIAG_UNROLL; ++uu) {
            if (!ok[uu])
                continue;
            const double ux = f[uu][0], uy = f[uu][1], uz = f[uu][2];
            const double ax = f[uu][3], ay = f[uu][4], az = f[uu][5], lr = f[uu][6];
            const double u2 = ux * ux + uy * uy + uz * uz, a2 = ax * ax + ay * ay + az * az;
            v[0] = dmax(v[0], u2);
            v[1] += u2;
            v[2] = dmax(v[2], a2);
            v[3] += a2;
            v[4] = dmax(v[4], lr);
            v[5] = dmin(v[5], lr);
            v[6] += lr * lr;
        }
    }
#pragma unroll
    for (int q = 0; q < NDIAG; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_xor(x, off, 64);
            x              = q == 5 ? dmin(x, y) : is_max(q) ? dmax(x, y) : x + y;
        }
        if (lane == 0)
            lds[wave][q] = x;
    }
    __syncthreads();
    if (threadIdx.x < NDIAG) {
        const int q = threadIdx.x;
        double x    = lds[0][q];
        for (int w = 1; w < 4; ++w)
            x = q == 5 ? dmin(x, lds[w][q]) : is_max(q) ? dmax(x, lds[w][q]) : x + lds[w][q];
        partials[(long)blockIdx.x * NDIAG + q] = x;
    }
}

// Per box and quantity, the box's block partials folded in block order.
__global__ void
diagnostics_finalize(const double* __restrict__ partials, const DiagBox* __restrict__ boxes,
                     const int nboxes, const int nblk, double* __restrict__ results)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nboxes * NDIAG)
        return;
    const int box = t / NDIAG, q = t % NDIAG;
    const int b0 = boxes[box].first_blk, b1 = box + 1 < nboxes ? boxes[box + 1].first_blk : nblk;
    double x = partials[(long)b0 * NDIAG + q];
    for (int i = b0 + 1; i < b1; ++i) {
        const double y = partials[(long)i * NDIAG + q];
        x              = q == 5 ? dmin(x, y) : is_max(q) ? dmax(x, y) : x + y;
    }
    results[t] = x;
}

} // namespace

struct ensemble_s {
    int id;
    int nboxes;
    std::vector<AcMeshInfo> info;
    std::vector<AcReal*> fields[2][NF]; // [set][field][box]; set 0/1 = in/out at parity 0
    AcReal* base;
    size_t base_bytes;
    hipStream_t stream;
    int parity; // number of swaps mod 2: set `parity` holds the current state
    // batched solve: one tile shape per launch; the z-depth of a box's work items is that of its
    // size class (its column length n_z) when the tuner set one, else z_depth (0: whole columns)
    int tile_x, tile_y, z_depth, nblocks;
    std::map<int, int> class_depth;
    BatchItem* d_items;
    int nitems;
    SubstepArgs* d_args[3][2]; // per (substep, parity), for dt_args
    AcReal dt_args;
    bool args_valid;
    // batched fill and diagnostics, per parity
    FillBox* d_fill[2];
    long fill_total;
    DiagBox* d_diag[2];
    int diag_blocks; // blocks of the diagnostics launch, DIAG_CELLS interior cells each
    double* d_partials;
    double* d_results;
    double* h_results; // pinned
    std::map<unsigned long long, hipGraphExec_t> graphs[2];
};

namespace {

int3
dims_of(const AcMeshInfo& i)
{
    return make_int3(i.int_params[AC_nx], i.int_params[AC_ny], i.int_params[AC_nz]);
}

size_t
cells_of(const AcMeshInfo& i)
{
    const int3 n = dims_of(i);
    return (size_t)(n.x + 6) * (n.y + 6) * (n.z + 6);
}

bool
ens_ok(const Ensemble e, int box)
{
    return e && box >= 0 && box < e->nboxes;
}

void
drop_graphs(Ensemble e)
{
    for (int p = 0; p < 2; ++p) {
        for (auto& kv : e->graphs[p])
            (void)hipGraphExecDestroy(kv.second);
        e->graphs[p].clear();
    }
}

// Output planes per work item of a box whose columns are nz planes long.
int
box_depth(const Ensemble e, int nz)
{
    const auto it = e->class_depth.find(nz);
    const int zd  = it != e->class_depth.end() ? it->second : e->z_depth;
    return (zd <= 0 || zd >= nz) ? nz : zd;
}

AcResult
build_items(Ensemble e)
{
    struct Item {
        BatchItem w;
        long cost;
    };
    std::vector<Item> items;
    const int TX = e->tile_x, TY = e->tile_y;
    for (int b = 0; b < e->nboxes; ++b) {
        const int3 n  = dims_of(e->info[b]);
        const int zd = box_depth(e, n.z);
        const int ntx = (n.x + TX - 1) / TX, nty = (n.y + TY - 1) / TY, ntz = (n.z + zd - 1) / zd;
        for (int tz = 0; tz < ntz; ++tz) {
            const int zs    = NGHOST + tz * zd;
            const int depth = std::min(zd, n.z + NGHOST - zs);
            // planes marched: 6 warm-up planes plus the outputs
            const long cost = 6 + depth;
            for (int ty = 0; ty < nty; ++ty)
                for (int tx = 0; tx < ntx; ++tx)
                    items.push_back({{b, tx, ty, zs, zd}, cost});
        }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const Item& a, const Item& b) { return a.cost > b.cost; });
    std::vector<BatchItem> w(items.size());
    for (size_t i = 0; i < items.size(); ++i)
        w[i] = items[i].w;
    if (e->d_items)
        AC_HIP_CHECK(hipFree(e->d_items));
    e->d_items = nullptr;
    e->nitems  = (int)w.size();
    if (!w.empty()) {
        AC_HIP_CHECK(hipMalloc((void**)&e->d_items, w.size() * sizeof(BatchItem)));
        AC_HIP_CHECK(hipMemcpy(e->d_items, w.data(), w.size() * sizeof(BatchItem),
                               hipMemcpyHostToDevice));
    }
    // the persistent grid: one block per resident block slot of the batched kernel
    AC_HIP_CHECK(hipSetDevice(e->id));
    e->nblocks = batch_resident_blocks(TX, TY);
    if (e->nblocks < 1)
        return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
build_args(Ensemble e, AcReal dt)
{
    drop_graphs(e); // recorded graphs read these argument arrays
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    // (a plain array: std::vector cannot relocate a struct of __restrict__ pointer arrays)
    std::unique_ptr<SubstepArgs[]> a(new SubstepArgs[e->nboxes]);
    for (int s = 0; s < 3; ++s)
        for (int par = 0; par < 2; ++par) {
            // substep s at buffer parity `par` reads set `par` and writes the other set
            for (int b = 0; b < e->nboxes; ++b) {
                const int3 n    = dims_of(e->info[b]);
                SubstepArgs& x  = a[b];
                for (int q = 0; q < NF; ++q) {
                    x.in[q]   = e->fields[par][q][b];
                    x.out[q]  = e->fields[1 - par][q][b];
                    x.cin[q]  = nullptr;
                    x.cout[q] = nullptr;
                }
                x.mx    = n.x + 6;
                x.my    = n.y + 6;
                x.mz    = n.z + 6;
                x.mxy   = x.mx * x.my;
                x.start = make_int3(NGHOST, NGHOST, NGHOST);
                x.end   = make_int3(n.x + NGHOST, n.y + NGHOST, n.z + NGHOST);
                x.wrap  = 1;
                x.phys  = make_phys_params(e->info[b]);
                x.rk    = make_rk3_coeffs(s, dt);
            }
            AC_HIP_CHECK(hipMemcpy(e->d_args[s][par], a.get(), e->nboxes * sizeof(SubstepArgs),
                                   hipMemcpyHostToDevice));
        }
    e->dt_args    = dt;
    e->args_valid = true;
    return AC_SUCCESS;
}

AcResult
ensure_args(Ensemble e, AcReal dt)
{
    if (e->args_valid && e->dt_args == dt)
        return AC_SUCCESS;
    return build_args(e, dt);
}

hipError_t
enqueue_fill(Ensemble e, int par)
{
    if (e->fill_total == 0)
        return hipSuccess;
    const long blocks = e->fill_total / FILL_BLOCK;
    if (blocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    periodic_fill_batch<<<dim3((unsigned)blocks), dim3(FILL_BLOCK), 0, e->stream>>>(e->d_fill[par],
                                                                                    e->nboxes);
    return hipGetLastError();
}

hipError_t
enqueue_diagnostics(Ensemble e, int par)
{
    diagnostics_batch<<<dim3((unsigned)e->diag_blocks), dim3(256), 0, e->stream>>>(
        e->d_diag[par], e->nboxes, e->d_partials);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return err;
    const int n = e->nboxes * NDIAG;
    diagnostics_finalize<<<dim3((n + 255) / 256), dim3(256), 0, e->stream>>>(
        e->d_partials, e->d_diag[par], e->nboxes, e->diag_blocks, e->d_results);
    return hipGetLastError();
}

// The step at the current parity, enqueued on the ensemble stream (also what a graph records).
AcResult
enqueue_step(Ensemble e, int par0)
{
    // The batched solve reads periodic images instead of ghost cells, so the substeps need no
    // fill; one fill after the last substep leaves every box's ghosts current.
    int par = par0;
    for (int s = 0; s < 3; ++s) {
        AC_HIP_CHECK(launch_substep_batch(e->tile_x, e->tile_y, e->d_args[s][par], e->d_items,
                                          e->nitems, e->nblocks, e->stream));
        par = 1 - par;
    }
    AC_HIP_CHECK(enqueue_fill(e, par));
    AC_HIP_CHECK(enqueue_diagnostics(e, par));
    return AC_SUCCESS;
}

bool
valid_dt(AcReal dt)
{
    return std::isfinite((double)dt) && dt > 0;
}

// Shortest of `samples` launches of substep 0 of the batched solve over the current work list,
// after `warm` untimed ones. Substep 0 reads the current set and writes the other one, which
// holds no state between steps.
float
time_batch(Ensemble e, int warm, int samples)
{
    hipEvent_t t0, t1;
    AC_HIP_CHECK(hipEventCreate(&t0));
    AC_HIP_CHECK(hipEventCreate(&t1));
    const SubstepArgs* args = e->d_args[0][e->parity];
    for (int i = 0; i < warm; ++i)
        AC_HIP_CHECK(launch_substep_batch(e->tile_x, e->tile_y, args, e->d_items, e->nitems,
                                          e->nblocks, e->stream));
    float best = INFINITY;
    for (int i = 0; i < samples; ++i) {
        AC_HIP_CHECK(hipEventRecord(t0, e->stream));
        AC_HIP_CHECK(launch_substep_batch(e->tile_x, e->tile_y, args, e->d_items, e->nitems,
                                          e->nblocks, e->stream));
        AC_HIP_CHECK(hipEventRecord(t1, e->stream));
        AC_HIP_CHECK(hipEventSynchronize(t1));
        float ms = 0;
        AC_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
        best = std::min(best, ms);
    }
    AC_HIP_CHECK(hipEventDestroy(t0));
    AC_HIP_CHECK(hipEventDestroy(t1));
    return best;
}

// Tuned choices per ensemble composition (device, box sizes in order), kept for the process.
struct EnsembleTuning {
    int tile_x, tile_y;
    std::map<int, int> class_depth;
};

std::map<std::vector<int>, EnsembleTuning>&
tuning_cache()
{
    static std::map<std::vector<int>, EnsembleTuning> cache;
    return cache;
}

std::mutex&
tuning_mutex()
{
    static std::mutex m;
    return m;
}

} // namespace

extern "C" {

AcResult
acEnsembleCreate(const int device_id, const int nboxes, const AcMeshInfo* configs,
                 Ensemble* ensemble)
{
    if (!ensemble || !configs || nboxes < 1)
        return AC_FAILURE;
    *ensemble = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device_id < 0 || device_id >= count) {
        fprintf(stderr, "acEnsembleCreate: no HIP device %d\n", device_id);
        return AC_FAILURE;
    }
    for (int b = 0; b < nboxes; ++b) {
        const int3 n = dims_of(configs[b]);
        // The batched solve reads the periodic images of the ghost cells instead of filled
        // ghosts; an image lies in the interior only if every axis holds NGHOST cells or more.
        if (n.x < NGHOST || n.y < NGHOST || n.z < NGHOST) {
            fprintf(stderr, "acEnsembleCreate: box %d is %d x %d x %d; every axis needs at least %d cells\n",
                    b, n.x, n.y, n.z, NGHOST);
            return AC_FAILURE;
        }
    }
    AC_HIP_CHECK(hipSetDevice(device_id));
    Ensemble e   = new ensemble_s();
    e->id        = device_id;
    e->nboxes    = nboxes;
    e->info.assign(configs, configs + nboxes);
    for (auto& i : e->info)
        acHostUpdateBuiltinParams(&i);
    e->parity   = 0;
    // 16 x 16 tiles and whole columns, the fastest measured on the 96-box workload, until the
    // tuner (below) measures this composition
    e->tile_x   = 16;
    e->tile_y   = 16;
    e->z_depth  = 0;
    e->d_items  = nullptr;
    e->args_valid = false;
    e->dt_args  = 0;
    // one allocation: box after box, 16 fields each (8 per set), 256-byte aligned
    size_t bytes = 0;
    std::vector<size_t> offs(nboxes);
    for (int b = 0; b < nboxes; ++b) {
        offs[b] = bytes;
        const size_t fb = (cells_of(e->info[b]) * sizeof(AcReal) + 255) / 256 * 256;
        bytes += 2 * NF * fb;
    }
    e->base_bytes = bytes;
    AC_HIP_CHECK(hipMalloc((void**)&e->base, bytes));
    AC_HIP_CHECK(hipMemset(e->base, 0, bytes));
    for (int set = 0; set < 2; ++set)
        for (int q = 0; q < NF; ++q)
            e->fields[set][q].resize(nboxes);
    for (int b = 0; b < nboxes; ++b) {
        const size_t fb = (cells_of(e->info[b]) * sizeof(AcReal) + 255) / 256 * 256;
        for (int set = 0; set < 2; ++set)
            for (int q = 0; q < NF; ++q)
                e->fields[set][q][b] = (AcReal*)((char*)e->base + offs[b] + (set * NF + q) * fb);
    }
    AC_HIP_CHECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (int s = 0; s < 3; ++s)
        for (int p = 0; p < 2; ++p)
            AC_HIP_CHECK(hipMalloc((void**)&e->d_args[s][p], nboxes * sizeof(SubstepArgs)));
    // fill and diagnostics descriptors per parity
    std::vector<FillBox> fb(nboxes);
    std::vector<DiagBox> db(nboxes);
    int diag_blocks = 0;
    for (int p = 0; p < 2; ++p) {
        AC_HIP_CHECK(hipMalloc((void**)&e->d_fill[p], nboxes * sizeof(FillBox)));
        AC_HIP_CHECK(hipMalloc((void**)&e->d_diag[p], nboxes * sizeof(DiagBox)));
        long first  = 0;
        diag_blocks = 0; // the same for both parities
        for (int b = 0; b < nboxes; ++b) {
            const int3 n = dims_of(e->info[b]);
            for (int q = 0; q < NF; ++q) {
                fb[b].f[q] = e->fields[p][q][b];
                db[b].f[q] = e->fields[p][q][b];
            }
            fb[b].g     = GhostShell::of(n.x, n.y, n.z);
            if (!fb[b].g.fits_32bit()) {
                fprintf(stderr, "acEnsembleCreate: box %d is too large for the batched fill\n", b);
                return AC_FAILURE;
            }
            if ((long)n.x * n.y * n.z >= (1L << 31)) { // the diagnostics' 32-bit cell index
                fprintf(stderr, "acEnsembleCreate: box %d is too large for the batched diagnostics\n", b);
                return AC_FAILURE;
            }
            fb[b].first = first;
            first += (fb[b].g.cells() + FILL_BLOCK - 1) / FILL_BLOCK * FILL_BLOCK;
            db[b].nx  = n.x;
            db[b].ny  = n.y;
            db[b].nz  = n.z;
            db[b].mx  = n.x + 6;
            db[b].mxy = (long)(n.x + 6) * (n.y + 6);
            db[b].first_blk = diag_blocks;
            diag_blocks += (int)(((long)n.x * n.y * n.z + DIAG_CELLS - 1) / DIAG_CELLS);
        }
        e->fill_total = first;
        AC_HIP_CHECK(hipMemcpy(e->d_fill[p], fb.data(), nboxes * sizeof(FillBox),
                               hipMemcpyHostToDevice));
        AC_HIP_CHECK(hipMemcpy(e->d_diag[p], db.data(), nboxes * sizeof(DiagBox),
                               hipMemcpyHostToDevice));
    }
    e->diag_blocks = diag_blocks;
    AC_HIP_CHECK(hipMalloc((void**)&e->d_partials, (size_t)diag_blocks * NDIAG * sizeof(double)));
    AC_HIP_CHECK(hipMalloc((void**)&e->d_results, (size_t)nboxes * NDIAG * sizeof(double)));
    AC_HIP_CHECK(hipMemset(e->d_results, 0, (size_t)nboxes * NDIAG * sizeof(double)));
    AC_HIP_CHECK(hipHostMalloc((void**)&e->h_results, (size_t)nboxes * NDIAG * sizeof(double)));
    if (build_items(e) != AC_SUCCESS)
        return AC_FAILURE;
    // Tile shape and per-class z-depths measured on this composition, as Devices tune at
    // creation; AC_AUTOTUNE=0 keeps the defaults.
    const char* tune = getenv("AC_AUTOTUNE");
    if (!(tune && tune[0] == '0') && acEnsembleAutoOptimize(e) != AC_SUCCESS)
        return AC_FAILURE;
    AC_HIP_CHECK(hipDeviceSynchronize());
    *ensemble = e;
    return AC_SUCCESS;
}

AcResult
acEnsembleDestroy(Ensemble e)
{
    if (!e)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    drop_graphs(e);
    AC_HIP_CHECK(hipFree(e->base));
    for (int s = 0; s < 3; ++s)
        for (int p = 0; p < 2; ++p)
            AC_HIP_CHECK(hipFree(e->d_args[s][p]));
    for (int p = 0; p < 2; ++p) {
        AC_HIP_CHECK(hipFree(e->d_fill[p]));
        AC_HIP_CHECK(hipFree(e->d_diag[p]));
    }
    if (e->d_items)
        AC_HIP_CHECK(hipFree(e->d_items));
    AC_HIP_CHECK(hipFree(e->d_partials));
    AC_HIP_CHECK(hipFree(e->d_results));
    AC_HIP_CHECK(hipHostFree(e->h_results));
    AC_HIP_CHECK(hipStreamDestroy(e->stream));
    delete e;
    return AC_SUCCESS;
}

int
acEnsembleBoxCount(const Ensemble e)
{
    return e ? e->nboxes : 0;
}

AcResult
acEnsembleSetKernelConfig(Ensemble e, const AcKernelConfig config)
{
    if (!e || config.variant != 3 || !batch_tile_supported(config.tile_x, config.tile_y) ||
        config.z_depth < 0)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->tile_x  = config.tile_x;
    e->tile_y  = config.tile_y;
    e->z_depth = config.z_depth;
    e->class_depth.clear(); // one z-depth for every box
    drop_graphs(e);
    return build_items(e);
}

AcResult
acEnsembleAutoOptimize(Ensemble e)
{
    if (!e)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    drop_graphs(e);
    std::vector<int> key{e->id};
    for (const AcMeshInfo& i : e->info) {
        const int3 n = dims_of(i);
        key.insert(key.end(), {n.x, n.y, n.z});
    }
    {
        std::lock_guard<std::mutex> lock(tuning_mutex());
        const auto it = tuning_cache().find(key);
        if (it != tuning_cache().end()) {
            e->tile_x      = it->second.tile_x;
            e->tile_y      = it->second.tile_y;
            e->z_depth     = 0;
            e->class_depth = it->second.class_depth;
            return build_items(e);
        }
    }
    // the timed launches need argument arrays; any dt times the same
    if (!e->args_valid && build_args(e, AcReal(1e-3)) != AC_SUCCESS)
        return AC_FAILURE;
    // 1. The launch's tile shape, every box in whole columns.
    const int shapes[3][2] = {{16, 16}, {16, 8}, {32, 4}};
    e->z_depth = 0;
    e->class_depth.clear();
    float best = INFINITY;
    int bx = 16, by = 16;
    for (const auto& s : shapes) {
        e->tile_x = s[0];
        e->tile_y = s[1];
        if (build_items(e) != AC_SUCCESS)
            continue;
        const float ms = time_batch(e, 2, 3);
        if (ms < best) {
            best = ms;
            bx   = s[0];
            by   = s[1];
        }
    }
    e->tile_x = bx;
    e->tile_y = by;
    if (build_items(e) != AC_SUCCESS)
        return AC_FAILURE;
    // 2. Per size class (the column length n_z), the class with the most cells first: whole
    // columns, or columns cut into 2, 3 or 4 equal z ranges. A cut shortens the longest work
    // items (the launch's tail) at the cost of 6 warm-up planes per extra item; it is kept only
    // when the whole launch gets at least 2 % faster.
    std::map<int, double> work;
    for (const AcMeshInfo& i : e->info) {
        const int3 n = dims_of(i);
        work[n.z] += (double)n.x * n.y * n.z;
    }
    std::vector<std::pair<double, int>> classes;
    for (const auto& kv : work)
        classes.push_back({kv.second, kv.first});
    std::sort(classes.rbegin(), classes.rend());
    best = time_batch(e, 1, 3);
    for (const auto& c : classes) {
        const int nz = c.second;
        int keep     = 0;
        for (const int parts : {2, 3, 4}) {
            const int d = (nz + parts - 1) / parts;
            if (d < 8 || d >= nz)
                continue;
            e->class_depth[nz] = d;
            if (build_items(e) != AC_SUCCESS)
                return AC_FAILURE;
            const float ms = time_batch(e, 1, 3);
            if (ms < 0.98f * best) {
                best = ms;
                keep = d;
            }
        }
        if (keep > 0)
            e->class_depth[nz] = keep;
        else
            e->class_depth.erase(nz);
        if (build_items(e) != AC_SUCCESS)
            return AC_FAILURE;
    }
    std::lock_guard<std::mutex> lock(tuning_mutex());
    tuning_cache()[key] = EnsembleTuning{e->tile_x, e->tile_y, e->class_depth};
    return AC_SUCCESS;
}

int
acEnsembleBoxZDepth(const Ensemble e, const int box)
{
    if (!ens_ok(e, box))
        return -1;
    const int nz = dims_of(e->info[box]).z;
    const int zd = box_depth(e, nz);
    return zd >= nz ? 0 : zd;
}

AcResult
acEnsembleGetKernelConfig(const Ensemble e, AcKernelConfig* config)
{
    if (!e || !config)
        return AC_FAILURE;
    config->variant = 3;
    config->tile_x  = e->tile_x;
    config->tile_y  = e->tile_y;
    config->z_depth = e->z_depth;
    return AC_SUCCESS;
}

AcResult
acEnsembleRandomizeUniform(Ensemble e, const int box, const VertexBufferHandle handle,
                           const double lo, const double hi, const unsigned long long seed)
{
    if (!ens_ok(e, box) || handle < 0 || handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    AC_HIP_CHECK(launch_fill_uniform(e->fields[e->parity][handle][box], e->info[box], lo, hi,
                                     seed, e->stream));
    return AC_SUCCESS;
}

AcResult
acEnsembleSetVertexBuffer(Ensemble e, const int box, const VertexBufferHandle handle,
                          const AcReal value)
{
    if (!ens_ok(e, box) || handle < 0 || handle >= NF || !std::isfinite((double)value))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    const long n = (long)cells_of(e->info[box]);
    for (int set = 0; set < 2; ++set)
        AC_HIP_CHECK(launch_fill_constant(e->fields[set][handle][box], n, value, e->stream));
    return AC_SUCCESS;
}

AcResult
acEnsembleLoadMesh(Ensemble e, const int box, const AcMesh host)
{
    if (!ens_ok(e, box))
        return AC_FAILURE;
    const int3 n = dims_of(host.info);
    if (n.x != dims_of(e->info[box]).x || n.y != dims_of(e->info[box]).y ||
        n.z != dims_of(e->info[box]).z)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    const size_t bytes = cells_of(e->info[box]) * sizeof(AcReal);
    for (int q = 0; q < NF; ++q)
        AC_HIP_CHECK(hipMemcpyAsync(e->fields[e->parity][q][box], host.vertex_buffer[q], bytes,
                                    hipMemcpyHostToDevice, e->stream));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    return AC_SUCCESS;
}

AcResult
acEnsembleStoreMesh(Ensemble e, const int box, AcMesh* host)
{
    if (!ens_ok(e, box) || !host)
        return AC_FAILURE;
    const int3 n = dims_of(host->info);
    if (n.x != dims_of(e->info[box]).x || n.y != dims_of(e->info[box]).y ||
        n.z != dims_of(e->info[box]).z)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    const size_t bytes = cells_of(e->info[box]) * sizeof(AcReal);
    for (int q = 0; q < NF; ++q)
        AC_HIP_CHECK(hipMemcpyAsync(host->vertex_buffer[q], e->fields[e->parity][q][box], bytes,
                                    hipMemcpyDeviceToHost, e->stream));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    return AC_SUCCESS;
}

AcResult
acEnsemblePeriodicBoundconds(Ensemble e)
{
    if (!e)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    AC_HIP_CHECK(enqueue_fill(e, e->parity));
    return AC_SUCCESS;
}

AcResult
acEnsembleIntegrateStep(Ensemble e, const AcReal dt)
{
    if (!e || !valid_dt(dt))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    if (ensure_args(e, dt) != AC_SUCCESS || enqueue_step(e, e->parity) != AC_SUCCESS)
        return AC_FAILURE;
    e->parity = 1 - e->parity;
    return AC_SUCCESS;
}

AcResult
acEnsembleIntegrateStepGraph(Ensemble e, const AcReal dt)
{
    if (!e || !valid_dt(dt))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    if (ensure_args(e, dt) != AC_SUCCESS) // a new dt rebuilds the arguments and drops the graphs
        return AC_FAILURE;
    unsigned long long key = 0;
    memcpy(&key, &dt, sizeof(dt));
    auto& cache = e->graphs[e->parity];
    auto it     = cache.find(key);
    if (it == cache.end()) {
        AC_HIP_CHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        const AcResult r = enqueue_step(e, e->parity);
        hipGraph_t graph;
        AC_HIP_CHECK(hipStreamEndCapture(e->stream, &graph));
        if (r != AC_SUCCESS)
            return AC_FAILURE;
        hipGraphExec_t exec;
        AC_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        AC_HIP_CHECK(hipGraphDestroy(graph));
        it = cache.emplace(key, exec).first;
    }
    AC_HIP_CHECK(hipGraphLaunch(it->second, e->stream));
    e->parity = 1 - e->parity;
    return AC_SUCCESS;
}

AcResult
acEnsembleIntegrateStepLooped(Ensemble e, const AcReal dt)
{
    if (!e || !valid_dt(dt))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    int par = e->parity;
    for (int s = 0; s < 3; ++s) {
        for (int b = 0; b < e->nboxes; ++b) {
            const int3 n = dims_of(e->info[b]);
            AcReal* cin[NF];
            AcReal* cout[NF];
            for (int q = 0; q < NF; ++q) {
                cin[q]  = e->fields[par][q][b];
                cout[q] = e->fields[1 - par][q][b];
            }
            // each box as a single-box Device steps it: its default configuration, reading the
            // ghost cells' periodic images where that is the default form, else fill + launch
            const AcKernelConfig cfg = default_kernel_config(n.x, n.y, n.z);
            const bool images        = kernel_config_reads_images(cfg) &&
                                default_reads_images(n.x, n.y, n.z) && n.x >= NGHOST &&
                                n.y >= NGHOST && n.z >= NGHOST;
            const int3 zero = make_int3(0, 0, 0), m = make_int3(n.x + 6, n.y + 6, n.z + 6);
            if (!images)
                AC_HIP_CHECK(launch_periodic_fill(cin, NF, n.x, n.y, n.z, zero, m, e->stream));
            SubstepArgs a;
            for (int q = 0; q < NF; ++q) {
                a.in[q]   = cin[q];
                a.out[q]  = cout[q];
                a.cin[q]  = nullptr;
                a.cout[q] = nullptr;
            }
            a.mx    = n.x + 6;
            a.my    = n.y + 6;
            a.mz    = n.z + 6;
            a.mxy   = a.mx * a.my;
            a.start = make_int3(NGHOST, NGHOST, NGHOST);
            a.end   = make_int3(n.x + NGHOST, n.y + NGHOST, n.z + NGHOST);
            a.wrap  = images;
            a.phys  = make_phys_params(e->info[b]);
            a.rk    = make_rk3_coeffs(s, dt);
            AC_HIP_CHECK(launch_substep(cfg, a, e->stream));
        }
        par = 1 - par;
    }
    AC_HIP_CHECK(enqueue_diagnostics(e, par));
    e->parity = par;
    return AC_SUCCESS;
}

AcResult
acEnsembleDiagnostics(Ensemble e, double* results)
{
    if (!e || !results)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    const size_t n = (size_t)e->nboxes * NDIAG;
    AC_HIP_CHECK(hipMemcpyAsync(e->h_results, e->d_results, n * sizeof(double),
                                hipMemcpyDeviceToHost, e->stream));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    for (int b = 0; b < e->nboxes; ++b) {
        const int3 nn   = dims_of(e->info[b]);
        const double N  = (double)nn.x * nn.y * nn.z;
        const double* r = e->h_results + (size_t)b * NDIAG;
        double* o       = results + (size_t)b * NDIAG;
        o[0] = std::sqrt(r[0]);     // max |u|
        o[1] = std::sqrt(r[1] / N); // rms |u|
        o[2] = std::sqrt(r[2]);     // max |A|
        o[3] = std::sqrt(r[3] / N); // rms |A|
        o[4] = r[4];                // max lnrho
        o[5] = r[5];                // min lnrho
        o[6] = std::sqrt(r[6] / N); // rms lnrho
    }
    return AC_SUCCESS;
}

AcResult
acEnsembleTimestepCFL(Ensemble e, double* dts)
{
    if (!e || !dts)
        return AC_FAILURE;
    std::vector<double> d((size_t)e->nboxes * NDIAG);
    if (acEnsembleDiagnostics(e, d.data()) != AC_SUCCESS)
        return AC_FAILURE;
    for (int b = 0; b < e->nboxes; ++b)
        dts[b] = cfl_timestep(d[(size_t)b * NDIAG], 0.0, e->info[b]);
    return AC_SUCCESS;
}

AcResult
acEnsembleSynchronize(const Ensemble e)
{
    if (!e)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(e->id));
    AC_HIP_CHECK(hipStreamSynchronize(e->stream));
    return AC_SUCCESS;
}

void*
acEnsembleGetStreamHandle(const Ensemble e)
{
    return e ? (void*)e->stream : nullptr;
}

} // extern "C"
