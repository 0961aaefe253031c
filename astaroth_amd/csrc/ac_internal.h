// Internal declarations shared by the translation units of libastaroth_{fp32,fp64}.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <utility>

#include "astaroth.h"
#include "astaroth_utils.h"

#define AC_HIP_CHECK(call)                                                                         \
    do {                                                                                           \
        hipError_t err_ = (call);                                                                  \
        if (err_ != hipSuccess) {                                                                  \
            fprintf(stderr, "HIP error %s (%d) at %s:%d: %s\n", hipGetErrorString(err_),          \
                    (int)err_, __FILE__, __LINE__, #call);                                         \
            abort();                                                                               \
        }                                                                                          \
    } while (0)

#define AC_FATAL(...)                                                                              \
    do {                                                                                           \
        fprintf(stderr, "astaroth fatal error at %s:%d: ", __FILE__, __LINE__);                    \
        fprintf(stderr, __VA_ARGS__);                                                              \
        fprintf(stderr, "\n");                                                                     \
        abort();                                                                                   \
    } while (0)

namespace ac {

constexpr int NF = NUM_VTXBUF_HANDLES; // 8 evolved fields

struct VertexBufferArray {
    AcReal* in[NF];
    AcReal* out[NF];
};

// Everything the solve kernel reads besides the buffers: passed by value at every launch, so
// a captured graph and concurrent Devices each carry their own copy.
struct PhysParams {
    // FD coefficients with the grid spacing folded in:
    //   first  : d1[a][i] = a_i / ds_a
    //   second : d2[a][i] = b_i / ds_a^2
    //   cross  : cxy[i] = c_i / (dsx dsy), cxz, cyz likewise
    AcReal d1[3][4];
    AcReal d2[3][4];
    AcReal d2c; // b_0 (1/dsx^2 + 1/dsy^2 + 1/dsz^2): centre weight of a Laplacian
    AcReal cxy[4], cxz[4], cyz[4];
    // physics
    AcReal nu_visc, eta, inv_mu0, eta_mu0, cp_sound, inv_cp, gamma, gamma_m1, gamma_inv_cp;
    AcReal lnT0, exp_lnT0, lnrho0, zeta, cs2_sound, k_thermal;
};

// RK3 substep coefficients (integration.cuh:14-37 form): for s = 0
//   out = in + beta_dt * F
// for s > 0
//   out = in + beta_s * (ab * (in - prev) + dt * F),  ab = alpha_s / beta_{s-1}
struct RK3Coeffs {
    int step;
    AcReal dt;
    AcReal beta;
    AcReal ab;
    AcReal beta_dt;
};

struct SubstepArgs {
    const AcReal* __restrict__ in[NF];
    AcReal* __restrict__ out[NF];
    // Compensated integration only (null otherwise): rounding-error terms of `in` (read) and of
    // `out` (read as the previous stage's, then overwritten with the new ones).
    const AcReal* __restrict__ cin[NF];
    AcReal* __restrict__ cout[NF];
    int mx, my, mz, mxy;
    int3 start, end;
    // 1: the tile halos and the planes beyond the interior are read at their periodic images
    // inside the interior, so no ghost cell is read (configurations of kernel_config_reads_images)
    int wrap;
    PhysParams phys;
    RK3Coeffs rk;
};

// One work item of the batched solve: tile column (btx, bty) of box `box`, output planes
// [zs, zs + zdepth) clipped to the box.
struct BatchItem {
    int box, btx, bty, zs, zdepth;
};

PhysParams make_phys_params(const AcMeshInfo& info);
RK3Coeffs make_rk3_coeffs(int step, AcReal dt);

// Solve-kernel launch (solve.hip). Returns hipSuccess or the launch error; `config` must be
// one of the compiled configurations.
hipError_t launch_substep(const AcKernelConfig& config, const SubstepArgs& args, hipStream_t stream);
bool kernel_config_supported(const AcKernelConfig& config);
// True when `config` has a build that reads periodic images (SubstepArgs::wrap = 1).
bool kernel_config_reads_images(const AcKernelConfig& config);
// Whether a periodic step of an n_x x n_y x n_z box reads the images rather than filling the ghost
// cells, before the autotuner has measured both (the measured choice for the default builds).
bool default_reads_images(int nx, int ny, int nz);
// Batched solve (persistent lean z-march over a work list): `boxes` and `items` are device
// arrays; tile shapes 32x4, 16x8, 16x16.
bool batch_tile_supported(int tile_x, int tile_y);
// Blocks of the batched solve build for this tile that are resident at once on the current
// device (occupancy per CU x CUs), or -1 on error.
int batch_resident_blocks(int tile_x, int tile_y);
hipError_t launch_substep_batch(int tile_x, int tile_y, const SubstepArgs* boxes,
                                const BatchItem* items, int nitems, int nblocks,
                                hipStream_t stream);
int kernel_config_count();
AcKernelConfig kernel_config_at(int index);
AcKernelConfig default_kernel_config(int nx, int ny, int nz);
// z-depth of equal z-chunks for a z-march of tile_x x tile_y columns with `blocks_per_cu` blocks
// resident per CU (0: whole columns), chosen by the launch's rounds of resident blocks.
int balanced_z_depth(int nx, int ny, int nz, int tile_x, int tile_y, int blocks_per_cu);

// Periodic fill of the ghost cells inside [start, end) of `nfields` buffers (boundconds.hip).
hipError_t launch_periodic_fill(AcReal* const* fields, int nfields, int nx, int ny, int nz,
                                int3 start, int3 end, hipStream_t stream);
// Ghost cells of [start, end) by face type (AC_BOUNDCOND_*; bot/top per axis x, y, z).
hipError_t launch_general_fill(AcReal* const* fields, int nfields, int nx, int ny, int nz,
                               int3 start, int3 end, const int bot[3], const int top[3],
                               hipStream_t stream);

// Reductions (reductions.hip). `vals` are 1, 3 or 4 field pointers. Writes the FP64 result to
// *d_result (device) and returns the launch status.
// One-pass diagnostics of the `in` fields: max, min and sum of squares of |u|, |A| and lnrho
// over the interior, FP64, into d_result[9] (reductions.hip).
int diagnostics_partials_needed();
hipError_t launch_diagnostics(const AcReal* const* fields, const AcMeshInfo& info,
                              double* d_partials, double* d_result, hipStream_t stream);
hipError_t launch_reduce(int rtype, int nvals, const AcReal* const* vals, const AcMeshInfo& info,
                         double* d_partials, int max_partials, double* d_result,
                         hipStream_t stream);
int reduce_partials_needed();
// Partials capacity (FP64 values) of the library's per-GPU reduction scratch array, and its
// device address on `device_id`, which must be the current device (looked up once per device).
int reduce_scratch_capacity();
hipError_t reduce_scratch_address(int device_id, double** out);

// Device-side PRNG fill (init.hip): value = lo + (hi - lo) * u (+ amp sin(kdz (k - 3))),
// u = splitmix64(seed, cell) as a double in [0, 1), over the interior [n_min, n_max) of one
// buffer.
hipError_t launch_fill_uniform(AcReal* field, const AcMeshInfo& info, double lo, double hi,
                               unsigned long long seed, hipStream_t stream, double amp = 0.0,
                               double kdz = 0.0);
hipError_t launch_fill_constant(AcReal* field, long n, AcReal value, hipStream_t stream);
// Halo pack / unpack (halo.hip): moves the eight fields of `nblocks` blocks (origins, dims)
// to (pack) or from (unpack) a buffer laid out block after block, each [f][z][y][x].
hipError_t launch_pack(bool pack, AcReal* const* fields, int mx, int my, AcReal* packed,
                       int nblocks, const int3* origins, const int3* dims, hipStream_t stream);
// The 26 ghost blocks of an n-cell box (8 corners, 12 edges, 6 sides) with their periodic
// interior images; returns the block count.
int halo_blocks(int3 n, int3* ghost_origin, int3* interior_origin, int3* dims);
// CFL time step of the reference's standalone driver from the maximum speed and Alfven speed
// (host.cc).
double cfl_timestep(double umax, double vamax, const AcMeshInfo& info);
// Plain element copy (one read and one write per element).
hipError_t launch_copy(AcReal* dst, const AcReal* src, long n, hipStream_t stream);

} // namespace ac

struct device_s {
    int id;
    AcMeshInfo local_config; // creation-time configuration: grid sizes and allocation
    AcMeshInfo uniforms;     // parameters the kernels use (the reference's d_mesh_info)
    hipStream_t streams[NUM_STREAMS];
    ac::VertexBufferArray vba;
    double* reduce_partials;
    double* reduce_scratch; // this GPU's copy of the library's reduction scratch array
    double* reduce_result;
    double* reduce_result_host; // pinned
    double* diag_partials;      // acDeviceDiagnostics: per-block partials, results, pinned copy
    double* diag_result;
    double* diag_result_host;
    AcKernelConfig kernel_config;
    // periodic steps read the ghost cells' images where the configuration has such a build
    // (default by size, measured by acDeviceAutoOptimize, set by acDeviceSetReadPeriodicImages)
    bool read_images;
    // compensated integration (acDeviceSetCompensatedIntegration): comp[0] pairs with vba.in,
    // comp[1] with vba.out; null when off
    AcReal* comp[2][ac::NF];
    // acDeviceIntegrateStepGraph cache: (stream, dt bits) -> (graph exec, which in/out parity
    // the graph was recorded for)
    std::map<std::pair<int, unsigned long long>, hipGraphExec_t> graphs[2];
    int parity; // number of swaps mod 2 since creation
    AcReal* halo_buffer;     // acDeviceHaloExchangeSelf's pack buffer, allocated on first use
    size_t halo_buffer_len;  // in AcReal elements
    AcReal* base_in[ac::NF];
};
