// Device layer: the C ABI of include/astaroth.h for one MI355X.
//
// Replaces src/core/device.cc:1-537 and the uniform plumbing of src/core/kernels/kernels.cu:
// 102-251. Differences in mechanism (results are the same):
//  - Parameters ("uniforms") are kept in a host-side AcMeshInfo per Device and passed to every
//    kernel by value (PhysParams), instead of a __constant__ symbol updated with
//    cudaMemcpyToSymbolAsync. Each Device therefore has its own parameters, and launches can be
//    captured into HIP graphs.
//  - All NUM_STREAMS streams are created at the default priority.
//  - The autotuner result is per Device (cached per grid size for the process), not one static
//    block shape shared by every device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>
#include <vector>

#include "ac_internal.h"

using namespace ac;

namespace {

bool
valid_real(AcReal v)
{
    return !std::isnan(v) && !std::isinf(v);
}

bool
stream_ok(const Device d, Stream s)
{
    return d && s >= 0 && s < NUM_STREAMS;
}

int3
mk3(int x, int y, int z)
{
    int3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

int
nx_of(const Device d)
{
    return d->local_config.int_params[AC_nx];
}
int
ny_of(const Device d)
{
    return d->local_config.int_params[AC_ny];
}
int
nz_of(const Device d)
{
    return d->local_config.int_params[AC_nz];
}

// Parameters used by the kernels (the Device's copy of every uniform).
AcMeshInfo&
uniforms_of(const Device d)
{
    return d->uniforms;
}

void
invalidate_graphs(Device d)
{
    for (int par = 0; par < 2; ++par) {
        for (auto& kv : d->graphs[par])
            (void)hipGraphExecDestroy(kv.second);
        d->graphs[par].clear();
    }
}

// Autotuner results per (nx, ny, nz), shared by every Device of the process: the configuration
// and whether periodic steps read the ghost cells' images.
std::map<std::tuple<int, int, int>, std::pair<AcKernelConfig, bool>>&
tune_cache()
{
    static std::map<std::tuple<int, int, int>, std::pair<AcKernelConfig, bool>> cache;
    return cache;
}

SubstepArgs
make_args(const Device d, const AcReal* const* in, AcReal* const* out, int step, int3 start,
          int3 end, AcReal dt, int par = 0)
{
    // `par` = 1 when `in` is the device's current `out` set (inside a recorded step)
    SubstepArgs a;
    for (int q = 0; q < NF; ++q) {
        a.in[q]   = in[q];
        a.out[q]  = out[q];
        a.cin[q]  = d->comp[par][q];
        a.cout[q] = d->comp[1 - par][q];
    }
    a.mx    = nx_of(d) + 6;
    a.my    = ny_of(d) + 6;
    a.mz    = nz_of(d) + 6;
    a.mxy   = a.mx * a.my;
    a.start = start;
    a.end   = end;
    a.wrap  = 0;
    a.phys  = make_phys_params(uniforms_of(d));
    a.rk    = make_rk3_coeffs(step, dt);
    return a;
}

bool
region_ok(const Device d, int3 start, int3 end)
{
    const int n[3] = {nx_of(d), ny_of(d), nz_of(d)};
    const int s[3] = {start.x, start.y, start.z};
    const int e[3] = {end.x, end.y, end.z};
    for (int a = 0; a < 3; ++a)
        if (s[a] < NGHOST || e[a] > n[a] + NGHOST || s[a] > e[a])
            return false;
    return true;
}

float
time_config(Device d, const AcKernelConfig& c, int iters, bool wrap = false)
{
    const int3 start = mk3(NGHOST, NGHOST, NGHOST);
    const int3 end   = mk3(nx_of(d) + NGHOST, ny_of(d) + NGHOST, nz_of(d) + NGHOST);
    SubstepArgs a =
        make_args(d, d->vba.in, d->vba.out, 2, start, end, uniforms_of(d).real_params[AC_dt]);
    a.wrap = wrap; // the build that reads the ghost cells' periodic images
    hipStream_t s = d->streams[STREAM_DEFAULT];
    if (launch_substep(c, a, s) != hipSuccess) { // warm-up, also rejects unusable shapes
        (void)hipGetLastError();
        return -1.f;
    }
    hipEvent_t t0, t1;
    AC_HIP_CHECK(hipEventCreate(&t0));
    AC_HIP_CHECK(hipEventCreate(&t1));
    AC_HIP_CHECK(hipEventRecord(t0, s));
    for (int i = 0; i < iters; ++i)
        if (launch_substep(c, a, s) != hipSuccess)
            break;
    AC_HIP_CHECK(hipEventRecord(t1, s));
    AC_HIP_CHECK(hipEventSynchronize(t1));
    float ms = 0;
    AC_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
    AC_HIP_CHECK(hipEventDestroy(t0));
    AC_HIP_CHECK(hipEventDestroy(t1));
    return ms / (iters > 0 ? iters : 1);
}

// Mean time of one periodic fill of all fields, timed on the `out` buffers (scratch while tuning,
// as the timed launches' output is) so that `in` is unchanged.
float
time_fill(Device d, int iters)
{
    const int nx = nx_of(d), ny = ny_of(d), nz = nz_of(d);
    const int3 zero = mk3(0, 0, 0), m = mk3(nx + 6, ny + 6, nz + 6);
    hipStream_t s   = d->streams[STREAM_DEFAULT];
    hipEvent_t t0, t1;
    AC_HIP_CHECK(hipEventCreate(&t0));
    AC_HIP_CHECK(hipEventCreate(&t1));
    AC_HIP_CHECK(launch_periodic_fill(d->vba.out, NF, nx, ny, nz, zero, m, s));
    AC_HIP_CHECK(hipEventRecord(t0, s));
    for (int i = 0; i < iters; ++i)
        AC_HIP_CHECK(launch_periodic_fill(d->vba.out, NF, nx, ny, nz, zero, m, s));
    AC_HIP_CHECK(hipEventRecord(t1, s));
    AC_HIP_CHECK(hipEventSynchronize(t1));
    float ms = 0;
    AC_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
    AC_HIP_CHECK(hipEventDestroy(t0));
    AC_HIP_CHECK(hipEventDestroy(t1));
    return ms / (iters > 0 ? iters : 1);
}

AcResult
reduce_impl(const Device d, Stream stream, ReductionType rtype, int nvals, const int* handles,
            double* result)
{
    if (!stream_ok(d, stream)) {
        fprintf(stderr, "WARNING: invalid stream %d\n", stream);
        return AC_FAILURE;
    }
    bool ok;
    if (nvals == 4)
        ok = rtype == RTYPE_ALFVEN_MAX || rtype == RTYPE_ALFVEN_MIN || rtype == RTYPE_ALFVEN_RMS;
    else
        ok = rtype == RTYPE_MAX || rtype == RTYPE_MIN || rtype == RTYPE_RMS ||
             rtype == RTYPE_RMS_EXP || rtype == RTYPE_SUM;
    if (!ok)
        AC_FATAL("Unrecognized rtype %d for a %d-field reduction", (int)rtype, nvals);
    const AcReal* vals[4];
    for (int q = 0; q < nvals; ++q) {
        if (handles[q] < 0 || handles[q] >= NF) {
            fprintf(stderr, "WARNING: invalid VertexBufferHandle %d\n", handles[q]);
            return AC_FAILURE;
        }
        vals[q] = d->vba.in[handles[q]];
    }
    AC_HIP_CHECK(hipSetDevice(d->id));
    hipStream_t s = d->streams[stream];
    AC_HIP_CHECK(launch_reduce((int)rtype, nvals, vals, d->local_config, d->reduce_partials,
                               reduce_partials_needed(), d->reduce_result, s));
    AC_HIP_CHECK(hipMemcpyAsync(d->reduce_result_host, d->reduce_result, sizeof(double),
                                hipMemcpyDeviceToHost, s));
    AC_HIP_CHECK(hipStreamSynchronize(s));
    *result = *d->reduce_result_host;
    return AC_SUCCESS;
}

} // namespace

namespace ac {

PhysParams
make_phys_params(const AcMeshInfo& u)
{
    PhysParams p;
    const double inv[3] = {(double)u.real_params[AC_inv_dsx], (double)u.real_params[AC_inv_dsy],
                           (double)u.real_params[AC_inv_dsz]};
    // stencil_kernel.ac / stdderiv.h:13-84 weights (6th order)
    const double a1[4] = {0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
    const double b2[4] = {-49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};
    const double cc[4] = {0.0, 270.0 / 720.0, -27.0 / 720.0, 2.0 / 720.0};
    for (int ax = 0; ax < 3; ++ax)
        for (int i = 0; i < 4; ++i) {
            p.d1[ax][i] = (AcReal)(a1[i] * inv[ax]);
            p.d2[ax][i] = (AcReal)(b2[i] * inv[ax] * inv[ax]);
        }
    p.d2c = (AcReal)(b2[0] * (inv[0] * inv[0] + inv[1] * inv[1] + inv[2] * inv[2]));
    for (int i = 0; i < 4; ++i) {
        p.cxy[i] = (AcReal)(cc[i] * inv[0] * inv[1]);
        p.cxz[i] = (AcReal)(cc[i] * inv[0] * inv[2]);
        p.cyz[i] = (AcReal)(cc[i] * inv[1] * inv[2]);
    }
    const double cp = u.real_params[AC_cp_sound], gamma = u.real_params[AC_gamma];
    p.nu_visc      = u.real_params[AC_nu_visc];
    p.eta          = u.real_params[AC_eta];
    p.inv_mu0      = (AcReal)(1.0 / (double)u.real_params[AC_mu0]);
    p.eta_mu0      = (AcReal)((double)u.real_params[AC_eta] * (double)u.real_params[AC_mu0]);
    p.cp_sound     = (AcReal)cp;
    p.inv_cp       = (AcReal)(1.0 / cp);
    p.gamma        = (AcReal)gamma;
    p.gamma_m1     = (AcReal)(gamma - 1.0);
    p.gamma_inv_cp = (AcReal)(gamma / cp);
    p.lnT0         = u.real_params[AC_lnT0];
    p.exp_lnT0     = (AcReal)std::exp((double)u.real_params[AC_lnT0]);
    p.lnrho0       = u.real_params[AC_lnrho0];
    p.zeta         = u.real_params[AC_zeta];
    p.cs2_sound    = u.real_params[AC_cs2_sound];
    p.k_thermal    = (AcReal)0.001; // AC_THERMAL_CONDUCTIVITY, stencil_kernel.ac:15
    return p;
}

RK3Coeffs
make_rk3_coeffs(int step, AcReal dt)
{
    // Williamson (1980) 2N-storage coefficients (integration.cuh:20-21)
    const AcReal alpha[3] = {AcReal(0.), AcReal(-5. / 9.), AcReal(-153. / 128.)};
    const AcReal beta[3]  = {AcReal(1. / 3.), AcReal(15. / 16.), AcReal(8. / 15.)};
    RK3Coeffs rk;
    rk.step    = step;
    rk.dt      = dt;
    rk.beta    = beta[step];
    rk.ab      = step > 0 ? alpha[step] * (AcReal(1.) / beta[step - 1]) : AcReal(0);
    rk.beta_dt = beta[step] * dt;
    return rk;
}

} // namespace ac

extern "C" {

int
acAbiRealSize(size_t* meshinfo_size, size_t* mesh_size)
{
    if (meshinfo_size)
        *meshinfo_size = sizeof(AcMeshInfo);
    if (mesh_size)
        *mesh_size = sizeof(AcMesh);
    return (int)sizeof(AcReal);
}

AcResult
acDevicePrintInfo(const Device device)
{
    hipDeviceProp_t props;
    AC_HIP_CHECK(hipGetDeviceProperties(&props, device->id));
    size_t free_bytes = 0, total_bytes = 0;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
    printf("--------------------------------------------------\n");
    printf("Device Number: %d\n", device->id);
    printf("    Device name: %s (%s)\n", props.name, props.gcnArchName);
    printf("  Compute\n");
    printf("    Clock rate (GHz): %g\n", props.clockRate / 1e6);
    printf("    Compute units: %d\n", props.multiProcessorCount);
    printf("  Global memory\n");
    printf("    Memory Clock Rate (MHz): %d\n", props.memoryClockRate / 1000);
    printf("    Memory Bus Width (bits): %d\n", props.memoryBusWidth);
    printf("    Total global mem: %.2f GiB\n", props.totalGlobalMem / (1024.0 * 1024 * 1024));
    printf("    Gmem used (GiB): %.2f\n", (total_bytes - free_bytes) / (1024.0 * 1024 * 1024));
    printf("  Caches\n");
    printf("    L2 size: %d KiB\n", props.l2CacheSize / 1024);
    printf("    LDS per block: %zu KiB\n", props.sharedMemPerBlock / 1024);
    printf("  Other\n");
    printf("    Wavefront size: %d\n", props.warpSize);
    printf("    Precision: %s\n", sizeof(AcReal) == 8 ? "FP64" : "FP32");
    printf("    Solve kernel: variant %d, tile (%d, %d), z-depth %d\n",
           device->kernel_config.variant, device->kernel_config.tile_x,
           device->kernel_config.tile_y, device->kernel_config.z_depth);
    printf("--------------------------------------------------\n");
    return AC_SUCCESS;
}

AcResult
acDeviceSynchronizeStream(const Device device, const Stream stream)
{
    AC_HIP_CHECK(hipSetDevice(device->id));
    if (stream == STREAM_ALL) {
        AC_HIP_CHECK(hipDeviceSynchronize());
        return AC_SUCCESS;
    }
    if (!stream_ok(device, stream)) {
        fprintf(stderr, "WARNING: invalid stream %d\n", stream);
        return AC_FAILURE;
    }
    AC_HIP_CHECK(hipStreamSynchronize(device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceLoadDefaultUniforms(const Device device)
{
    // DSL defaults: every uniform without an initialiser is 0 (code_generator.c:345-346,565-569);
    // AC_ds* = 0.04908738521, AC_inv_ds* = 1/AC_ds* (stdderiv.h:6-11); AC_cs_sound = 1 and
    // AC_cs2_sound = AC_cs_sound^2 (stencil_kernel.ac:69,112).
    AcMeshInfo& u = uniforms_of(device);
    memset(&u, 0, sizeof(u));
    const AcReal ds = (AcReal)0.04908738521;
    u.real_params[AC_dsx] = u.real_params[AC_dsy] = u.real_params[AC_dsz] = ds;
    u.real_params[AC_inv_dsx] = u.real_params[AC_inv_dsy] = u.real_params[AC_inv_dsz] =
        (AcReal)(1.0 / (double)ds);
    u.real_params[AC_cs_sound]  = (AcReal)1.0;
    u.real_params[AC_cs2_sound] = (AcReal)1.0;
    invalidate_graphs(device);
    return AC_SUCCESS;
}

AcResult
acDeviceLoadScalarUniform(const Device device, const Stream stream, const AcRealParam param,
                          const AcReal value)
{
    (void)stream;
    if (param < 0 || param >= NUM_REAL_PARAMS) {
        fprintf(stderr, "WARNING: invalid AcRealParam %d.\n", param);
        return AC_FAILURE;
    }
    if (!valid_real(value)) {
        fprintf(stderr, "WARNING: Passed an invalid value %g to device constant %s. Skipping.\n",
                (double)value, realparam_names[param]);
        return AC_FAILURE;
    }
    uniforms_of(device).real_params[param] = value;
    if (param != AC_dt)
        invalidate_graphs(device);
    return AC_SUCCESS;
}

AcResult
acDeviceLoadVectorUniform(const Device device, const Stream stream, const AcReal3Param param,
                          const AcReal3 value)
{
    (void)device;
    (void)stream;
    (void)value;
    fprintf(stderr, "WARNING: invalid AcReal3Param %d\n", (int)param); // the module has none
    return AC_FAILURE;
}

AcResult
acDeviceLoadIntUniform(const Device device, const Stream stream, const AcIntParam param,
                       const int value)
{
    (void)stream;
    if (param < 0 || param >= NUM_INT_PARAMS) {
        fprintf(stderr, "WARNING: invalid AcIntParam %d\n", param);
        return AC_FAILURE;
    }
    uniforms_of(device).int_params[param] = value;
    return AC_SUCCESS;
}

AcResult
acDeviceLoadInt3Uniform(const Device device, const Stream stream, const AcInt3Param param,
                        const int3 value)
{
    (void)stream;
    if (param < 0 || param >= NUM_INT3_PARAMS) {
        fprintf(stderr, "WARNING: invalid AcInt3Param %d\n", param);
        return AC_FAILURE;
    }
    uniforms_of(device).int3_params[param] = value;
    return AC_SUCCESS;
}

AcResult
acDeviceLoadScalarArray(const Device device, const Stream stream, const ScalarArrayHandle handle,
                        const size_t start, const AcReal* data, const size_t num)
{
    (void)device;
    (void)stream;
    (void)start;
    (void)data;
    (void)num;
    (void)handle;
    return AC_FAILURE; // NUM_SCALARARRAY_HANDLES == 0 (device.cc:230-231)
}

AcResult
acDeviceLoadMeshInfo(const Device device, const AcMeshInfo config)
{
    if (config.int_params[AC_nx] != device->local_config.int_params[AC_nx] ||
        config.int_params[AC_ny] != device->local_config.int_params[AC_ny] ||
        config.int_params[AC_nz] != device->local_config.int_params[AC_nz])
        AC_FATAL("acDeviceLoadMeshInfo: grid size differs from the one the Device was created with");
    for (int i = 0; i < NUM_INT_PARAMS; ++i)
        acDeviceLoadIntUniform(device, STREAM_DEFAULT, (AcIntParam)i, config.int_params[i]);
    for (int i = 0; i < NUM_INT3_PARAMS; ++i)
        acDeviceLoadInt3Uniform(device, STREAM_DEFAULT, (AcInt3Param)i, config.int3_params[i]);
    for (int i = 0; i < NUM_REAL_PARAMS; ++i) {
        const AcReal v = config.real_params[i];
        if (valid_real(v)) // invalid (unset) values keep the current value, silently
            uniforms_of(device).real_params[i] = v;
    }
    // The reference keeps the DSL default AC_cs2_sound = 1 when the configuration leaves it
    // unset, while its host model solver always uses cs_sound^2 (modelsolver.c:996-997). Derive
    // it here so both sides use the same sound speed.
    if (!valid_real(config.real_params[AC_cs2_sound]) && valid_real(config.real_params[AC_cs_sound]))
        uniforms_of(device).real_params[AC_cs2_sound] =
            config.real_params[AC_cs_sound] * config.real_params[AC_cs_sound];
    invalidate_graphs(device);
    return AC_SUCCESS;
}

AcResult
acDeviceCreate(const int id, const AcMeshInfo config, Device* device_handle)
{
    const int nx = config.int_params[AC_nx], ny = config.int_params[AC_ny],
              nz = config.int_params[AC_nz];
    if (nx < 1 || ny < 1 || nz < 1) {
        fprintf(stderr, "acDeviceCreate: invalid grid %d x %d x %d\n", nx, ny, nz);
        return AC_FAILURE;
    }
    // The buffers hold (n + 2 NGHOST) cells per axis, while loads, stores and fills size their
    // copies from AC_mx/my/mz (acVertexBufferSize, acVertexBufferIdx): a configuration whose
    // built-in sizes are unset or stale would copy past the allocation or copy nothing.
    if (config.int_params[AC_mx] != nx + 2 * NGHOST || config.int_params[AC_my] != ny + 2 * NGHOST ||
        config.int_params[AC_mz] != nz + 2 * NGHOST ||
        config.int_params[AC_mxy] != config.int_params[AC_mx] * config.int_params[AC_my]) {
        fprintf(stderr,
                "acDeviceCreate: AC_mx, AC_my, AC_mz = %d, %d, %d (AC_mxy %d) do not match the grid "
                "%d x %d x %d with %d ghost cells per side; call acHostUpdateBuiltinParams first\n",
                config.int_params[AC_mx], config.int_params[AC_my], config.int_params[AC_mz],
                config.int_params[AC_mxy], nx, ny, nz, NGHOST);
        return AC_FAILURE;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        fprintf(stderr, "acDeviceCreate: no HIP device is available\n");
        return AC_FAILURE;
    }
    if (id < 0 || id >= count) {
        fprintf(stderr, "acDeviceCreate: no GPU with id %d (%d visible)\n", id, count);
        return AC_FAILURE;
    }
    AC_HIP_CHECK(hipSetDevice(id));
    hipDeviceProp_t props;
    AC_HIP_CHECK(hipGetDeviceProperties(&props, id));
    if (strncmp(props.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "acDeviceCreate: device %d is %s; this library is built for gfx950\n", id,
                props.gcnArchName);
        return AC_FAILURE;
    }

    device_s* d     = new device_s();
    d->id           = id;
    d->local_config = config;
    d->parity       = 0;
    d->halo_buffer     = nullptr;
    d->halo_buffer_len = 0;
    for (int q = 0; q < NF; ++q)
        d->comp[0][q] = d->comp[1][q] = nullptr;
    for (int i = 0; i < NUM_STREAMS; ++i)
        AC_HIP_CHECK(hipStreamCreateWithFlags(&d->streams[i], hipStreamNonBlocking));

    const size_t m     = (size_t)(nx + 6) * (ny + 6) * (nz + 6);
    const size_t bytes = m * sizeof(AcReal);
    for (int q = 0; q < NF; ++q) {
        AC_HIP_CHECK(hipMalloc((void**)&d->vba.in[q], bytes));
        AC_HIP_CHECK(hipMalloc((void**)&d->vba.out[q], bytes));
        AC_HIP_CHECK(launch_fill_constant(d->vba.in[q], (long)m, AcReal(0), d->streams[0]));
        AC_HIP_CHECK(launch_fill_constant(d->vba.out[q], (long)m, AcReal(0), d->streams[0]));
        d->base_in[q] = d->vba.in[q];
    }
    AC_HIP_CHECK(hipMalloc((void**)&d->reduce_partials, sizeof(double) * reduce_partials_needed()));
    AC_HIP_CHECK(hipMalloc((void**)&d->reduce_result, sizeof(double)));
    AC_HIP_CHECK(hipHostMalloc((void**)&d->reduce_result_host, sizeof(double), hipHostMallocDefault));
    AC_HIP_CHECK(hipMalloc((void**)&d->diag_partials, sizeof(double) * diagnostics_partials_needed()));
    AC_HIP_CHECK(hipMalloc((void**)&d->diag_result, 9 * sizeof(double)));
    AC_HIP_CHECK(hipHostMalloc((void**)&d->diag_result_host, 9 * sizeof(double), hipHostMallocDefault));
    AC_HIP_CHECK(hipStreamSynchronize(d->streams[0]));
    // this GPU's copy of the reduction scratch array (acDeviceReduceNoAlloc), with `id` current
    AC_HIP_CHECK(reduce_scratch_address(id, &d->reduce_scratch));

    acDeviceLoadDefaultUniforms(d);
    acDeviceLoadMeshInfo(d, config);
    d->kernel_config = default_kernel_config(nx, ny, nz);
    d->read_images   = default_reads_images(nx, ny, nz);
    *device_handle   = d;

    const char* tune = getenv("AC_AUTOTUNE");
    if (!(tune && tune[0] == '0'))
        acDeviceAutoOptimize(d);
    return AC_SUCCESS;
}

AcResult
acDeviceDestroy(Device device)
{
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(hipDeviceSynchronize());
    invalidate_graphs(device);
    for (int q = 0; q < NF; ++q) {
        AC_HIP_CHECK(hipFree(device->vba.in[q]));
        AC_HIP_CHECK(hipFree(device->vba.out[q]));
    }
    AC_HIP_CHECK(hipFree(device->reduce_partials));
    AC_HIP_CHECK(hipFree(device->reduce_result));
    AC_HIP_CHECK(hipHostFree(device->reduce_result_host));
    AC_HIP_CHECK(hipFree(device->diag_partials));
    AC_HIP_CHECK(hipFree(device->diag_result));
    AC_HIP_CHECK(hipHostFree(device->diag_result_host));
    if (device->halo_buffer)
        AC_HIP_CHECK(hipFree(device->halo_buffer));
    for (int set = 0; set < 2; ++set)
        for (int q = 0; q < NF; ++q)
            if (device->comp[set][q])
                AC_HIP_CHECK(hipFree(device->comp[set][q]));
    for (int i = 0; i < NUM_STREAMS; ++i)
        AC_HIP_CHECK(hipStreamDestroy(device->streams[i]));
    delete device;
    return AC_SUCCESS;
}

AcResult
acDeviceSwapBuffers(const Device device)
{
    for (int q = 0; q < NF; ++q) {
        std::swap(device->vba.in[q], device->vba.out[q]);
        std::swap(device->comp[0][q], device->comp[1][q]);
    }
    device->parity ^= 1;
    return AC_SUCCESS;
}

// Values written into `in` from outside the integrator are exact: their compensation terms
// (compensated integration only) restart at zero.
static hipError_t
clear_compensation(const Device d, int handle, hipStream_t s)
{
    if (!d->comp[0][handle])
        return hipSuccess;
    const long n = (long)acVertexBufferSize(d->local_config);
    for (int set = 0; set < 2; ++set) {
        const hipError_t e = launch_fill_constant(d->comp[set][handle], n, AcReal(0), s);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

AcResult
acDeviceLoadVertexBufferWithOffset(const Device device, const Stream stream, const AcMesh host_mesh,
                                   const VertexBufferHandle vtxbuf_handle, const int3 src,
                                   const int3 dst, const int num_vertices)
{
    if (!stream_ok(device, stream) || vtxbuf_handle < 0 || vtxbuf_handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    const size_t src_idx = acVertexBufferIdx(src.x, src.y, src.z, host_mesh.info);
    const size_t dst_idx = acVertexBufferIdx(dst.x, dst.y, dst.z, device->local_config);
    if (dst_idx + (size_t)num_vertices > acVertexBufferSize(device->local_config))
        AC_FATAL("load of %d vertices at %zu overruns the device buffer", num_vertices, dst_idx);
    AC_HIP_CHECK(hipMemcpyAsync(device->vba.in[vtxbuf_handle] + dst_idx,
                                host_mesh.vertex_buffer[vtxbuf_handle] + src_idx,
                                sizeof(AcReal) * (size_t)num_vertices, hipMemcpyHostToDevice,
                                device->streams[stream]));
    AC_HIP_CHECK(clear_compensation(device, vtxbuf_handle, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceLoadMeshWithOffset(const Device device, const Stream stream, const AcMesh host_mesh,
                           const int3 src, const int3 dst, const int num_vertices)
{
    for (int q = 0; q < NF; ++q)
        if (acDeviceLoadVertexBufferWithOffset(device, stream, host_mesh, (VertexBufferHandle)q,
                                               src, dst, num_vertices) != AC_SUCCESS)
            return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
acDeviceLoadVertexBuffer(const Device device, const Stream stream, const AcMesh host_mesh,
                         const VertexBufferHandle vtxbuf_handle)
{
    const int3 zero = mk3(0, 0, 0);
    return acDeviceLoadVertexBufferWithOffset(device, stream, host_mesh, vtxbuf_handle, zero, zero,
                                              (int)acVertexBufferSize(device->local_config));
}

AcResult
acDeviceLoadMesh(const Device device, const Stream stream, const AcMesh host_mesh)
{
    for (int q = 0; q < NF; ++q)
        if (acDeviceLoadVertexBuffer(device, stream, host_mesh, (VertexBufferHandle)q) != AC_SUCCESS)
            return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
acDeviceSetVertexBuffer(const Device device, const Stream stream, const VertexBufferHandle handle,
                        const AcReal value)
{
    if (!stream_ok(device, stream) || handle < 0 || handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    const long n = (long)acVertexBufferSize(device->local_config);
    AC_HIP_CHECK(launch_fill_constant(device->vba.in[handle], n, value, device->streams[stream]));
    AC_HIP_CHECK(launch_fill_constant(device->vba.out[handle], n, value, device->streams[stream]));
    AC_HIP_CHECK(clear_compensation(device, handle, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceStoreVertexBufferWithOffset(const Device device, const Stream stream,
                                    const VertexBufferHandle vtxbuf_handle, const int3 src,
                                    const int3 dst, const int num_vertices, AcMesh* host_mesh)
{
    if (!stream_ok(device, stream) || vtxbuf_handle < 0 || vtxbuf_handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    const size_t src_idx = acVertexBufferIdx(src.x, src.y, src.z, device->local_config);
    const size_t dst_idx = acVertexBufferIdx(dst.x, dst.y, dst.z, host_mesh->info);
    if (src_idx + (size_t)num_vertices > acVertexBufferSize(device->local_config))
        AC_FATAL("store of %d vertices at %zu overruns the device buffer", num_vertices, src_idx);
    AC_HIP_CHECK(hipMemcpyAsync(host_mesh->vertex_buffer[vtxbuf_handle] + dst_idx,
                                device->vba.in[vtxbuf_handle] + src_idx,
                                sizeof(AcReal) * (size_t)num_vertices, hipMemcpyDeviceToHost,
                                device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceStoreMeshWithOffset(const Device device, const Stream stream, const int3 src,
                            const int3 dst, const int num_vertices, AcMesh* host_mesh)
{
    for (int q = 0; q < NF; ++q)
        if (acDeviceStoreVertexBufferWithOffset(device, stream, (VertexBufferHandle)q, src, dst,
                                                num_vertices, host_mesh) != AC_SUCCESS)
            return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
acDeviceStoreVertexBuffer(const Device device, const Stream stream,
                          const VertexBufferHandle vtxbuf_handle, AcMesh* host_mesh)
{
    const int3 zero = mk3(0, 0, 0);
    return acDeviceStoreVertexBufferWithOffset(device, stream, vtxbuf_handle, zero, zero,
                                               (int)acVertexBufferSize(device->local_config),
                                               host_mesh);
}

AcResult
acDeviceStoreMesh(const Device device, const Stream stream, AcMesh* host_mesh)
{
    for (int q = 0; q < NF; ++q)
        if (acDeviceStoreVertexBuffer(device, stream, (VertexBufferHandle)q, host_mesh) != AC_SUCCESS)
            return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
acDeviceTransferVertexBufferWithOffset(const Device src_device, const Stream stream,
                                       const VertexBufferHandle vtxbuf_handle, const int3 src,
                                       const int3 dst, const int num_vertices, Device dst_device)
{
    if (!stream_ok(src_device, stream) || !dst_device || vtxbuf_handle < 0 || vtxbuf_handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(src_device->id));
    const size_t src_idx = acVertexBufferIdx(src.x, src.y, src.z, src_device->local_config);
    const size_t dst_idx = acVertexBufferIdx(dst.x, dst.y, dst.z, dst_device->local_config);
    AC_HIP_CHECK(hipMemcpyPeerAsync(dst_device->vba.in[vtxbuf_handle] + dst_idx, dst_device->id,
                                    src_device->vba.in[vtxbuf_handle] + src_idx, src_device->id,
                                    sizeof(AcReal) * (size_t)num_vertices,
                                    src_device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceTransferMeshWithOffset(const Device src_device, const Stream stream, const int3 src,
                               const int3 dst, const int num_vertices, Device* dst_device)
{
    for (int q = 0; q < NF; ++q)
        if (acDeviceTransferVertexBufferWithOffset(src_device, stream, (VertexBufferHandle)q, src,
                                                   dst, num_vertices, *dst_device) != AC_SUCCESS)
            return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
acDeviceTransferVertexBuffer(const Device src_device, const Stream stream,
                             const VertexBufferHandle vtxbuf_handle, Device dst_device)
{
    const int3 zero = mk3(0, 0, 0);
    return acDeviceTransferVertexBufferWithOffset(src_device, stream, vtxbuf_handle, zero, zero,
                                                  (int)acVertexBufferSize(src_device->local_config),
                                                  dst_device);
}

AcResult
acDeviceTransferMesh(const Device src_device, const Stream stream, Device dst_device)
{
    for (int q = 0; q < NF; ++q)
        if (acDeviceTransferVertexBuffer(src_device, stream, (VertexBufferHandle)q, dst_device) !=
            AC_SUCCESS)
            return AC_FAILURE;
    return AC_SUCCESS;
}

AcResult
acDeviceIntegrateSubstep(const Device device, const Stream stream, const int step_number,
                         const int3 start, const int3 end, const AcReal dt)
{
    if (step_number < 0 || step_number > 2)
        AC_FATAL("acDeviceIntegrateSubstep: step_number %d is not 0, 1 or 2", step_number);
    if (!stream_ok(device, stream)) {
        fprintf(stderr, "WARNING: invalid stream %d\n", stream);
        return AC_FAILURE;
    }
    if (!region_ok(device, start, end)) {
        fprintf(stderr,
                "acDeviceIntegrateSubstep: region [(%d,%d,%d), (%d,%d,%d)) is outside the "
                "computational domain\n",
                start.x, start.y, start.z, end.x, end.y, end.z);
        return AC_FAILURE;
    }
    if (acDeviceLoadScalarUniform(device, stream, AC_dt, dt) != AC_SUCCESS)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    const SubstepArgs a = make_args(device, device->vba.in, device->vba.out, step_number, start,
                                    end, dt);
    AC_HIP_CHECK(launch_substep(device->kernel_config, a, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDevicePeriodicBoundcondStep(const Device device, const Stream stream,
                              const VertexBufferHandle vtxbuf_handle, const int3 start,
                              const int3 end)
{
    if (!stream_ok(device, stream) || vtxbuf_handle < 0 || vtxbuf_handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AcReal* f[1] = {device->vba.in[vtxbuf_handle]};
    AC_HIP_CHECK(launch_periodic_fill(f, 1, nx_of(device), ny_of(device), nz_of(device), start,
                                      end, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDevicePeriodicBoundconds(const Device device, const Stream stream, const int3 start,
                           const int3 end)
{
    if (!stream_ok(device, stream))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_periodic_fill(device->vba.in, NF, nx_of(device), ny_of(device),
                                      nz_of(device), start, end, device->streams[stream]));
    return AC_SUCCESS;
}

// ---- halo pack / unpack (packing.cuh:3-92; Grid-layer blocks device.cc:1375-1381,1758-1818) ----

static bool
block_ok(const Device d, int3 start, int3 dims)
{
    const int m[3] = {nx_of(d) + 2 * NGHOST, ny_of(d) + 2 * NGHOST, nz_of(d) + 2 * NGHOST};
    const int s[3] = {start.x, start.y, start.z};
    const int e[3] = {dims.x, dims.y, dims.z};
    for (int a = 0; a < 3; ++a)
        if (s[a] < 0 || e[a] < 0 || s[a] + e[a] > m[a])
            return false;
    return true;
}

AcResult
acDevicePackData(const Device device, const Stream stream, const int3 vba_start, const int3 dims,
                 AcReal* packed)
{
    if (!stream_ok(device, stream) || !block_ok(device, vba_start, dims) || !packed)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_pack(true, device->vba.in, nx_of(device) + 2 * NGHOST,
                             ny_of(device) + 2 * NGHOST, packed, 1, &vba_start, &dims,
                             device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceUnpackData(const Device device, const Stream stream, const AcReal* packed,
                   const int3 vba_start, const int3 dims)
{
    if (!stream_ok(device, stream) || !block_ok(device, vba_start, dims) || !packed)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_pack(false, device->vba.in, nx_of(device) + 2 * NGHOST,
                             ny_of(device) + 2 * NGHOST, const_cast<AcReal*>(packed), 1,
                             &vba_start, &dims, device->streams[stream]));
    return AC_SUCCESS;
}

// The 26 blocks need n >= NGHOST on every axis so that each ghost block's periodic image is one
// interior block (as in the reference Grid layer); smaller boxes get 0 blocks and AC_FAILURE.
static int
device_halo_blocks(const Device d, int3* ghost, int3* interior, int3* dims)
{
    if (nx_of(d) < NGHOST || ny_of(d) < NGHOST || nz_of(d) < NGHOST)
        return 0;
    return halo_blocks(make_int3(nx_of(d), ny_of(d), nz_of(d)), ghost, interior, dims);
}

size_t
acDeviceHaloPackedSize(const Device device)
{
    int3 g[26], in[26], dims[26];
    const int nb = device_halo_blocks(device, g, in, dims);
    size_t cells = 0;
    for (int i = 0; i < nb; ++i)
        cells += (size_t)dims[i].x * dims[i].y * dims[i].z;
    return cells * NF;
}

AcResult
acDevicePackHalo(const Device device, const Stream stream, AcReal* packed)
{
    if (!stream_ok(device, stream) || !packed)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    int3 g[26], in[26], dims[26];
    const int nb = device_halo_blocks(device, g, in, dims);
    if (nb == 0)
        return AC_FAILURE;
    AC_HIP_CHECK(launch_pack(true, device->vba.in, nx_of(device) + 2 * NGHOST,
                             ny_of(device) + 2 * NGHOST, packed, nb, in, dims,
                             device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceUnpackHalo(const Device device, const Stream stream, const AcReal* packed)
{
    if (!stream_ok(device, stream) || !packed)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    int3 g[26], in[26], dims[26];
    const int nb = device_halo_blocks(device, g, in, dims);
    if (nb == 0)
        return AC_FAILURE;
    AC_HIP_CHECK(launch_pack(false, device->vba.in, nx_of(device) + 2 * NGHOST,
                             ny_of(device) + 2 * NGHOST, const_cast<AcReal*>(packed), nb, g,
                             dims, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceHaloExchangeSelf(const Device device, const Stream stream)
{
    if (!stream_ok(device, stream))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    const size_t len = acDeviceHaloPackedSize(device);
    if (len == 0)
        return AC_FAILURE;
    if (device->halo_buffer_len < len) {
        if (device->halo_buffer) {
            AC_HIP_CHECK(hipStreamSynchronize(device->streams[stream]));
            AC_HIP_CHECK(hipFree(device->halo_buffer));
        }
        AC_HIP_CHECK(hipMalloc((void**)&device->halo_buffer, len * sizeof(AcReal)));
        device->halo_buffer_len = len;
    }
    if (acDevicePackHalo(device, stream, device->halo_buffer) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceUnpackHalo(device, stream, device->halo_buffer);
}

// ref device.cc:457-475 over boundconds.cuh:192-242; the face types come from `config`
// (AC_bc_type_{bot,top}_{x,y,z}). bindex = (-1, -1, -1) is the single-device form the Node layer
// passes (node.cc:672-690): this Device owns every face. The reference's multi-process forms
// (bindex >= 0, disabled there) are not provided and return AC_FAILURE, as do unknown face
// types and mirrored axes shorter than 4 cells (their mirror images would be ghost cells).
static bool
general_bc_faces(const Device device, const AcMeshInfo& config, const int3 bindex, int bot[3],
                 int top[3])
{
    if (bindex.x >= 0 || bindex.y >= 0 || bindex.z >= 0) {
        fprintf(stderr, "acDeviceGeneralBoundcond*: only bindex = (-1, -1, -1) is provided\n");
        return false;
    }
    const AcIntParam b[3] = {AC_bc_type_bot_x, AC_bc_type_bot_y, AC_bc_type_bot_z};
    const AcIntParam t[3] = {AC_bc_type_top_x, AC_bc_type_top_y, AC_bc_type_top_z};
    const int n[3]        = {nx_of(device), ny_of(device), nz_of(device)};
    for (int d = 0; d < 3; ++d) {
        bot[d] = config.int_params[b[d]];
        top[d] = config.int_params[t[d]];
        for (int v : {bot[d], top[d]}) {
            if (v != AC_BOUNDCOND_PERIODIC && v != AC_BOUNDCOND_SYMMETRIC &&
                v != AC_BOUNDCOND_ANTISYMMETRIC) {
                fprintf(stderr, "acDeviceGeneralBoundcond*: unknown boundary type %d\n", v);
                return false;
            }
            if (v != AC_BOUNDCOND_PERIODIC && n[d] < 4)
                return false;
        }
        if (n[d] < NGHOST)
            return false;
    }
    return true;
}

AcResult
acDeviceGeneralBoundcondStep(const Device device, const Stream stream,
                             const VertexBufferHandle vtxbuf_handle, const int3 start,
                             const int3 end, const AcMeshInfo config, const int3 bindex)
{
    int bot[3], top[3];
    if (!stream_ok(device, stream) || vtxbuf_handle < 0 || vtxbuf_handle >= NF ||
        !general_bc_faces(device, config, bindex, bot, top))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AcReal* f[1] = {device->vba.in[vtxbuf_handle]};
    AC_HIP_CHECK(launch_general_fill(f, 1, nx_of(device), ny_of(device), nz_of(device), start,
                                     end, bot, top, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceGeneralBoundconds(const Device device, const Stream stream, const int3 start,
                          const int3 end, const AcMeshInfo config, const int3 bindex)
{
    int bot[3], top[3];
    if (!stream_ok(device, stream) || !general_bc_faces(device, config, bindex, bot, top))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_general_fill(device->vba.in, NF, nx_of(device), ny_of(device),
                                     nz_of(device), start, end, bot, top,
                                     device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceReduceScal(const Device device, const Stream stream, const ReductionType rtype,
                   const VertexBufferHandle vtxbuf_handle, AcReal* result)
{
    double r = 0;
    const int h[1] = {(int)vtxbuf_handle};
    const AcResult res = reduce_impl(device, stream, rtype, 1, h, &r);
    *result = (AcReal)r;
    return res;
}

AcResult
acDeviceReduceVec(const Device device, const Stream stream, const ReductionType rtype,
                  const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                  const VertexBufferHandle vtxbuf2, AcReal* result)
{
    double r = 0;
    const int h[3] = {(int)vtxbuf0, (int)vtxbuf1, (int)vtxbuf2};
    const AcResult res = reduce_impl(device, stream, rtype, 3, h, &r);
    *result = (AcReal)r;
    return res;
}

AcResult
acDeviceReduceVecScal(const Device device, const Stream stream, const ReductionType rtype,
                      const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                      const VertexBufferHandle vtxbuf2, const VertexBufferHandle vtxbuf3,
                      AcReal* result)
{
    double r = 0;
    const int h[4] = {(int)vtxbuf0, (int)vtxbuf1, (int)vtxbuf2, (int)vtxbuf3};
    const AcResult res = reduce_impl(device, stream, rtype, 4, h, &r);
    *result = (AcReal)r;
    return res;
}

AcResult
acDeviceReduceScalDouble(const Device device, const Stream stream, const ReductionType rtype,
                         const VertexBufferHandle vtxbuf, double* result)
{
    const int h[1] = {(int)vtxbuf};
    return reduce_impl(device, stream, rtype, 1, h, result);
}

AcResult
acDeviceReduceVecDouble(const Device device, const Stream stream, const ReductionType rtype,
                        const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                        const VertexBufferHandle vtxbuf2, double* result)
{
    const int h[3] = {(int)vtxbuf0, (int)vtxbuf1, (int)vtxbuf2};
    return reduce_impl(device, stream, rtype, 3, h, result);
}

AcResult
acDeviceReduceVecScalDouble(const Device device, const Stream stream, const ReductionType rtype,
                            const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                            const VertexBufferHandle vtxbuf2, const VertexBufferHandle vtxbuf3,
                            double* result)
{
    const int h[4] = {(int)vtxbuf0, (int)vtxbuf1, (int)vtxbuf2, (int)vtxbuf3};
    return reduce_impl(device, stream, rtype, 4, h, result);
}

// The reductions of acDeviceReduce{Scal,Vec,VecScal}Double with the same two launches, their
// partials in the library's per-GPU scratch array instead of the Device's allocation, and the
// FP64 result written to the caller's device memory without waiting: nothing is allocated and
// nothing blocks the host. Any grid size fits: the row walk is grid-stride and the first launch
// writes at most reduce_scratch_capacity() partials.
AcResult
acDeviceReduceNoAlloc(const Device device, const Stream stream, const ReductionType rtype,
                      const int num_fields, const VertexBufferHandle* fields, double* d_result)
{
    if (!device || !fields || !d_result) {
        fprintf(stderr, "acDeviceReduceNoAlloc: null device, field list or result pointer\n");
        return AC_FAILURE;
    }
    if (num_fields != 1 && num_fields != 3 && num_fields != 4) {
        fprintf(stderr, "acDeviceReduceNoAlloc: %d fields; 1 (scalar), 3 (vector) or 4 (Alfven) expected\n",
                num_fields);
        return AC_FAILURE;
    }
    if (!stream_ok(device, stream)) {
        fprintf(stderr, "acDeviceReduceNoAlloc: invalid stream %d\n", stream);
        return AC_FAILURE;
    }
    const bool alfven = rtype == RTYPE_ALFVEN_MAX || rtype == RTYPE_ALFVEN_MIN || rtype == RTYPE_ALFVEN_RMS;
    const bool plain  = rtype == RTYPE_MAX || rtype == RTYPE_MIN || rtype == RTYPE_RMS ||
                       rtype == RTYPE_RMS_EXP || rtype == RTYPE_SUM;
    if (num_fields == 4 ? !alfven : !plain) {
        fprintf(stderr, "acDeviceReduceNoAlloc: reduction type %d does not apply to %d field(s)\n",
                (int)rtype, num_fields);
        return AC_FAILURE;
    }
    const AcReal* vals[4];
    for (int q = 0; q < num_fields; ++q) {
        if (fields[q] < 0 || fields[q] >= NF) {
            fprintf(stderr, "acDeviceReduceNoAlloc: invalid VertexBufferHandle %d\n", (int)fields[q]);
            return AC_FAILURE;
        }
        vals[q] = device->vba.in[fields[q]];
    }
    AC_HIP_CHECK(hipSetDevice(device->id));
    const hipError_t e = launch_reduce((int)rtype, num_fields, vals, device->local_config,
                                       device->reduce_scratch, reduce_scratch_capacity(), d_result,
                                       device->streams[stream]);
    if (e != hipSuccess) {
        fprintf(stderr, "acDeviceReduceNoAlloc: launch failed: %s\n", hipGetErrorString(e));
        return AC_FAILURE;
    }
    return AC_SUCCESS;
}

AcResult
acDeviceRunMPITest(void)
{
    fprintf(stderr, "acDeviceRunMPITest: this library has no MPI layer\n");
    return AC_FAILURE;
}

// The step's substeps read the periodic images of the ghost cells instead of filling them
// first: the configuration has such a build, compensated integration is off and the interior
// spans at least one ghost width on each axis.
static bool
step_reads_images(const Device d)
{
    if (!d->read_images || d->comp[0][0] != nullptr || nx_of(d) < NGHOST || ny_of(d) < NGHOST ||
        nz_of(d) < NGHOST)
        return false;
    return kernel_config_reads_images(d->kernel_config);
}

AcResult
acDeviceSetReadPeriodicImages(const Device device, const int enable)
{
    if (enable && !kernel_config_reads_images(device->kernel_config))
        return AC_FAILURE;
    device->read_images = enable != 0;
    invalidate_graphs(device);
    return AC_SUCCESS;
}

AcResult
acDeviceIntegrateSubstepPeriodic(const Device device, const Stream stream, const int step_number,
                                 const AcReal dt)
{
    const int3 nmin = mk3(NGHOST, NGHOST, NGHOST);
    const int3 nmax = mk3(nx_of(device) + NGHOST, ny_of(device) + NGHOST, nz_of(device) + NGHOST);
    if (!step_reads_images(device)) {
        const int3 m = mk3(nx_of(device) + 6, ny_of(device) + 6, nz_of(device) + 6);
        if (acDevicePeriodicBoundconds(device, stream, mk3(0, 0, 0), m) != AC_SUCCESS)
            return AC_FAILURE;
        return acDeviceIntegrateSubstep(device, stream, step_number, nmin, nmax, dt);
    }
    if (step_number < 0 || step_number > 2)
        AC_FATAL("acDeviceIntegrateSubstepPeriodic: step_number %d is not 0, 1 or 2", step_number);
    if (!stream_ok(device, stream) ||
        acDeviceLoadScalarUniform(device, stream, AC_dt, dt) != AC_SUCCESS)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    SubstepArgs a = make_args(device, device->vba.in, device->vba.out, step_number, nmin, nmax, dt);
    a.wrap        = 1;
    AC_HIP_CHECK(launch_substep(device->kernel_config, a, device->streams[stream]));
    return AC_SUCCESS;
}

int
acDeviceSubstepReadsPeriodicImages(const Device device)
{
    return step_reads_images(device) ? 1 : 0;
}

AcResult
acDeviceIntegrateStep(const Device device, const Stream stream, const AcReal dt)
{
    const int3 zero = mk3(0, 0, 0);
    const int3 m    = mk3(nx_of(device) + 6, ny_of(device) + 6, nz_of(device) + 6);
    const int3 nmin = mk3(NGHOST, NGHOST, NGHOST);
    const int3 nmax = mk3(nx_of(device) + NGHOST, ny_of(device) + NGHOST, nz_of(device) + NGHOST);
    if (step_reads_images(device)) {
        // The build reads the periodic images of the ghost cells itself: the substeps need no
        // fill between them and compute what they compute after one.
        for (int s = 0; s < 3; ++s) {
            if (acDeviceIntegrateSubstepPeriodic(device, stream, s, dt) != AC_SUCCESS)
                return AC_FAILURE;
            acDeviceSwapBuffers(device);
        }
        return AC_SUCCESS;
    }
    for (int s = 0; s < 3; ++s) {
        if (acDevicePeriodicBoundconds(device, stream, zero, m) != AC_SUCCESS ||
            acDeviceIntegrateSubstep(device, stream, s, nmin, nmax, dt) != AC_SUCCESS)
            return AC_FAILURE;
        acDeviceSwapBuffers(device);
    }
    return AC_SUCCESS;
}

// The step of acNodeIntegrateGBC (node.cc:808-880) on one Device: 3 x {general boundary fill by
// the face types of `config`, substep over the interior, swap}. With every face periodic it is
// acDeviceIntegrateStep.
AcResult
acDeviceIntegrateStepGBC(const Device device, const Stream stream, const AcMeshInfo config,
                         const AcReal dt)
{
    int bot[3], top[3];
    if (!general_bc_faces(device, config, mk3(-1, -1, -1), bot, top))
        return AC_FAILURE;
    bool periodic = true;
    for (int d = 0; d < 3; ++d)
        periodic = periodic && bot[d] == AC_BOUNDCOND_PERIODIC && top[d] == AC_BOUNDCOND_PERIODIC;
    if (periodic)
        return acDeviceIntegrateStep(device, stream, dt);
    const int3 zero  = mk3(0, 0, 0);
    const int3 m     = mk3(nx_of(device) + 6, ny_of(device) + 6, nz_of(device) + 6);
    const int3 nmin  = mk3(NGHOST, NGHOST, NGHOST);
    const int3 nmax  = mk3(nx_of(device) + NGHOST, ny_of(device) + NGHOST, nz_of(device) + NGHOST);
    const int3 owner = mk3(-1, -1, -1);
    for (int s = 0; s < 3; ++s) {
        if (acDeviceGeneralBoundconds(device, stream, zero, m, config, owner) != AC_SUCCESS ||
            acDeviceIntegrateSubstep(device, stream, s, nmin, nmax, dt) != AC_SUCCESS)
            return AC_FAILURE;
        acDeviceSwapBuffers(device);
    }
    return AC_SUCCESS;
}

AcResult
acDeviceIntegrateStepGraph(const Device device, const Stream stream, const AcReal dt)
{
    if (!stream_ok(device, stream) || !valid_real(dt))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    unsigned long long dt_bits = 0;
    memcpy(&dt_bits, &dt, sizeof(dt));
    const auto key = std::make_pair((int)stream, dt_bits);
    auto& cache    = device->graphs[device->parity];
    auto it        = cache.find(key);
    hipStream_t s  = device->streams[stream];
    if (it == cache.end()) {
        uniforms_of(device).real_params[AC_dt] = dt;
        const int nx = nx_of(device), ny = ny_of(device), nz = nz_of(device);
        const int3 zero = mk3(0, 0, 0), m = mk3(nx + 6, ny + 6, nz + 6);
        const int3 nmin = mk3(NGHOST, NGHOST, NGHOST);
        const int3 nmax = mk3(nx + NGHOST, ny + NGHOST, nz + NGHOST);
        AcReal* bufs[2][NF];
        for (int q = 0; q < NF; ++q) {
            bufs[0][q] = device->vba.in[q];
            bufs[1][q] = device->vba.out[q];
        }
        const bool images = step_reads_images(device);
        AC_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int st = 0; st < 3; ++st) {
            AcReal* const* cin  = bufs[st & 1];
            AcReal* const* cout = bufs[(st + 1) & 1];
            if (!images)
                AC_HIP_CHECK(launch_periodic_fill(cin, NF, nx, ny, nz, zero, m, s));
            SubstepArgs a = make_args(device, cin, cout, st, nmin, nmax, dt, st & 1);
            a.wrap        = images;
            AC_HIP_CHECK(launch_substep(device->kernel_config, a, s));
        }
        hipGraph_t graph;
        AC_HIP_CHECK(hipStreamEndCapture(s, &graph));
        hipGraphExec_t exec;
        AC_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        AC_HIP_CHECK(hipGraphDestroy(graph));
        it = cache.emplace(key, exec).first;
    }
    AC_HIP_CHECK(hipGraphLaunch(it->second, s));
    for (int st = 0; st < 3; ++st)
        acDeviceSwapBuffers(device);
    return AC_SUCCESS;
}

int
acKernelConfigCount(void)
{
    return kernel_config_count();
}

AcResult
acKernelConfigGet(const int index, AcKernelConfig* config)
{
    if (index < 0 || index >= kernel_config_count())
        return AC_FAILURE;
    *config = kernel_config_at(index);
    return AC_SUCCESS;
}

AcResult
acDeviceGetKernelConfig(const Device device, AcKernelConfig* config)
{
    *config = device->kernel_config;
    return AC_SUCCESS;
}

AcResult
acDeviceSetKernelConfig(const Device device, const AcKernelConfig config)
{
    if (!kernel_config_supported(config))
        return AC_FAILURE;
    device->kernel_config = config;
    invalidate_graphs(device);
    return AC_SUCCESS;
}

float
acDeviceTimeKernelConfig(const Device device, const AcKernelConfig config, const int iters)
{
    if (!kernel_config_supported(config))
        return -1.f;
    AC_HIP_CHECK(hipSetDevice(device->id));
    return time_config(device, config, iters);
}

// ref device.cc:83-98 / integration.cuh:130-215. Sweeps every compiled tile shape and, for the
// z-march variants, z-depths below n_z plus n_z itself; keeps the fastest substep. Overwrites the
// `out` buffers (as the reference's sweep does), so call it between full RK3 steps.
AcResult
acDeviceAutoOptimize(const Device device)
{
    AC_HIP_CHECK(hipSetDevice(device->id));
    const auto key = std::make_tuple(nx_of(device), ny_of(device), nz_of(device));
    {
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
            device->kernel_config = it->second.first;
            device->read_images   = it->second.second;
            invalidate_graphs(device);
            return AC_SUCCESS;
        }
    }
    const long cells = (long)nx_of(device) * ny_of(device) * nz_of(device);
    const int iters  = cells >= (1L << 24) ? 3 : 5;
    AcKernelConfig best = default_kernel_config(nx_of(device), ny_of(device), nz_of(device));
    float best_ms       = INFINITY;
    // untimed launches first, so that the first configuration timed does not pay for the clocks
    // ramping up from idle
    (void)time_config(device, best, 4 * iters);
    // z-depths: powers of two for variant 1, whole rounds of the unrolled 8-plane loop for
    // variant 13, and for the other lean marches (whose plane loop ends with a partial round)
    // equal chunks: the box's depth cut in 2, 3, 4 or 8
    const int nz        = nz_of(device);
    const int depths1[] = {8, 16, 32, 64, 128, 0};
    const int depths7[] = {(nz + 1) / 2, (nz + 2) / 3, (nz + 3) / 4, (nz + 7) / 8, 0};
    const int depths8[] = {16, 32, 64, 128, 0};
    for (int i = 0; i < kernel_config_count(); ++i) {
        AcKernelConfig c = kernel_config_at(i);
        if (c.tile_x > 2 * nx_of(device) || c.tile_y > 2 * ny_of(device))
            continue;
        const int* depths = c.variant == 1 ? depths1 : c.variant == 13 ? depths8 : depths7;
        const int ndepth  = c.variant == 0 ? 1 : c.variant == 1 ? 6 : 5;
        for (int di = 0; di < ndepth; ++di) {
            c.z_depth = c.variant == 0 ? 0 : depths[di];
            if (c.z_depth >= nz_of(device))
                continue;
            const float ms = time_config(device, c, iters);
            if (ms > 0 && ms < best_ms) {
                best_ms = ms;
                best    = c;
            }
        }
    }
    if (best.variant == 1 && best.z_depth == 0)
        best.z_depth = nz_of(device);
    // A periodic substep costs a fill plus the launch, or one launch of a build that reads the
    // ghost cells' periodic images (the defaults have one): keep the cheaper of the fastest
    // configuration with its fill and the image-reading builds of it and of the default.
    bool images = false;
    if (nx_of(device) >= NGHOST && ny_of(device) >= NGHOST && nz_of(device) >= NGHOST) {
        float best_substep = best_ms + time_fill(device, iters);
        const AcKernelConfig candidates[2] = {
            best, default_kernel_config(nx_of(device), ny_of(device), nz_of(device))};
        for (const AcKernelConfig& c : candidates) {
            if (!kernel_config_reads_images(c))
                continue;
            const float ms = time_config(device, c, iters, true);
            if (ms > 0 && ms < best_substep) {
                best_substep = ms;
                best         = c;
                images       = true;
            }
        }
    }
    device->kernel_config = best;
    device->read_images   = images;
    tune_cache()[key]     = std::make_pair(best, images);
    invalidate_graphs(device);
    return AC_SUCCESS;
}

AcResult
acDeviceRandomizeUniform(const Device device, const Stream stream, const VertexBufferHandle handle,
                         const double lo, const double hi, const unsigned long long seed)
{
    if (!stream_ok(device, stream) || handle < 0 || handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_fill_uniform(device->vba.in[handle], device->local_config, lo, hi, seed,
                                     device->streams[stream]));
    AC_HIP_CHECK(clear_compensation(device, handle, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceRandomizeUniformSinZ(const Device device, const Stream stream,
                             const VertexBufferHandle handle, const double lo, const double hi,
                             const unsigned long long seed, const double amplitude,
                             const double phase_per_cell)
{
    if (!stream_ok(device, stream) || handle < 0 || handle >= NF || !std::isfinite(amplitude) ||
        !std::isfinite(phase_per_cell))
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_fill_uniform(device->vba.in[handle], device->local_config, lo, hi, seed,
                                     device->streams[stream], amplitude, phase_per_cell));
    AC_HIP_CHECK(clear_compensation(device, handle, device->streams[stream]));
    return AC_SUCCESS;
}

AcResult
acDeviceTimestepCFL(const Device device, const Stream stream, double* dt)
{
    if (!dt)
        return AC_FAILURE;
    double umax = 0;
    if (acDeviceReduceVecDouble(device, stream, RTYPE_MAX, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ,
                                &umax) != AC_SUCCESS)
        return AC_FAILURE;
    *dt = cfl_timestep(umax, 0.0, uniforms_of(device));
    return AC_SUCCESS;
}

AcResult
acDeviceSetCompensatedIntegration(const Device device, const int enable)
{
    if (!device)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(hipDeviceSynchronize());
    const bool on = device->comp[0][0] != nullptr;
    if (enable && !on) {
        const size_t bytes = acVertexBufferSizeBytes(device->local_config);
        for (int set = 0; set < 2; ++set)
            for (int q = 0; q < NF; ++q) {
                AC_HIP_CHECK(hipMalloc((void**)&device->comp[set][q], bytes));
                AC_HIP_CHECK(hipMemset(device->comp[set][q], 0, bytes));
            }
    }
    else if (!enable && on) {
        for (int set = 0; set < 2; ++set)
            for (int q = 0; q < NF; ++q) {
                AC_HIP_CHECK(hipFree(device->comp[set][q]));
                device->comp[set][q] = nullptr;
            }
    }
    invalidate_graphs(device);
    return AC_SUCCESS;
}

AcResult
acDeviceCopyVertexBuffer(const Device device, const Stream stream, const VertexBufferHandle handle)
{
    if (!stream_ok(device, stream) || handle < 0 || handle >= NF)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    AC_HIP_CHECK(launch_copy(device->vba.out[handle], device->vba.in[handle],
                             (long)acVertexBufferSize(device->local_config), device->streams[stream]));
    return AC_SUCCESS;
}

void*
acDeviceGetStreamHandle(const Device device, const Stream stream)
{
    if (!stream_ok(device, stream))
        return nullptr;
    return (void*)device->streams[stream];
}

AcResult
acDeviceDiagnostics(const Device device, const Stream stream, double* out)
{
    if (!stream_ok(device, stream) || !out)
        return AC_FAILURE;
    AC_HIP_CHECK(hipSetDevice(device->id));
    hipStream_t s = device->streams[stream];
    const AcReal* f[NF];
    for (int q = 0; q < NF; ++q)
        f[q] = device->vba.in[q];
    AC_HIP_CHECK(launch_diagnostics(f, device->local_config, device->diag_partials,
                                    device->diag_result, s));
    AC_HIP_CHECK(hipMemcpyAsync(device->diag_result_host, device->diag_result, 9 * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    AC_HIP_CHECK(hipStreamSynchronize(s));
    const double n = (double)nx_of(device) * ny_of(device) * nz_of(device);
    for (int i = 0; i < 9; ++i)
        out[i] = i % 3 == 2 ? std::sqrt(device->diag_result_host[i] / n) : device->diag_result_host[i];
    return AC_SUCCESS;
}

AcResult
acDeviceGetMeshInfo(const Device device, AcMeshInfo* info)
{
    if (!device || !info)
        return AC_FAILURE;
    *info = device->uniforms;
    return AC_SUCCESS;
}

} // extern "C"
