/*
 * astaroth_ext.h — additions of the MI355X library that have no counterpart in the reference
 * header. Reference signatures are untouched; every symbol here is new.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ABI self-description (lets bindings check struct layouts at load time) ---- */
/* Returns sizeof(AcReal) and writes sizeof(AcMeshInfo) and sizeof(AcMesh). */
int acAbiRealSize(size_t* meshinfo_size, size_t* mesh_size);

/* ---- FP64 reduction results (same semantics as acDeviceReduce*, no rounding to AcReal) ---- */
AcResult acDeviceReduceScalDouble(const Device device, const Stream stream,
                                  const ReductionType rtype, const VertexBufferHandle vtxbuf,
                                  double* result);
AcResult acDeviceReduceVecDouble(const Device device, const Stream stream,
                                 const ReductionType rtype, const VertexBufferHandle vtxbuf0,
                                 const VertexBufferHandle vtxbuf1,
                                 const VertexBufferHandle vtxbuf2, double* result);
AcResult acDeviceReduceVecScalDouble(const Device device, const Stream stream,
                                     const ReductionType rtype, const VertexBufferHandle vtxbuf0,
                                     const VertexBufferHandle vtxbuf1,
                                     const VertexBufferHandle vtxbuf2,
                                     const VertexBufferHandle vtxbuf3, double* result);

/* ---- Whole RK3 step ----
 * Three times {periodic fill of `in` over the whole m^3 box, substep over the interior, swap},
 * the caller sequence of the reference's single-GPU node loop (src/core/node.cc:729-806).
 * The state ends in `in` without a final fill, as in the reference. */
AcResult acDeviceIntegrateStep(const Device device, const Stream stream, const AcReal dt);

/* Same step recorded once into a HIP graph per (stream, dt) and replayed afterwards. */
AcResult acDeviceIntegrateStepGraph(const Device device, const Stream stream, const AcReal dt);

/* ---- Solve-kernel configuration and autotuning ----
 * variant 0: one thread per cell, stencils read straight from global memory.
 * variant 1: x-y tile in LDS with a 3-cell halo, z-march of depth z_depth, register queues for
 *            the z stencils and accumulators for the x-z / y-z cross derivatives. */
typedef struct {
    int variant;
    int tile_x;
    int tile_y;
    int z_depth;
} AcKernelConfig;

/* Number of compiled configurations and the i-th of them (z_depth = 0 means "any"). */
int acKernelConfigCount(void);
AcResult acKernelConfigGet(const int index, AcKernelConfig* config);
AcResult acDeviceGetKernelConfig(const Device device, AcKernelConfig* config);
/* Rejects configurations that are not compiled in or do not fit the grid. */
AcResult acDeviceSetKernelConfig(const Device device, const AcKernelConfig config);
/* Times one configuration: mean milliseconds of one substep (step 2) over `iters` launches on
 * the device's current buffers. Returns a negative value for an unusable configuration. */
float acDeviceTimeKernelConfig(const Device device, const AcKernelConfig config, const int iters);

/* ---- Synthetic initial conditions ----
 * Fills the interior of `in[handle]` with lo + (hi - lo) u, u in [0, 1) from a counter-based
 * splitmix64 stream keyed by `seed` and the linear interior cell index (asynchronous on
 * `stream`). Ghost cells are untouched; apply acDevicePeriodicBoundconds afterwards. */
AcResult acDeviceRandomizeUniform(const Device device, const Stream stream,
                                  const VertexBufferHandle handle, const double lo,
                                  const double hi, const unsigned long long seed);

/* ---- Plain copy of in[handle] into out[handle] (all m^3 cells), asynchronous on `stream`.
 * One read and one write per cell: the known-size kernel used to calibrate memory counters. */
AcResult acDeviceCopyVertexBuffer(const Device device, const Stream stream,
                                  const VertexBufferHandle handle);

/* ---- Stream access for timing from the caller (returns the hipStream_t as void*) ---- */
void* acDeviceGetStreamHandle(const Device device, const Stream stream);

#ifdef __cplusplus
} /* extern "C" */
#endif
