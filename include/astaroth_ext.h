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

/* ---- Allocation-free reduction ----
 * No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The reduction of
 * acDeviceReduceScalDouble (num_fields = 1), acDeviceReduceVecDouble (3) or
 * acDeviceReduceVecScalDouble (4, the Alfven types) over fields[0..num_fields), with the same
 * FP64 result bit for bit (RMS types: the raw sum, as at Device level), written to d_result, a
 * device pointer the caller owns, on `stream` without waiting. Its per-block partials live in a
 * 32 KiB array of the library's code object, one per GPU (any grid size fits: the walk is
 * grid-stride and writes at most 4096 partials), instead of memory the Device allocated. That
 * array is shared by every call on the GPU and keeps the last call's partials: calls that overlap
 * on one GPU (several streams or Devices without ordering between them) race. Returns
 * AC_FAILURE for a null pointer, a field count other than 1, 3 or 4, a reduction type that does
 * not apply to that count, or an invalid stream or handle. */
AcResult acDeviceReduceNoAlloc(const Device device, const Stream stream, const ReductionType rtype,
                               const int num_fields, const VertexBufferHandle* fields,
                               double* d_result);

/* ---- Whole RK3 step ----
 * The interior result of three times {periodic fill of `in` over the whole m^3 box, substep over
 * the interior, swap}, the caller sequence of the reference's single-GPU node loop
 * (src/core/node.cc:729-806). With the default kernel configuration (FP32 variant 19, FP64
 * variant 8; 16 x 16 tiles), compensated integration off and n >= 3 on every axis, the substeps
 * can instead read the periodic images of the ghost cells directly and make no fill: always in
 * FP64, in FP32 below 384^3 cells (above, its image-reading build is slower than the fills it
 * saves); acDeviceAutoOptimize measures both forms and keeps the faster, and
 * acDeviceSetReadPeriodicImages overrides the choice. The state ends in `in` without a final
 * fill, as in the reference: its ghost cells are to be filled before they are read. */
AcResult acDeviceIntegrateStep(const Device device, const Stream stream, const AcReal dt);

/* One substep of acDeviceIntegrateStep: over the whole interior of a periodic box, reading the
 * periodic images of the ghost cells where the step does (acDeviceSubstepReadsPeriodicImages
 * returns 1), otherwise after a periodic fill of `in`. No swap. */
AcResult acDeviceIntegrateSubstepPeriodic(const Device device, const Stream stream,
                                          const int step_number, const AcReal dt);
int acDeviceSubstepReadsPeriodicImages(const Device device);
/* Whether periodic steps read the ghost cells' images (1) or fill them (0); AC_FAILURE when 1 is
 * asked of a configuration without an image-reading build. */
AcResult acDeviceSetReadPeriodicImages(const Device device, const int enable);

/* The step with general boundaries (acNodeIntegrateGBC, node.cc:808-880, on one Device):
 * 3 x {acDeviceGeneralBoundconds over the whole box with the face types of `config`, substep,
 * swap}; acDeviceIntegrateStep when every face is periodic. */
AcResult acDeviceIntegrateStepGBC(const Device device, const Stream stream,
                                  const AcMeshInfo config, const AcReal dt);

/* Same step recorded once into a HIP graph per (stream, dt) and replayed afterwards. */
AcResult acDeviceIntegrateStepGraph(const Device device, const Stream stream, const AcReal dt);

/* ---- Solve-kernel configuration and autotuning ----
 * variant 0: one thread per cell, stencils read straight from global memory.
 * variant 1: x-y tile in LDS with a 3-cell halo, z-march of depth z_depth, register queues for
 *            the z stencils and accumulators for the x-z / y-z cross derivatives.
 * variant 2: variant 1 with the plane loop unrolled by the 7-plane queue length.
 * variants 3, 4: the lean z-march (all eight fields in one LDS ring, 3-register z queues, six
 *            combined cross-derivative sums), compiled for one and two waves per SIMD.
 * variants 5, 6: the lean z-march with a cell-major LDS ring (FP32).
 * variant 7: the lean z-march with a 5-slot ring, one barrier per plane.
 * variants 8, 9: the lean z-march loading the previous stage first and storing each output
 *            plane one step later; 4 and 5 ring slots (variant 8 is the FP64 default).
 * variants 10-12: two-waves-per-SIMD builds of the lean z-march on 16 x 16 tiles (10: deferred
 *            stores, 11: previous stage first, 12: both).
 * variant 13: variant 12 with an 8-plane loop, whose LDS ring slots are compile-time (FP32).
 * variant 16: variant 12 with the 5-slot ring (one barrier per plane).
 * variants 17, 19, 22: variants 16, 12 and 10 reading each field's LDS stencil window through
 *            one pinned base register (FP32; variant 19 is the FP32 default).
 * variant 23: variant 17 with each plane loaded straight into its ring slot by direct-to-LDS
 *            loads (no staging registers or LDS writes; FP32). */
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

/* ---- CFL time step ----
 * dt of the reference's standalone driver (samples/standalone/model/host_timestep.cc:34-56):
 * min(cdt dsmin / (|u|_max + cs), cdtv dsmin^2 / max(nu, eta, gamma)), with |u|_max from an FP64
 * device reduction (synchronous) and dsmin = AC_dsmin or, when unset, min(dsx, dsy, dsz). */
AcResult acDeviceTimestepCFL(const Device device, const Stream stream, double* dt);

/* ---- Compensated integration (the precision study's FP32 mode) ----
 * With enable != 0 the Device keeps, for every field, the rounding error of its stored state in
 * a second buffer (two more buffer sets, allocated and zeroed here) and each substep adds its
 * increment with an error-free two-sum: u_out + c_out = u_in + d exactly, where d includes
 * c_in. The right-hand side still reads the rounded state only. Loads, sets and randomisations
 * of a field reset its compensation to zero. The substep then always uses the lean z-march in
 * its default tile. enable == 0 frees the buffers. */
AcResult acDeviceSetCompensatedIntegration(const Device device, const int enable);

/* ---- Synthetic initial conditions ----
 * Fills the interior of `in[handle]` with lo + (hi - lo) u, u in [0, 1) from a counter-based
 * splitmix64 stream keyed by `seed` and the linear interior cell index (asynchronous on
 * `stream`). Ghost cells are untouched; apply acDevicePeriodicBoundconds afterwards. */
AcResult acDeviceRandomizeUniform(const Device device, const Stream stream,
                                  const VertexBufferHandle handle, const double lo,
                                  const double hi, const unsigned long long seed);

/* Same stream plus amplitude * sin(phase_per_cell * (k - 3)), evaluated in FP64 before the
 * single rounding to AcReal. With phase_per_cell = 2 dz this is the 1e-2 sin(2 z) perturbation
 * of u_x in the 512^3 precision study; dz is passed in FP64 because AC_dsz may be FP32. */
AcResult acDeviceRandomizeUniformSinZ(const Device device, const Stream stream,
                                      const VertexBufferHandle handle, const double lo,
                                      const double hi, const unsigned long long seed,
                                      const double amplitude, const double phase_per_cell);

/* ---- Plain copy of in[handle] into out[handle] (all m^3 cells), asynchronous on `stream`.
 * One read and one write per cell: the known-size kernel used to calibrate memory counters. */
AcResult acDeviceCopyVertexBuffer(const Device device, const Stream stream,
                                  const VertexBufferHandle handle);

/* ---- Halo pack / unpack ----
 * Device-level forms of the reference's acKernelPackData / acKernelUnpackData
 * (src/core/kernels/kernels.h:65-70, packing.cuh:64-92): copy the block of `dims` cells at
 * `vba_start` (local m-coordinates) of the eight `in` fields to / from `packed`, a device buffer
 * of 8 * dims.x * dims.y * dims.z AcReal laid out [field][z][y][x]. Asynchronous on `stream`.
 * A block that does not lie inside the m^3 box is rejected with AC_FAILURE. */
AcResult acDevicePackData(const Device device, const Stream stream, const int3 vba_start,
                          const int3 dims, AcReal* packed);
AcResult acDeviceUnpackData(const Device device, const Stream stream, const AcReal* packed,
                            const int3 vba_start, const int3 dims);

/* The 26 halo blocks of the reference's Grid layer (src/core/device.cc:1375-1381 extents,
 * 1758-1818 ghost-side origins; corners included) packed in one launch: 8 corners, 4 x-edges,
 * 4 y-edges, 4 z-edges, 2 xy-sides, 2 xz-sides, 2 yz-sides, each block [field][z][y][x], blocks
 * back to back. acDevicePackHalo reads each block's periodic interior image
 * (a0 = ((b0 - 3) mod n) + 3, device.cc:1102-1116); acDeviceUnpackHalo writes the ghost blocks
 * b0. acDeviceHaloPackedSize is the buffer length in AcReal. Boxes with fewer than 3 interior
 * cells on an axis have no such blocks: size 0 and AC_FAILURE. */
size_t acDeviceHaloPackedSize(const Device device);
AcResult acDevicePackHalo(const Device device, const Stream stream, AcReal* packed);
AcResult acDeviceUnpackHalo(const Device device, const Stream stream, const AcReal* packed);
/* Pack followed by unpack through a device-owned buffer (allocated on the first call): the
 * halo exchange of a box that is its own periodic neighbour on every side. Fills every ghost
 * cell exactly as acDevicePeriodicBoundconds over the whole m^3 box does. */
AcResult acDeviceHaloExchangeSelf(const Device device, const Stream stream);

/* ---- Batched ensembles of independent periodic boxes ----
 * No counterpart in the reference (one box per Device there); for callers that advance many
 * small boxes of uneven sizes, e.g. parameter scans. Each box has its own AcMeshInfo (grid size,
 * spacing, viscosity, ...) and the reference layout. One RK3 step is 3 persistent solve launches
 * over a (box, tile column, z range) work list, which read the periodic images of the ghost
 * cells directly (the same values a periodic fill would give), then one periodic-fill launch over
 * all boxes (so the ghosts are current after the step) and per-box FP64 diagnostics, on the
 * ensemble's own stream. All boxes share dt.
 * Box fields are set with acEnsembleRandomizeUniform / acEnsembleSetVertexBuffer (device) or
 * acEnsembleLoadMesh (host, synchronous); acEnsemblePeriodicBoundconds fills the ghosts. */
typedef struct ensemble_s* Ensemble;
AcResult acEnsembleCreate(const int device_id, const int nboxes, const AcMeshInfo* configs,
                          Ensemble* ensemble);
AcResult acEnsembleDestroy(Ensemble ensemble);
int acEnsembleBoxCount(const Ensemble ensemble);
/* Batched solve configuration: variant 3, tile 32x4, 16x8 or 16x16, z_depth (0: whole box) for
 * every box (it replaces the tuner's per-class z-depths). */
AcResult acEnsembleSetKernelConfig(Ensemble ensemble, const AcKernelConfig config);
AcResult acEnsembleGetKernelConfig(const Ensemble ensemble, AcKernelConfig* config);
/* The ensemble's tuner, run by acEnsembleCreate unless the environment sets AC_AUTOTUNE=0 (then
 * 16 x 16 tiles and whole columns): times the batched launch over the whole work list for each
 * tile shape, then, per size class (boxes of one column length n_z, the class with the most cells
 * first), whole columns against columns cut into 2 or 4 z ranges, keeping a cut only when the
 * launch gets 2 % faster. Results are kept per composition (device and box sizes) for the
 * process. acEnsembleBoxZDepth gives a box's output planes per work item (0: whole columns). */
AcResult acEnsembleAutoOptimize(Ensemble ensemble);
int acEnsembleBoxZDepth(const Ensemble ensemble, const int box);
AcResult acEnsembleRandomizeUniform(Ensemble ensemble, const int box,
                                    const VertexBufferHandle handle, const double lo,
                                    const double hi, const unsigned long long seed);
AcResult acEnsembleSetVertexBuffer(Ensemble ensemble, const int box,
                                   const VertexBufferHandle handle, const AcReal value);
AcResult acEnsembleLoadMesh(Ensemble ensemble, const int box, const AcMesh host);
AcResult acEnsembleStoreMesh(Ensemble ensemble, const int box, AcMesh* host);
AcResult acEnsemblePeriodicBoundconds(Ensemble ensemble);
/* One RK3 step of every box: eager launches, or one HIP graph per (dt, buffer parity) replayed. */
AcResult acEnsembleIntegrateStep(Ensemble ensemble, const AcReal dt);
AcResult acEnsembleIntegrateStepGraph(Ensemble ensemble, const AcReal dt);
/* The same step as a loop of single-box launches, each box stepped as a single-box Device steps
 * it (its default kernel configuration; a fill + solve per substep, or the solve reading the ghost
 * cells' periodic images where that is the default form): the baseline the batched step is
 * measured against. */
AcResult acEnsembleIntegrateStepLooped(Ensemble ensemble, const AcReal dt);
/* Diagnostics of the last step, 7 per box: max |u|, rms |u|, max |A|, rms |A|, max lnrho,
 * min lnrho, rms lnrho (FP64, over the interior; synchronous). */
AcResult acEnsembleDiagnostics(Ensemble ensemble, double* results);
/* The CFL time step of every box from its last-step max |u| (see acDeviceTimestepCFL). */
AcResult acEnsembleTimestepCFL(Ensemble ensemble, double* dts);
AcResult acEnsembleSynchronize(const Ensemble ensemble);
void* acEnsembleGetStreamHandle(const Ensemble ensemble);

/* ---- Stream access for timing from the caller (returns the hipStream_t as void*) ---- */
void* acDeviceGetStreamHandle(const Device device, const Stream stream);

/* Diagnostics of the `in` state in one pass over the interior (synchronous), FP64:
 * out[0..2] = max, min, rms of |u|; out[3..5] = the same of |A|; out[6..8] = max, min, rms of
 * lnrho (rms = sqrt(sum of squares / (nx ny nz))). The maxima and minima equal those of
 * acDeviceReduceVecDouble / acDeviceReduceScalDouble bit for bit; the sums differ from theirs
 * only by summation order. One read of 7 fields instead of 21 for the nine single reductions. */
AcResult acDeviceDiagnostics(const Device device, const Stream stream, double* out);

/* The parameters in effect on the Device (its creation configuration with every uniform loaded
 * since). */
AcResult acDeviceGetMeshInfo(const Device device, AcMeshInfo* info);

/* ---- Snapshots, diagnostics time series and the time-stepping driver ----
 * Files of the reference's standalone simulation (samples/standalone/simulation.cc), in `dir`
 * (NULL or "" for the working directory):
 *   mesh_info.list        every parameter by name (simulation.cc:62-110);
 *   <FIELD>_<step>.mesh   the time stamp then the whole m^3 buffer, ghosts included, AcReal,
 *                         x fastest, host byte order (simulation.cc:115-147, read 205-245);
 *   timeseries.ts         title row and one row per diagnostics step (simulation.cc:251-313,
 *                         345-362).
 * The reference's reader (analysis/python/astar/data/read.py) and astaroth_amd/snapshot.py read
 * them. I/O errors, missing files and files of the wrong length return AC_FAILURE. */
AcResult acHostWriteMeshInfo(const AcMeshInfo info, const char* dir);
AcResult acHostSaveMeshSnapshot(const AcMesh mesh, const int step, const AcReal t, const char* dir);
AcResult acHostReadMeshSnapshot(AcMesh mesh, const int step, AcReal* t, const char* dir);
/* Appends the title row when start_step == 0, then a line break (as the reference does on every
 * start, also a resumed one). */
AcResult acTimeseriesBegin(const char* path, const int start_step);
/* Appends the row of `step`: step, t, dt, then min, rms, max of |u| and of every field over the
 * interior (FP64 device reductions, rms = sqrt(sum / (nx ny nz)); synchronous). Sets *found_nan
 * to 1 when a field's statistic is NaN. */
AcResult acDeviceTimeseriesAppend(const Device device, const Stream stream, const int step,
                                  const AcReal dt, const AcReal t, const char* path,
                                  int* found_nan);
/* Boundary fill of the `in` buffers by the Device's face types (acDeviceGeneralBoundconds),
 * store to the host and write the snapshot of `step`. */
AcResult acDeviceSaveMeshSnapshot(const Device device, const Stream stream, const int step,
                                  const AcReal t, const char* dir);
/* Read the snapshot of `step` into the `in` buffers (synchronous); *t gets its time stamp. */
AcResult acDeviceLoadMeshSnapshot(const Device device, const Stream stream, const int step,
                                  AcReal* t, const char* dir);

typedef enum {
    AC_STOP_MAX_STEPS = 0, /* reached AC_max_steps */
    AC_STOP_MAX_TIME  = 1, /* t >= AC_max_time (> 0) */
    AC_STOP_SMALL_DT  = 2, /* dt below 1e-5 of the early-run dt more than 10 steps in a row */
    AC_STOP_NAN       = 3, /* a diagnostics row held NaN */
    AC_STOP_FILE      = 4  /* a file named STOP exists in `dir` */
} AcStopReason;

typedef struct {
    int last_step; /* last step taken */
    AcReal t;      /* simulated time after it */
    int stop;      /* AcStopReason */
} AcSimulationResult;

/* The reference driver loop (simulation.cc:316-560) on the Device's loaded state, with the
 * Device's boundary face types (acDeviceIntegrateStepGBC; periodic by default): with
 * AC_start_step > 0 it first reads that step's snapshot; it writes
 * mesh_info.list, appends to timeseries.ts (row 0 and snapshot 0 on a fresh start), then for
 * steps start+1 .. AC_max_steps-1: dt = acDeviceTimestepCFL, acDeviceIntegrateStep, t += dt;
 * a diagnostics row every AC_save_steps; a snapshot every AC_bin_steps or each time t passes the
 * next multiple of AC_bin_save_t; it stops at AC_max_time, or with a final snapshot on a too
 * small dt, a NaN or a STOP file. */
AcResult acDeviceRunSimulation(const Device device, const char* dir, AcSimulationResult* result);

#ifdef __cplusplus
} /* extern "C" */
#endif
