/*
 * astaroth.h — C interface of the MI355X single-device stencil library.
 *
 * Drop-in for the Device-layer part of the reference header include/astaroth.h: the same
 * types (AcReal switch :30-42, AcMeshInfo :109-114, AcMesh :116-119, opaque Device :122),
 * the same enum values (tables in user_defines.h), the same index helpers (:153-183) and the
 * same Device-interface prototypes (:507-642) and host helpers (:650-659). Programs written
 * against the reference Device interface compile and link against this header and
 * libastaroth_fp32.so / libastaroth_fp64.so unchanged (AC_DOUBLE_PRECISION selects which).
 *
 * The reference's Node layer (:355-500), a wrapper that splits the box over the GPUs of a node,
 * and its legacy single-call interface (acInit ... acGetNode, :208-283) are provided over one
 * Device (one GPU). The Grid layer (MPI over several nodes) is not.
 */
#pragma once
#include <float.h>  /* FLT_EPSILON, etc */
#include <stdio.h>  /* printf */
#include <stdlib.h> /* size_t */

/* Vector types. HIP's int3/float3/double3 are trivially copyable 12/12/24-byte structs with
 * the same by-value calling convention as the plain structs below, so C callers and the
 * HIP-compiled library agree on the ABI. */
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_vector_types.h>
#else
typedef struct {
    int x, y, z;
} int3;
typedef struct {
    float x, y, z;
} float3;
typedef struct {
    double x, y, z;
} double3;
#endif

/* Library flags (reference include/astaroth.h:25-27) */
#define STENCIL_ORDER (6)
#define NGHOST (STENCIL_ORDER / 2)

#ifndef AC_DOUBLE_PRECISION
#define AC_DOUBLE_PRECISION (0)
#endif

#if AC_DOUBLE_PRECISION == 1
typedef double AcReal;
typedef double3 AcReal3;
#define AC_REAL_MAX (DBL_MAX)
#define AC_REAL_MIN (DBL_MIN)
#define AC_REAL_EPSILON (DBL_EPSILON)
#else
typedef float AcReal;
typedef float3 AcReal3;
#define AC_REAL_MAX (FLT_MAX)
#define AC_REAL_MIN (FLT_MIN)
#define AC_REAL_EPSILON (FLT_EPSILON)
#endif

typedef struct {
    AcReal3 row[3];
} AcMatrix;

#include "user_defines.h"

typedef enum { AC_SUCCESS = 0, AC_FAILURE = 1 } AcResult;

typedef enum {
    AC_BOUNDCOND_PERIODIC      = 0,
    AC_BOUNDCOND_SYMMETRIC     = 1,
    AC_BOUNDCOND_ANTISYMMETRIC = 2
} AcBoundcond;

#define AC_GEN_ID(X) X,
typedef enum { AC_FOR_RTYPES(AC_GEN_ID) NUM_RTYPES } ReductionType;
typedef enum { AC_FOR_USER_INT_PARAM_TYPES(AC_GEN_ID) NUM_INT_PARAMS } AcIntParam;
typedef enum { AC_FOR_USER_INT3_PARAM_TYPES(AC_GEN_ID) NUM_INT3_PARAMS } AcInt3Param;
typedef enum { AC_FOR_USER_REAL_PARAM_TYPES(AC_GEN_ID) NUM_REAL_PARAMS } AcRealParam;
typedef enum { AC_FOR_USER_REAL3_PARAM_TYPES(AC_GEN_ID) NUM_REAL3_PARAMS } AcReal3Param;
typedef enum { AC_FOR_SCALARARRAY_HANDLES(AC_GEN_ID) NUM_SCALARARRAY_HANDLES } ScalarArrayHandle;
typedef enum { AC_FOR_VTXBUF_HANDLES(AC_GEN_ID) NUM_VTXBUF_HANDLES } VertexBufferHandle;
#undef AC_GEN_ID

#define _UNUSED __attribute__((unused))
#define AC_GEN_STR(X) #X,
static const char* rtype_names[] _UNUSED       = {AC_FOR_RTYPES(AC_GEN_STR) "-end-"};
static const char* intparam_names[] _UNUSED    = {AC_FOR_USER_INT_PARAM_TYPES(AC_GEN_STR) "-end-"};
static const char* int3param_names[] _UNUSED   = {AC_FOR_USER_INT3_PARAM_TYPES(AC_GEN_STR) "-end-"};
static const char* realparam_names[] _UNUSED   = {AC_FOR_USER_REAL_PARAM_TYPES(AC_GEN_STR) "-end-"};
static const char* real3param_names[] _UNUSED  = {AC_FOR_USER_REAL3_PARAM_TYPES(AC_GEN_STR) "-end-"};
static const char* scalararray_names[] _UNUSED = {AC_FOR_SCALARARRAY_HANDLES(AC_GEN_STR) "-end-"};
static const char* vtxbuf_names[] _UNUSED      = {AC_FOR_VTXBUF_HANDLES(AC_GEN_STR) "-end-"};
#undef AC_GEN_STR
#undef _UNUSED

/* 424 bytes (FP32) / 728 bytes (FP64), as in the reference. The zero-length real3 array is
 * the reference's layout for a module without vector uniforms (GNU extension). */
typedef struct {
    int int_params[NUM_INT_PARAMS];
    int3 int3_params[NUM_INT3_PARAMS];
    AcReal real_params[NUM_REAL_PARAMS];
    AcReal3 real3_params[NUM_REAL3_PARAMS];
} AcMeshInfo;

typedef struct {
    AcReal* vertex_buffer[NUM_VTXBUF_HANDLES];
    AcMeshInfo info;
} AcMesh;

/* Opaque handle to one GPU's state (reference include/astaroth.h:122). */
typedef struct device_s* Device;

/* Opaque handle to a Node (reference :126): here one Device. */
typedef struct node_s* Node;

/* Grid dimensions with (m) and without (n) the ghost zones (reference :131-134). */
typedef struct {
    int3 m;
    int3 n;
} GridDims;

/* What acNodeQueryDeviceConfiguration reports (reference :136-142). */
typedef struct {
    int num_devices;
    Device* devices;
    GridDims grid;
    GridDims subgrid;
} DeviceConfiguration;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Helper functions (reference include/astaroth.h:153-183) ---- */
static inline size_t
acVertexBufferSize(const AcMeshInfo info)
{
    return (size_t)info.int_params[AC_mx] * (size_t)info.int_params[AC_my] *
           (size_t)info.int_params[AC_mz];
}

static inline size_t
acVertexBufferSizeBytes(const AcMeshInfo info)
{
    return sizeof(AcReal) * acVertexBufferSize(info);
}

static inline size_t
acVertexBufferCompdomainSize(const AcMeshInfo info)
{
    return (size_t)info.int_params[AC_nx] * (size_t)info.int_params[AC_ny] *
           (size_t)info.int_params[AC_nz];
}

static inline size_t
acVertexBufferCompdomainSizeBytes(const AcMeshInfo info)
{
    return sizeof(AcReal) * acVertexBufferCompdomainSize(info);
}

static inline size_t
acVertexBufferIdx(const int i, const int j, const int k, const AcMeshInfo info)
{
    return (size_t)i + (size_t)j * (size_t)info.int_params[AC_mx] +
           (size_t)k * (size_t)info.int_params[AC_mx] * (size_t)info.int_params[AC_my];
}

static inline void
acPrintMeshInfo(const AcMeshInfo config)
{
    for (int i = 0; i < NUM_INT_PARAMS; ++i)
        printf("[%s]: %d\n", intparam_names[i], config.int_params[i]);
    for (int i = 0; i < NUM_INT3_PARAMS; ++i)
        printf("[%s]: (%d, %d, %d)\n", int3param_names[i], config.int3_params[i].x,
               config.int3_params[i].y, config.int3_params[i].z);
    for (int i = 0; i < NUM_REAL_PARAMS; ++i)
        printf("[%s]: %g\n", realparam_names[i], (double)(config.real_params[i]));
}

/* ---- Device interface ---- */

/* ref include/astaroth.h:507. Selects GPU `id`, allocates the in/out 8-field buffer sets and
 * the reduction scratch, creates NUM_STREAMS streams (default priority), loads the DSL
 * defaults then `device_config`, and runs the autotuner. */
AcResult acDeviceCreate(const int id, const AcMeshInfo device_config, Device* device);

/* ref :510. Synchronises every stream and frees everything. */
AcResult acDeviceDestroy(Device device);

/* ref :513 */
AcResult acDevicePrintInfo(const Device device);

/* ref :516. Re-runs the tile/z-depth autotuner for this device's grid size. */
AcResult acDeviceAutoOptimize(const Device device);

/* ref :519. STREAM_ALL synchronises the whole device. */
AcResult acDeviceSynchronizeStream(const Device device, const Stream stream);

/* ref :522. Host-side swap of the in/out pointer sets. */
AcResult acDeviceSwapBuffers(const Device device);

/* ref :525-549. Uniform (parameter) loaders. Invalid indices and NaN/Inf values are rejected
 * with a warning on stderr and AC_FAILURE; the old value is kept. */
AcResult acDeviceLoadScalarUniform(const Device device, const Stream stream,
                                   const AcRealParam param, const AcReal value);
AcResult acDeviceLoadVectorUniform(const Device device, const Stream stream,
                                   const AcReal3Param param, const AcReal3 value);
AcResult acDeviceLoadIntUniform(const Device device, const Stream stream, const AcIntParam param,
                                const int value);
AcResult acDeviceLoadInt3Uniform(const Device device, const Stream stream, const AcInt3Param param,
                                 const int3 value);
AcResult acDeviceLoadScalarArray(const Device device, const Stream stream,
                                 const ScalarArrayHandle handle, const size_t start,
                                 const AcReal* data, const size_t num);
AcResult acDeviceLoadMeshInfo(const Device device, const AcMeshInfo device_config);
AcResult acDeviceLoadDefaultUniforms(const Device device);

/* ref :552-571. Host -> device copies into the `in` buffers, asynchronous on `stream`. */
AcResult acDeviceLoadVertexBufferWithOffset(const Device device, const Stream stream,
                                            const AcMesh host_mesh,
                                            const VertexBufferHandle vtxbuf_handle, const int3 src,
                                            const int3 dst, const int num_vertices);
AcResult acDeviceLoadMeshWithOffset(const Device device, const Stream stream,
                                    const AcMesh host_mesh, const int3 src, const int3 dst,
                                    const int num_vertices);
AcResult acDeviceLoadVertexBuffer(const Device device, const Stream stream, const AcMesh host_mesh,
                                  const VertexBufferHandle vtxbuf_handle);
AcResult acDeviceLoadMesh(const Device device, const Stream stream, const AcMesh host_mesh);
/* Sets both the `in` and the `out` buffer of `handle` to `value`. */
AcResult acDeviceSetVertexBuffer(const Device device, const Stream stream,
                                 const VertexBufferHandle handle, const AcReal value);

/* ref :574-588. Device -> host copies from the `in` buffers, asynchronous on `stream`; the
 * caller synchronises the stream before reading the host mesh. */
AcResult acDeviceStoreVertexBufferWithOffset(const Device device, const Stream stream,
                                             const VertexBufferHandle vtxbuf_handle, const int3 src,
                                             const int3 dst, const int num_vertices,
                                             AcMesh* host_mesh);
AcResult acDeviceStoreMeshWithOffset(const Device device, const Stream stream, const int3 src,
                                     const int3 dst, const int num_vertices, AcMesh* host_mesh);
AcResult acDeviceStoreVertexBuffer(const Device device, const Stream stream,
                                   const VertexBufferHandle vtxbuf_handle, AcMesh* host_mesh);
AcResult acDeviceStoreMesh(const Device device, const Stream stream, AcMesh* host_mesh);

/* ref :591-606. Device -> device copies of `in` buffers (also between two Devices on the same
 * GPU). */
AcResult acDeviceTransferVertexBufferWithOffset(const Device src_device, const Stream stream,
                                                const VertexBufferHandle vtxbuf_handle,
                                                const int3 src, const int3 dst,
                                                const int num_vertices, Device dst_device);
AcResult acDeviceTransferMeshWithOffset(const Device src_device, const Stream stream,
                                        const int3 src, const int3 dst, const int num_vertices,
                                        Device* dst_device);
AcResult acDeviceTransferVertexBuffer(const Device src_device, const Stream stream,
                                      const VertexBufferHandle vtxbuf_handle, Device dst_device);
AcResult acDeviceTransferMesh(const Device src_device, const Stream stream, Device dst_device);

/* ref :609-610. One substep of the 2N-RK3 integrator over cells [start, end) (local
 * m-coordinates, inside [NGHOST, n + NGHOST)): reads `in` (and `out` as the previous stage when
 * step_number > 0), writes `out`. No halo fill and no swap. step_number outside {0, 1, 2} is a
 * fatal error, as in the reference. */
AcResult acDeviceIntegrateSubstep(const Device device, const Stream stream, const int step_number,
                                  const int3 start, const int3 end, const AcReal dt);

/* ref :612-618. Periodic fill of the ghost cells of `in` that lie inside [start, end). */
AcResult acDevicePeriodicBoundcondStep(const Device device, const Stream stream,
                                       const VertexBufferHandle vtxbuf_handle, const int3 start,
                                       const int3 end);
AcResult acDevicePeriodicBoundconds(const Device device, const Stream stream, const int3 start,
                                    const int3 end);

/* ref :621-628. Ghost cells of `in` within [start, end) by the face types of `config`
 * (AC_bc_type_{bot,top}_{x,y,z}: AC_BOUNDCOND_PERIODIC wraps, _SYMMETRIC mirrors about the
 * first/last interior cell, _ANTISYMMETRIC mirrors with a sign change; edges and corners combine
 * their axes). bindex must be (-1, -1, -1), the single-device form; mirrored axes need n >= 4.
 * The reference's kernel (boundconds.cuh:3-128) writes 1.0 into every ghost cell and reads only
 * the x-top type; this is the fill its comments describe. */
AcResult acDeviceGeneralBoundcondStep(const Device device, const Stream stream,
                                      const VertexBufferHandle vtxbuf_handle, const int3 start,
                                      const int3 end, const AcMeshInfo config, const int3 bindex);
AcResult acDeviceGeneralBoundconds(const Device device, const Stream stream, const int3 start,
                                   const int3 end, const AcMeshInfo config, const int3 bindex);

/* ref :630-640. Synchronous reductions over the interior of `in`. MAX/MIN return the extreme
 * of the value (or vector length); RMS, RMS_EXP and ALFVEN_RMS return the raw sum (the caller
 * applies sqrt(sum / n^3)); vector SUM sums lengths. Accumulation is in FP64 and rounded to
 * AcReal here; acDeviceReduceScalDouble/VecDouble return the FP64 value. */
AcResult acDeviceReduceScal(const Device device, const Stream stream, const ReductionType rtype,
                            const VertexBufferHandle vtxbuf_handle, AcReal* result);
AcResult acDeviceReduceVec(const Device device, const Stream stream_type, const ReductionType rtype,
                           const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                           const VertexBufferHandle vtxbuf2, AcReal* result);
AcResult acDeviceReduceVecScal(const Device device, const Stream stream_type,
                               const ReductionType rtype, const VertexBufferHandle vtxbuf0,
                               const VertexBufferHandle vtxbuf1, const VertexBufferHandle vtxbuf2,
                               const VertexBufferHandle vtxbuf3, AcReal* result);

/* ref :642. There is no MPI in this library: returns AC_FAILURE. */
AcResult acDeviceRunMPITest(void);

/* ---- Host helpers (ref :650-659) ---- */
/* m = n + 6, n_min = 3, n_max = n + 3, inv_ds = 1/ds, mxy, nxy, nxyz. */
AcResult acHostUpdateBuiltinParams(AcMeshInfo* config);
AcResult acHostMeshCreate(const AcMeshInfo mesh_info, AcMesh* mesh);
/* Fills every cell of every field with U[0, 1) from a portable generator (the reference draws
 * from the C library's rand(), whose sequence is platform specific). */
AcResult acHostMeshRandomize(AcMesh* mesh);
AcResult acHostMeshDestroy(AcMesh* mesh);

/* ---- Legacy interface (ref :208-283, src/core/astaroth.cc:28-188) over one Device ----
 * acIntegrate is 3 x {periodic fill, substep, swap} and synchronises; acReduce* finalise the
 * RMS types as sqrt(sum / n_x n_y n_z), as the reference's Node does (node.cc:932-957);
 * acLoad is asynchronous on STREAM_DEFAULT, acStore synchronous. The *GBC variants use the
 * general boundary fill of acDeviceGeneralBoundconds with the face types of their `config`. */
AcResult acInit(const AcMeshInfo mesh_info);
AcResult acQuit(void);
AcResult acCheckDeviceAvailability(void);
AcResult acSynchronizeStream(const Stream stream);
AcResult acSynchronizeMesh(void);
AcResult acLoadDeviceConstant(const AcRealParam param, const AcReal value);
AcResult acLoad(const AcMesh host_mesh);
AcResult acSetVertexBuffer(const VertexBufferHandle handle, const AcReal value);
AcResult acStore(AcMesh* host_mesh);
AcResult acIntegrate(const AcReal dt);
AcResult acIntegrateGBC(const AcMeshInfo config, const AcReal dt);
AcResult acBoundcondStep(void);
AcResult acBoundcondStepGBC(const AcMeshInfo config);
AcReal acReduceScal(const ReductionType rtype, const VertexBufferHandle vtxbuf_handle);
AcReal acReduceVec(const ReductionType rtype, const VertexBufferHandle a,
                   const VertexBufferHandle b, const VertexBufferHandle c);
AcReal acReduceVecScal(const ReductionType rtype, const VertexBufferHandle a,
                       const VertexBufferHandle b, const VertexBufferHandle c,
                       const VertexBufferHandle d);
AcResult acStoreWithOffset(const int3 dst, const size_t num_vertices, AcMesh* host_mesh);
AcResult acIntegrateStep(const int isubstep, const AcReal dt);
AcResult acIntegrateStepWithOffset(const int isubstep, const AcReal dt, const int3 start,
                                   const int3 end);
AcResult acSynchronize(void);
AcResult acLoadWithOffset(const AcMesh host_mesh, const int3 src, const int num_vertices);
int acGetNumDevicesPerNode(void);
/* The Node view of the legacy interface's Device (valid between acInit and acQuit). */
Node acGetNode(void);

/* ---- Node interface (ref :355-500, src/core/node.cc) over one Device ----
 * acNodeCreate creates one Device on GPU 0 for the whole grid (its subgrid is the grid, so
 * acNodeQueryDeviceConfiguration reports num_devices = 1 and subgrid = grid) and every call is
 * the matching Device call, with these specifics:
 * - the *WithOffset load/store forms copy from `src` to `dst` as the Device forms do (the
 *   reference Node uses `src` for both, node.cc:452-455, 549-552);
 * - acNodeIntegrateSubstep clips the region to the interior and does nothing when it is empty;
 * - acNodeIntegrate / acNodeIntegrateGBC: the Device's RK3 step (acDeviceIntegrateStep /
 *   acDeviceIntegrateStepGBC on STREAM_0), synchronous. The reference splits each substep into
 *   an inner region integrated during the halo exchange between GPUs and six outer slabs; one
 *   GPU has no exchange to hide, and the split measured slower than the step;
 * - acNodeReduce* finalise RMS, RMS_EXP and ALFVEN_RMS as sqrt(sum / n_x n_y n_z);
 * - acNodePrintInfo and acNodeAutoOptimize are the Device's (the reference returns AC_FAILURE).
 * The load, store, set and reduction calls first synchronise their stream, as the reference
 * Node does. */
AcResult acNodeCreate(const int id, const AcMeshInfo node_config, Node* node);
AcResult acNodeDestroy(Node node);
AcResult acNodePrintInfo(const Node node);
AcResult acNodeQueryDeviceConfiguration(const Node node, DeviceConfiguration* config);
AcResult acNodeAutoOptimize(const Node node);
AcResult acNodeSynchronizeStream(const Node node, const Stream stream);
AcResult acNodeSynchronizeVertexBuffer(const Node node, const Stream stream,
                                       const VertexBufferHandle vtxbuf_handle);
AcResult acNodeSynchronizeMesh(const Node node, const Stream stream);
AcResult acNodeSwapBuffers(const Node node);
AcResult acNodeLoadConstant(const Node node, const Stream stream, const AcRealParam param,
                            const AcReal value);
AcResult acNodeLoadVertexBufferWithOffset(const Node node, const Stream stream,
                                          const AcMesh host_mesh,
                                          const VertexBufferHandle vtxbuf_handle, const int3 src,
                                          const int3 dst, const int num_vertices);
AcResult acNodeLoadMeshWithOffset(const Node node, const Stream stream, const AcMesh host_mesh,
                                  const int3 src, const int3 dst, const int num_vertices);
AcResult acNodeLoadVertexBuffer(const Node node, const Stream stream, const AcMesh host_mesh,
                                const VertexBufferHandle vtxbuf_handle);
AcResult acNodeLoadMesh(const Node node, const Stream stream, const AcMesh host_mesh);
AcResult acNodeSetVertexBuffer(const Node node, const Stream stream,
                               const VertexBufferHandle handle, const AcReal value);
AcResult acNodeStoreVertexBufferWithOffset(const Node node, const Stream stream,
                                           const VertexBufferHandle vtxbuf_handle, const int3 src,
                                           const int3 dst, const int num_vertices,
                                           AcMesh* host_mesh);
AcResult acNodeStoreMeshWithOffset(const Node node, const Stream stream, const int3 src,
                                   const int3 dst, const int num_vertices, AcMesh* host_mesh);
AcResult acNodeStoreVertexBuffer(const Node node, const Stream stream,
                                 const VertexBufferHandle vtxbuf_handle, AcMesh* host_mesh);
AcResult acNodeStoreMesh(const Node node, const Stream stream, AcMesh* host_mesh);
AcResult acNodeIntegrateSubstep(const Node node, const Stream stream, const int step_number,
                                const int3 start, const int3 end, const AcReal dt);
AcResult acNodeIntegrate(const Node node, const AcReal dt);
AcResult acNodeIntegrateGBC(const Node node, const AcMeshInfo config, const AcReal dt);
AcResult acNodePeriodicBoundcondStep(const Node node, const Stream stream,
                                     const VertexBufferHandle vtxbuf_handle);
AcResult acNodePeriodicBoundconds(const Node node, const Stream stream);
AcResult acNodeGeneralBoundcondStep(const Node node, const Stream stream,
                                    const VertexBufferHandle vtxbuf_handle,
                                    const AcMeshInfo config);
AcResult acNodeGeneralBoundconds(const Node node, const Stream stream, const AcMeshInfo config);
AcResult acNodeReduceScal(const Node node, const Stream stream, const ReductionType rtype,
                          const VertexBufferHandle vtxbuf_handle, AcReal* result);
AcResult acNodeReduceVec(const Node node, const Stream stream_type, const ReductionType rtype,
                         const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                         const VertexBufferHandle vtxbuf2, AcReal* result);
AcResult acNodeReduceVecScal(const Node node, const Stream stream_type, const ReductionType rtype,
                             const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                             const VertexBufferHandle vtxbuf2, const VertexBufferHandle vtxbuf3,
                             AcReal* result);

#ifdef __cplusplus
} /* extern "C" */
#endif

#include "astaroth_ext.h"
