/* Fortran binding of the Device interface (reference src/core/astaroth_fortran.h:1-52 and
 * src/core/astaroth_fortran.cc:1-125): one C symbol per call, lower case with a trailing
 * underscore, every argument by reference, so a Fortran program calls them with an implicit
 * interface (`call acdevicecreate(0, info, device)`) after
 *
 *     use, intrinsic :: iso_c_binding
 *     include "astaroth.f90"
 *
 * with include/fortran/fp32 or include/fortran/fp64 on the include path (the parameter tables,
 * the stream and reduction constants and the AcMeshInfo derived type of the matching build).
 * Errors that the C call reports are fatal here, as in the reference (the calls return nothing):
 * the message goes to stderr and the process exits with EXIT_FAILURE. */
#pragma once
#include "astaroth.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- utilities (astaroth_fortran.cc:7-20) ---- */
void achostupdatebuiltinparams_(AcMeshInfo* info);
void acgetdevicecount_(int* count);

/* ---- Device (astaroth_fortran.cc:22-125) ---- */
void acdevicecreate_(const int* id, const AcMeshInfo* info, Device* handle);
void acdevicedestroy_(Device* device);
void acdeviceprintinfo_(const Device* device);
void acdeviceloadmeshinfo_(const Device* device, const AcMeshInfo* info);
/* `farray` holds num_farrays >= NUM_VTXBUF_HANDLES fields of mx * my * mz values each, field
 * after field (a Fortran array f(mx, my, mz, nfields)); the first NUM_VTXBUF_HANDLES are moved. */
void acdeviceloadmesh_(const Device* device, const Stream* stream, const AcMeshInfo* info,
                       const int* num_farrays, AcReal* farray);
void acdevicestoremesh_(const Device* device, const Stream* stream, const AcMeshInfo* info,
                        const int* num_farrays, AcReal* farray);
void acdeviceintegratesubstep_(const Device* device, const Stream* stream, const int* step_number,
                               const int3* start, const int3* end, const AcReal* dt);
void acdeviceperiodicboundconds_(const Device* device, const Stream* stream, const int3* start,
                                 const int3* end);
void acdeviceswapbuffers_(const Device* device);
void acdevicereducescal_(const Device* device, const Stream* stream, const ReductionType* rtype,
                         const VertexBufferHandle* vtxbuf_handle, AcReal* result);
void acdevicereducevec_(const Device* device, const Stream* stream, const ReductionType* rtype,
                        const VertexBufferHandle* vtxbuf0, const VertexBufferHandle* vtxbuf1,
                        const VertexBufferHandle* vtxbuf2, AcReal* result);
void acdevicesynchronizestream_(const Device* device, const Stream* stream);

#ifdef __cplusplus
} /* extern "C" */
#endif
