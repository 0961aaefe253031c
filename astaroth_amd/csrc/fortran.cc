// Fortran binding (reference src/core/astaroth_fortran.cc:1-125): by-reference wrappers over the
// Device interface. Differences from the reference: the mesh calls size each field as
// mx * my * mz (the reference multiplies mx three times, astaroth_fortran.cc:52,71, which is
// wrong for any non-cubic box), and an AC_FAILURE from the wrapped call is fatal (the reference
// ignores the results of these calls except in its own checks).
#include "astaroth_fortran.h"

#include <cstdio>
#include <cstdlib>
#include <hip/hip_runtime_api.h>

namespace {

void
check(AcResult r, const char* fn)
{
    if (r != AC_SUCCESS) {
        fprintf(stderr, "%s: the call failed\n", fn);
        exit(EXIT_FAILURE);
    }
}

AcMesh
mesh_over(const AcMeshInfo* info, const int* num_farrays, AcReal* farray, const char* fn)
{
    if (*num_farrays < NUM_VTXBUF_HANDLES) {
        fprintf(stderr, "%s: %d field arrays given, %d needed\n", fn, *num_farrays,
                NUM_VTXBUF_HANDLES);
        exit(EXIT_FAILURE);
    }
    if (*num_farrays > NUM_VTXBUF_HANDLES)
        fprintf(stderr, "%s: %d field arrays given, the first %d are used\n", fn, *num_farrays,
                NUM_VTXBUF_HANDLES);
    const size_t mxyz = acVertexBufferSize(*info);
    AcMesh mesh;
    mesh.info = *info;
    for (int i = 0; i < NUM_VTXBUF_HANDLES; ++i)
        mesh.vertex_buffer[i] = farray + (size_t)i * mxyz;
    return mesh;
}

} // namespace

extern "C" {

void
achostupdatebuiltinparams_(AcMeshInfo* info)
{
    check(acHostUpdateBuiltinParams(info), "achostupdatebuiltinparams");
}

void
acgetdevicecount_(int* count)
{
    if (hipGetDeviceCount(count) != hipSuccess) {
        fprintf(stderr, "acgetdevicecount: hipGetDeviceCount failed\n");
        exit(EXIT_FAILURE);
    }
}

void
acdevicecreate_(const int* id, const AcMeshInfo* info, Device* handle)
{
    check(acDeviceCreate(*id, *info, handle), "acdevicecreate");
}

void
acdevicedestroy_(Device* device)
{
    check(acDeviceDestroy(*device), "acdevicedestroy");
    *device = nullptr;
}

void
acdeviceprintinfo_(const Device* device)
{
    check(acDevicePrintInfo(*device), "acdeviceprintinfo");
}

void
acdeviceloadmeshinfo_(const Device* device, const AcMeshInfo* info)
{
    check(acDeviceLoadMeshInfo(*device, *info), "acdeviceloadmeshinfo");
}

void
acdeviceloadmesh_(const Device* device, const Stream* stream, const AcMeshInfo* info,
                  const int* num_farrays, AcReal* farray)
{
    const AcMesh mesh = mesh_over(info, num_farrays, farray, "acdeviceloadmesh");
    check(acDeviceLoadMesh(*device, *stream, mesh), "acdeviceloadmesh");
}

void
acdevicestoremesh_(const Device* device, const Stream* stream, const AcMeshInfo* info,
                   const int* num_farrays, AcReal* farray)
{
    AcMesh mesh = mesh_over(info, num_farrays, farray, "acdevicestoremesh");
    check(acDeviceStoreMesh(*device, *stream, &mesh), "acdevicestoremesh");
}

void
acdeviceintegratesubstep_(const Device* device, const Stream* stream, const int* step_number,
                          const int3* start, const int3* end, const AcReal* dt)
{
    check(acDeviceIntegrateSubstep(*device, *stream, *step_number, *start, *end, *dt),
          "acdeviceintegratesubstep");
}

void
acdeviceperiodicboundconds_(const Device* device, const Stream* stream, const int3* start,
                            const int3* end)
{
    check(acDevicePeriodicBoundconds(*device, *stream, *start, *end),
          "acdeviceperiodicboundconds");
}

void
acdeviceswapbuffers_(const Device* device)
{
    check(acDeviceSwapBuffers(*device), "acdeviceswapbuffers");
}

void
acdevicereducescal_(const Device* device, const Stream* stream, const ReductionType* rtype,
                    const VertexBufferHandle* vtxbuf_handle, AcReal* result)
{
    check(acDeviceReduceScal(*device, *stream, *rtype, *vtxbuf_handle, result),
          "acdevicereducescal");
}

void
acdevicereducevec_(const Device* device, const Stream* stream, const ReductionType* rtype,
                   const VertexBufferHandle* vtxbuf0, const VertexBufferHandle* vtxbuf1,
                   const VertexBufferHandle* vtxbuf2, AcReal* result)
{
    check(acDeviceReduceVec(*device, *stream, *rtype, *vtxbuf0, *vtxbuf1, *vtxbuf2, result),
          "acdevicereducevec");
}

void
acdevicesynchronizestream_(const Device* device, const Stream* stream)
{
    check(acDeviceSynchronizeStream(*device, *stream), "acdevicesynchronizestream");
}

} // extern "C"
