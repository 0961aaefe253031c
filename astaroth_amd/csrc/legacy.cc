// The reference's legacy single-call interface (include/astaroth.h:208-283, implemented in
// src/core/astaroth.cc:28-188 over one Node) restated over one Device: a program written against
// acInit / acLoad / acIntegrate / acReduce* / acStore / acQuit runs unchanged on one GPU.
//
// The reference Node splits the box over the GPUs of a node and exchanges halos between them;
// on one GPU that exchange is the periodic fill, so:
//   acIntegrate    = 3 x {periodic fill, substep over the interior, swap} (node.cc:729-806),
//   acReduce*      = the Device reduction finalised as the Node does (node.cc:932-957):
//                    RMS, RMS_EXP and ALFVEN_RMS return sqrt(sum / n_x n_y n_z),
//   acSynchronizeMesh has nothing to exchange.
//   acIntegrateGBC = 3 x {general fill by the face types of `config`, substep, swap}
//                    (node.cc:808-880), acBoundcondStepGBC = the general fill alone
//                    (node.cc:923-930), both with the corrected fill of boundconds.hip.
// acGetNode returns a Node view of the Device (the Node interface of node.cc over it).
#include "astaroth.h"
#include "astaroth_ext.h"

#include <cmath>
#include <cstdio>
#include <hip/hip_runtime_api.h>

// node.cc: a Node that wraps an existing Device without owning it
Node ac_node_wrap(Device device, const AcMeshInfo& info);

namespace {

Device g_device        = nullptr;
Node g_node            = nullptr; // view of g_device
AcMeshInfo g_info      = {};

bool
ready(const char* fn)
{
    if (!g_device) {
        fprintf(stderr, "%s: acInit has not been called\n", fn);
        return false;
    }
    return true;
}

int3
i3(int x, int y, int z)
{
    int3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

AcReal
finalise(ReductionType rtype, AcReal sum)
{
    if (rtype == RTYPE_RMS || rtype == RTYPE_RMS_EXP || rtype == RTYPE_ALFVEN_RMS) {
        const AcReal inv_n = AcReal(1.) / (AcReal)((double)g_info.int_params[AC_nx] *
                                                    g_info.int_params[AC_ny] *
                                                    g_info.int_params[AC_nz]);
        return std::sqrt(inv_n * sum);
    }
    return sum;
}

} // namespace

extern "C" {

AcResult
acInit(const AcMeshInfo mesh_info)
{
    if (g_device) {
        acNodeDestroy(g_node); // the view only
        acDeviceDestroy(g_device);
        g_device = nullptr;
        g_node   = nullptr;
    }
    g_info = mesh_info;
    acHostUpdateBuiltinParams(&g_info);
    if (acDeviceCreate(0, g_info, &g_device) != AC_SUCCESS)
        return AC_FAILURE;
    g_node = ac_node_wrap(g_device, g_info);
    if (!g_node) {
        acDeviceDestroy(g_device);
        g_device = nullptr;
        return AC_FAILURE;
    }
    return AC_SUCCESS;
}

AcResult
acQuit(void)
{
    if (!ready("acQuit"))
        return AC_FAILURE;
    acNodeDestroy(g_node); // the view only
    const AcResult r = acDeviceDestroy(g_device);
    g_device         = nullptr;
    g_node           = nullptr;
    return r;
}

Node
acGetNode(void)
{
    if (!ready("acGetNode"))
        return nullptr;
    return g_node;
}

AcResult
acCheckDeviceAvailability(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess)
        return AC_FAILURE;
    return count > 0 ? AC_SUCCESS : AC_FAILURE;
}

int
acGetNumDevicesPerNode(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess)
        return 0;
    return count;
}

AcResult
acSynchronize(void)
{
    if (!ready("acSynchronize"))
        return AC_FAILURE;
    return acDeviceSynchronizeStream(g_device, STREAM_ALL);
}

AcResult
acSynchronizeStream(const Stream stream)
{
    if (!ready("acSynchronizeStream"))
        return AC_FAILURE;
    return acDeviceSynchronizeStream(g_device, stream);
}

AcResult
acSynchronizeMesh(void)
{
    if (!ready("acSynchronizeMesh"))
        return AC_FAILURE;
    return acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
}

AcResult
acLoadDeviceConstant(const AcRealParam param, const AcReal value)
{
    if (!ready("acLoadDeviceConstant"))
        return AC_FAILURE;
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    const AcResult r = acDeviceLoadScalarUniform(g_device, STREAM_DEFAULT, param, value);
    if (r == AC_SUCCESS && param >= 0 && param < NUM_REAL_PARAMS)
        g_info.real_params[param] = value;
    return r;
}

AcResult
acLoad(const AcMesh host_mesh)
{
    if (!ready("acLoad"))
        return AC_FAILURE;
    return acDeviceLoadMesh(g_device, STREAM_DEFAULT, host_mesh);
}

AcResult
acLoadWithOffset(const AcMesh host_mesh, const int3 src, const int num_vertices)
{
    if (!ready("acLoadWithOffset"))
        return AC_FAILURE;
    return acDeviceLoadMeshWithOffset(g_device, STREAM_DEFAULT, host_mesh, src, src, num_vertices);
}

AcResult
acSetVertexBuffer(const VertexBufferHandle handle, const AcReal value)
{
    if (!ready("acSetVertexBuffer"))
        return AC_FAILURE;
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    const AcResult r = acDeviceSetVertexBuffer(g_device, STREAM_DEFAULT, handle, value);
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    return r;
}

AcResult
acStore(AcMesh* host_mesh)
{
    if (!ready("acStore"))
        return AC_FAILURE;
    const AcResult r = acDeviceStoreMesh(g_device, STREAM_DEFAULT, host_mesh);
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    return r;
}

AcResult
acStoreWithOffset(const int3 dst, const size_t num_vertices, AcMesh* host_mesh)
{
    if (!ready("acStoreWithOffset"))
        return AC_FAILURE;
    const AcResult r = acDeviceStoreMeshWithOffset(g_device, STREAM_DEFAULT, dst, dst,
                                                   (int)num_vertices, host_mesh);
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    return r;
}

AcResult
acIntegrate(const AcReal dt)
{
    if (!ready("acIntegrate"))
        return AC_FAILURE;
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    const AcResult r = acDeviceIntegrateStep(g_device, STREAM_DEFAULT, dt);
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    return r;
}

AcResult
acIntegrateGBC(const AcMeshInfo config, const AcReal dt)
{
    if (!ready("acIntegrateGBC"))
        return AC_FAILURE;
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    const AcResult r = acDeviceIntegrateStepGBC(g_device, STREAM_DEFAULT, config, dt);
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    return r;
}

AcResult
acIntegrateStep(const int isubstep, const AcReal dt)
{
    if (!ready("acIntegrateStep"))
        return AC_FAILURE;
    const int3 start = i3(NGHOST, NGHOST, NGHOST);
    const int3 end   = i3(NGHOST + g_info.int_params[AC_nx], NGHOST + g_info.int_params[AC_ny],
                          NGHOST + g_info.int_params[AC_nz]);
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    return acDeviceIntegrateSubstep(g_device, STREAM_DEFAULT, isubstep, start, end, dt);
}

AcResult
acIntegrateStepWithOffset(const int isubstep, const AcReal dt, const int3 start, const int3 end)
{
    if (!ready("acIntegrateStepWithOffset"))
        return AC_FAILURE;
    acDeviceSynchronizeStream(g_device, STREAM_DEFAULT);
    return acDeviceIntegrateSubstep(g_device, STREAM_DEFAULT, isubstep, start, end, dt);
}

AcResult
acBoundcondStep(void)
{
    if (!ready("acBoundcondStep"))
        return AC_FAILURE;
    const int3 m = i3(g_info.int_params[AC_nx] + 2 * NGHOST, g_info.int_params[AC_ny] + 2 * NGHOST,
                      g_info.int_params[AC_nz] + 2 * NGHOST);
    return acDevicePeriodicBoundconds(g_device, STREAM_DEFAULT, i3(0, 0, 0), m);
}

AcResult
acBoundcondStepGBC(const AcMeshInfo config)
{
    if (!ready("acBoundcondStepGBC"))
        return AC_FAILURE;
    const int3 m = i3(g_info.int_params[AC_nx] + 2 * NGHOST, g_info.int_params[AC_ny] + 2 * NGHOST,
                      g_info.int_params[AC_nz] + 2 * NGHOST);
    return acDeviceGeneralBoundconds(g_device, STREAM_DEFAULT, i3(0, 0, 0), m, config,
                                     i3(-1, -1, -1));
}

AcReal
acReduceScal(const ReductionType rtype, const VertexBufferHandle vtxbuf_handle)
{
    if (!ready("acReduceScal"))
        return AcReal(NAN);
    AcReal r = AcReal(NAN);
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    if (acDeviceReduceScal(g_device, STREAM_DEFAULT, rtype, vtxbuf_handle, &r) != AC_SUCCESS)
        return AcReal(NAN);
    return finalise(rtype, r);
}

AcReal
acReduceVec(const ReductionType rtype, const VertexBufferHandle a, const VertexBufferHandle b,
            const VertexBufferHandle c)
{
    if (!ready("acReduceVec"))
        return AcReal(NAN);
    AcReal r = AcReal(NAN);
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    if (acDeviceReduceVec(g_device, STREAM_DEFAULT, rtype, a, b, c, &r) != AC_SUCCESS)
        return AcReal(NAN);
    return finalise(rtype, r);
}

AcReal
acReduceVecScal(const ReductionType rtype, const VertexBufferHandle a, const VertexBufferHandle b,
                const VertexBufferHandle c, const VertexBufferHandle d)
{
    if (!ready("acReduceVecScal"))
        return AcReal(NAN);
    AcReal r = AcReal(NAN);
    acDeviceSynchronizeStream(g_device, STREAM_ALL);
    if (acDeviceReduceVecScal(g_device, STREAM_DEFAULT, rtype, a, b, c, d, &r) != AC_SUCCESS)
        return AcReal(NAN);
    return finalise(rtype, r);
}

} // extern "C"
