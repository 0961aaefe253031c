// The reference's Node interface (include/astaroth.h:355-500, src/core/node.cc) on one GPU.
//
// The reference Node cuts the box along z into one subgrid per GPU of the node and exchanges the
// z halos between them. With one GPU the subgrid is the whole grid and there is nothing to
// exchange, so every call is the matching Device call:
//   acNodeLoad* / acNodeStore*: the Device copies. The offset forms copy `num_vertices` cells
//     from `src` to `dst`, as the Device forms do; the reference Node passes `src` for both
//     (node.cc:452-455, 549-552, marked there as still to fix), which only differs when
//     src != dst.
//   acNodeIntegrateSubstep: the region clipped to the interior, as node.cc:595-612 clips it to
//     each subgrid, and skipped when empty.
//   acNodeIntegrate / acNodeIntegrateGBC: the Device's RK3 step (acDeviceIntegrateStep /
//     acDeviceIntegrateStepGBC on STREAM_0), between synchronisations of every stream. The
//     reference splits each substep into an inner region, integrated while the halos are
//     exchanged between GPUs, and six outer slabs (node.cc:729-880). On one GPU there is no
//     exchange: the halo is the box's own periodic fill (or images read in its place), 4 % of a
//     512^3 FP32 step, and every split measured slower than the step it would overlap (512^3
//     FP32, same box, against 12.8 ms for the Device step): the reference's six 3-cell slabs
//     after an inner region [2N, n) on another stream 17.0 ms (a 3-cell slab leaves 13 of a
//     16-wide tile's lanes idle); a one-tile ring in x and y with the inner region over all of
//     z (reading its z ghosts' images) 15.9 ms with the ring in 28-plane z ranges, 28.9 ms in
//     whole columns queued piece after piece, 19.3 ms in whole columns in one launch (the GPU
//     dispatches the inner launch's remaining blocks before the ring's, so the ring runs as a
//     third round of columns). The result is the Device step's, bit for bit.
//   acNodeReduce*: the Device reduction, finalised as node.cc:932-957 finalises the per-GPU
//     results: RMS, RMS_EXP and ALFVEN_RMS return sqrt(sum / (n_x n_y n_z)).
//   acNodePrintInfo, acNodeAutoOptimize: the Device's (the reference returns AC_FAILURE for both,
//     node.cc:347-371).
#include "astaroth.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <new>

struct node_s {
    int id;
    AcMeshInfo config;
    Device device;
    bool owns_device; // false for the Node of the legacy interface, which wraps its Device
    GridDims grid;
};

namespace {

int3
i3(int x, int y, int z)
{
    int3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

GridDims
dims_of(const AcMeshInfo& info)
{
    GridDims g;
    g.n = i3(info.int_params[AC_nx], info.int_params[AC_ny], info.int_params[AC_nz]);
    g.m = i3(g.n.x + 2 * NGHOST, g.n.y + 2 * NGHOST, g.n.z + 2 * NGHOST);
    return g;
}

bool
valid(const Node node, const char* fn)
{
    if (!node || !node->device) {
        fprintf(stderr, "%s: invalid Node\n", fn);
        return false;
    }
    return true;
}

AcReal
finalise(const Node node, ReductionType rtype, AcReal sum)
{
    if (rtype == RTYPE_RMS || rtype == RTYPE_RMS_EXP || rtype == RTYPE_ALFVEN_RMS) {
        const GridDims& g  = node->grid;
        const AcReal inv_n = AcReal(1.) / (AcReal)((double)g.n.x * g.n.y * g.n.z);
        return std::sqrt(inv_n * sum);
    }
    return sum;
}

} // namespace

// The Node of the legacy interface: a view of its Device (destroying it leaves the Device).
Node
ac_node_wrap(Device device, const AcMeshInfo& info)
{
    node_s* node = new (std::nothrow) node_s;
    if (!node)
        return nullptr;
    node->id          = 0;
    node->config      = info;
    node->device      = device;
    node->owns_device = false;
    node->grid        = dims_of(info);
    return node;
}

extern "C" {

AcResult
acNodeCreate(const int id, const AcMeshInfo node_config, Node* node_handle)
{
    if (!node_handle)
        return AC_FAILURE;
    *node_handle = nullptr;
    node_s* node = new (std::nothrow) node_s;
    if (!node)
        return AC_FAILURE;
    node->id          = id;
    // the built-in sizes follow from the grid, as the reference derives its subgrid's
    // (node.cc:240-242)
    node->config = node_config;
    acHostUpdateBuiltinParams(&node->config);
    node->owns_device = true;
    node->grid        = dims_of(node_config);
    if (acDeviceCreate(0, node->config, &node->device) != AC_SUCCESS) {
        delete node;
        return AC_FAILURE;
    }
    *node_handle = node;
    return AC_SUCCESS;
}

AcResult
acNodeDestroy(Node node)
{
    if (!node)
        return AC_FAILURE;
    AcResult r = AC_SUCCESS;
    if (node->owns_device && node->device)
        r = acDeviceDestroy(node->device);
    delete node;
    return r;
}

AcResult
acNodePrintInfo(const Node node)
{
    if (!valid(node, "acNodePrintInfo"))
        return AC_FAILURE;
    printf("Node %d: 1 device, grid n = (%d, %d, %d), m = (%d, %d, %d)\n", node->id,
           node->grid.n.x, node->grid.n.y, node->grid.n.z, node->grid.m.x, node->grid.m.y,
           node->grid.m.z);
    return acDevicePrintInfo(node->device);
}

AcResult
acNodeQueryDeviceConfiguration(const Node node, DeviceConfiguration* config)
{
    if (!valid(node, "acNodeQueryDeviceConfiguration") || !config)
        return AC_FAILURE;
    config->num_devices = 1;
    config->devices     = &node->device;
    config->grid        = node->grid;
    config->subgrid     = node->grid;
    return AC_SUCCESS;
}

AcResult
acNodeAutoOptimize(const Node node)
{
    if (!valid(node, "acNodeAutoOptimize"))
        return AC_FAILURE;
    return acDeviceAutoOptimize(node->device);
}

AcResult
acNodeSynchronizeStream(const Node node, const Stream stream)
{
    if (!valid(node, "acNodeSynchronizeStream"))
        return AC_FAILURE;
    return acDeviceSynchronizeStream(node->device, stream);
}

AcResult
acNodeSynchronizeVertexBuffer(const Node node, const Stream stream,
                              const VertexBufferHandle vtxbuf_handle)
{
    (void)vtxbuf_handle; // one subgrid: no halo to exchange
    return acNodeSynchronizeStream(node, stream);
}

AcResult
acNodeSynchronizeMesh(const Node node, const Stream stream)
{
    return acNodeSynchronizeStream(node, stream);
}

AcResult
acNodeSwapBuffers(const Node node)
{
    if (!valid(node, "acNodeSwapBuffers"))
        return AC_FAILURE;
    return acDeviceSwapBuffers(node->device);
}

AcResult
acNodeLoadConstant(const Node node, const Stream stream, const AcRealParam param,
                   const AcReal value)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceLoadScalarUniform(node->device, stream, param, value);
}

AcResult
acNodeLoadVertexBufferWithOffset(const Node node, const Stream stream, const AcMesh host_mesh,
                                 const VertexBufferHandle vtxbuf_handle, const int3 src,
                                 const int3 dst, const int num_vertices)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceLoadVertexBufferWithOffset(node->device, stream, host_mesh, vtxbuf_handle, src,
                                              dst, num_vertices);
}

AcResult
acNodeLoadMeshWithOffset(const Node node, const Stream stream, const AcMesh host_mesh,
                         const int3 src, const int3 dst, const int num_vertices)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceLoadMeshWithOffset(node->device, stream, host_mesh, src, dst, num_vertices);
}

AcResult
acNodeLoadVertexBuffer(const Node node, const Stream stream, const AcMesh host_mesh,
                       const VertexBufferHandle vtxbuf_handle)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceLoadVertexBuffer(node->device, stream, host_mesh, vtxbuf_handle);
}

AcResult
acNodeLoadMesh(const Node node, const Stream stream, const AcMesh host_mesh)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceLoadMesh(node->device, stream, host_mesh);
}

AcResult
acNodeSetVertexBuffer(const Node node, const Stream stream, const VertexBufferHandle handle,
                      const AcReal value)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS ||
        acDeviceSetVertexBuffer(node->device, stream, handle, value) != AC_SUCCESS)
        return AC_FAILURE;
    return acNodeSynchronizeStream(node, stream);
}

AcResult
acNodeStoreVertexBufferWithOffset(const Node node, const Stream stream,
                                  const VertexBufferHandle vtxbuf_handle, const int3 src,
                                  const int3 dst, const int num_vertices, AcMesh* host_mesh)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceStoreVertexBufferWithOffset(node->device, stream, vtxbuf_handle, src, dst,
                                               num_vertices, host_mesh);
}

AcResult
acNodeStoreMeshWithOffset(const Node node, const Stream stream, const int3 src, const int3 dst,
                          const int num_vertices, AcMesh* host_mesh)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceStoreMeshWithOffset(node->device, stream, src, dst, num_vertices, host_mesh);
}

AcResult
acNodeStoreVertexBuffer(const Node node, const Stream stream,
                        const VertexBufferHandle vtxbuf_handle, AcMesh* host_mesh)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceStoreVertexBuffer(node->device, stream, vtxbuf_handle, host_mesh);
}

AcResult
acNodeStoreMesh(const Node node, const Stream stream, AcMesh* host_mesh)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceStoreMesh(node->device, stream, host_mesh);
}

AcResult
acNodeIntegrateSubstep(const Node node, const Stream stream, const int step_number,
                       const int3 start, const int3 end, const AcReal dt)
{
    if (acNodeSynchronizeStream(node, stream) != AC_SUCCESS)
        return AC_FAILURE;
    const int3 n  = node->grid.n;
    const int3 lo = i3(start.x > NGHOST ? start.x : NGHOST, start.y > NGHOST ? start.y : NGHOST,
                       start.z > NGHOST ? start.z : NGHOST);
    const int3 hi = i3(end.x < n.x + NGHOST ? end.x : n.x + NGHOST,
                       end.y < n.y + NGHOST ? end.y : n.y + NGHOST,
                       end.z < n.z + NGHOST ? end.z : n.z + NGHOST);
    if (hi.x <= lo.x || hi.y <= lo.y || hi.z <= lo.z)
        return AC_SUCCESS; // nothing of the region lies in the interior
    return acDeviceIntegrateSubstep(node->device, stream, step_number, lo, hi, dt);
}

AcResult
acNodeIntegrate(const Node node, const AcReal dt)
{
    if (!valid(node, "acNodeIntegrate"))
        return AC_FAILURE;
    const Device d = node->device;
    if (acDeviceSynchronizeStream(d, STREAM_ALL) != AC_SUCCESS ||
        acDeviceIntegrateStep(d, STREAM_0, dt) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceSynchronizeStream(d, STREAM_ALL);
}

AcResult
acNodeIntegrateGBC(const Node node, const AcMeshInfo config, const AcReal dt)
{
    if (!valid(node, "acNodeIntegrateGBC"))
        return AC_FAILURE;
    const Device d = node->device;
    if (acDeviceSynchronizeStream(d, STREAM_ALL) != AC_SUCCESS ||
        acDeviceIntegrateStepGBC(d, STREAM_0, config, dt) != AC_SUCCESS)
        return AC_FAILURE;
    return acDeviceSynchronizeStream(d, STREAM_ALL);
}

AcResult
acNodePeriodicBoundcondStep(const Node node, const Stream stream,
                            const VertexBufferHandle vtxbuf_handle)
{
    if (!valid(node, "acNodePeriodicBoundcondStep"))
        return AC_FAILURE;
    return acDevicePeriodicBoundcondStep(node->device, stream, vtxbuf_handle, i3(0, 0, 0),
                                         node->grid.m);
}

AcResult
acNodePeriodicBoundconds(const Node node, const Stream stream)
{
    if (!valid(node, "acNodePeriodicBoundconds"))
        return AC_FAILURE;
    return acDevicePeriodicBoundconds(node->device, stream, i3(0, 0, 0), node->grid.m);
}

AcResult
acNodeGeneralBoundcondStep(const Node node, const Stream stream,
                           const VertexBufferHandle vtxbuf_handle, const AcMeshInfo config)
{
    if (!valid(node, "acNodeGeneralBoundcondStep"))
        return AC_FAILURE;
    return acDeviceGeneralBoundcondStep(node->device, stream, vtxbuf_handle, i3(0, 0, 0),
                                        node->grid.m, config, i3(-1, -1, -1));
}

AcResult
acNodeGeneralBoundconds(const Node node, const Stream stream, const AcMeshInfo config)
{
    if (!valid(node, "acNodeGeneralBoundconds"))
        return AC_FAILURE;
    return acDeviceGeneralBoundconds(node->device, stream, i3(0, 0, 0), node->grid.m, config,
                                     i3(-1, -1, -1));
}

AcResult
acNodeReduceScal(const Node node, const Stream stream, const ReductionType rtype,
                 const VertexBufferHandle vtxbuf_handle, AcReal* result)
{
    if (acNodeSynchronizeStream(node, STREAM_ALL) != AC_SUCCESS)
        return AC_FAILURE;
    AcReal r = 0;
    if (acDeviceReduceScal(node->device, stream, rtype, vtxbuf_handle, &r) != AC_SUCCESS)
        return AC_FAILURE;
    *result = finalise(node, rtype, r);
    return AC_SUCCESS;
}

AcResult
acNodeReduceVec(const Node node, const Stream stream_type, const ReductionType rtype,
                const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                const VertexBufferHandle vtxbuf2, AcReal* result)
{
    if (acNodeSynchronizeStream(node, STREAM_ALL) != AC_SUCCESS)
        return AC_FAILURE;
    AcReal r = 0;
    if (acDeviceReduceVec(node->device, stream_type, rtype, vtxbuf0, vtxbuf1, vtxbuf2, &r) !=
        AC_SUCCESS)
        return AC_FAILURE;
    *result = finalise(node, rtype, r);
    return AC_SUCCESS;
}

AcResult
acNodeReduceVecScal(const Node node, const Stream stream_type, const ReductionType rtype,
                    const VertexBufferHandle vtxbuf0, const VertexBufferHandle vtxbuf1,
                    const VertexBufferHandle vtxbuf2, const VertexBufferHandle vtxbuf3,
                    AcReal* result)
{
    if (acNodeSynchronizeStream(node, STREAM_ALL) != AC_SUCCESS)
        return AC_FAILURE;
    AcReal r = 0;
    if (acDeviceReduceVecScal(node->device, stream_type, rtype, vtxbuf0, vtxbuf1, vtxbuf2,
                              vtxbuf3, &r) != AC_SUCCESS)
        return AC_FAILURE;
    *result = finalise(node, rtype, r);
    return AC_SUCCESS;
}

} // extern "C"
