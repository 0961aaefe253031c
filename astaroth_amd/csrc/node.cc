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
//   acNodeIntegrate / acNodeIntegrateGBC: per substep the interior is split as node.cc:729-880
//     splits it, into the inner region [2 NGHOST, n) and six outer slabs. The inner region reads
//     no ghost cell, so it runs on STREAM_16 while the ghost shell is filled on STREAM_0; the
//     slabs run on STREAM_0..5 once the fill has completed, then the buffers swap. (The reference
//     fills before it starts the inner region: what its inner region overlaps is the exchange
//     between GPUs, which one GPU does not have.) The result is that of one launch over the
//     interior to rounding: the z-march accumulates a cell's cross-derivative sums in the phase
//     of its unrolled plane loop that the column's first plane sets, and the phases are compiled
//     separately, so a few cells differ in their last bits. Boxes with an axis shorter than
//     2 NGHOST integrate the interior in one launch. Where the Device reads the ghost cells'
//     periodic images instead of filling them (acDeviceSubstepReadsPeriodicImages), there is no
//     fill to overlap and acNodeIntegrate runs each substep as one such launch.
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

// The step of node.cc:729-880 on one Device; `fill` fills the ghost shell of the `in` buffers
// on the given stream.
template <typename Fill>
AcResult
split_step(const Node node, const AcReal dt, Fill fill)
{
    const Device d = node->device;
    const int3 n   = node->grid.n;
    const int N    = NGHOST;
    const bool split = n.x >= 2 * N && n.y >= 2 * N && n.z >= 2 * N;
    if (acDeviceSynchronizeStream(d, STREAM_ALL) != AC_SUCCESS)
        return AC_FAILURE;
    for (int s = 0; s < 3; ++s) {
        if (fill(STREAM_0) != AC_SUCCESS)
            return AC_FAILURE;
        if (!split) {
            if (acDeviceIntegrateSubstep(d, STREAM_0, s, i3(N, N, N),
                                         i3(n.x + N, n.y + N, n.z + N), dt) != AC_SUCCESS)
                return AC_FAILURE;
        }
        else {
            // inner region: reads cells [N, n + N) only, concurrently with the fill
            if (acDeviceIntegrateSubstep(d, STREAM_16, s, i3(2 * N, 2 * N, 2 * N), n, dt) !=
                AC_SUCCESS)
                return AC_FAILURE;
            if (acDeviceSynchronizeStream(d, STREAM_0) != AC_SUCCESS) // the ghost shell is filled
                return AC_FAILURE;
            // The slabs are thin: z-march columns 14 planes deep give their launches enough
            // blocks to spread over the GPU (a whole-depth column per 16 x 16 tile leaves a side
            // slab of a 512^3 box 32 blocks). The inner launch has already taken its own depth.
            AcKernelConfig whole;
            if (acDeviceGetKernelConfig(d, &whole) != AC_SUCCESS)
                return AC_FAILURE;
            AcKernelConfig thin = whole;
            if (thin.variant != 0)
                thin.z_depth = 14;
            if (acDeviceSetKernelConfig(d, thin) != AC_SUCCESS)
                return AC_FAILURE;
            const struct {
                Stream stream;
                int3 start, end;
            } slabs[6] = {
                {STREAM_0, i3(N, N, N), i3(n.x + N, n.y + N, 2 * N)},            // front
                {STREAM_1, i3(N, N, n.z), i3(n.x + N, n.y + N, n.z + N)},        // back
                {STREAM_2, i3(N, N, 2 * N), i3(n.x + N, 2 * N, n.z)},            // bottom
                {STREAM_3, i3(N, n.y, 2 * N), i3(n.x + N, n.y + N, n.z)},        // top
                {STREAM_4, i3(N, 2 * N, 2 * N), i3(2 * N, n.y, n.z)},            // left
                {STREAM_5, i3(n.x, 2 * N, 2 * N), i3(n.x + N, n.y, n.z)},        // right
            };
            AcResult r = AC_SUCCESS;
            for (const auto& sl : slabs)
                if (r == AC_SUCCESS)
                    r = acDeviceIntegrateSubstep(d, sl.stream, s, sl.start, sl.end, dt);
            if (acDeviceSetKernelConfig(d, whole) != AC_SUCCESS || r != AC_SUCCESS)
                return AC_FAILURE;
        }
        if (acDeviceSynchronizeStream(d, STREAM_ALL) != AC_SUCCESS ||
            acDeviceSwapBuffers(d) != AC_SUCCESS)
            return AC_FAILURE;
    }
    return AC_SUCCESS;
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
    if (acDeviceSubstepReadsPeriodicImages(node->device)) {
        // the Device's build reads the ghost cells' periodic images: no fill to wait for or to
        // overlap, so each substep is one launch over the interior
        if (acDeviceSynchronizeStream(node->device, STREAM_ALL) != AC_SUCCESS)
            return AC_FAILURE;
        for (int s = 0; s < 3; ++s)
            if (acDeviceIntegrateSubstepPeriodic(node->device, STREAM_0, s, dt) != AC_SUCCESS ||
                acDeviceSwapBuffers(node->device) != AC_SUCCESS)
                return AC_FAILURE;
        return acDeviceSynchronizeStream(node->device, STREAM_ALL);
    }
    const int3 m = node->grid.m;
    return split_step(node, dt, [&](Stream s) {
        return acDevicePeriodicBoundconds(node->device, s, i3(0, 0, 0), m);
    });
}

AcResult
acNodeIntegrateGBC(const Node node, const AcMeshInfo config, const AcReal dt)
{
    if (!valid(node, "acNodeIntegrateGBC"))
        return AC_FAILURE;
    const int3 m = node->grid.m;
    return split_step(node, dt, [&](Stream s) {
        return acDeviceGeneralBoundconds(node->device, s, i3(0, 0, 0), m, config,
                                         i3(-1, -1, -1));
    });
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
