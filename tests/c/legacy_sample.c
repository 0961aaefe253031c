/* A C program in the style of the reference's samples/ctest, written against include/astaroth.h
 * (the legacy ac* interface, or with a second argument "node" the same sequence through the Node
 * interface). The CPU test suite compiles and links it; the GPU suite runs it and compares its
 * printed numbers with the Device-level results. */
#include <stdio.h>
#include <stdlib.h>

#include "astaroth.h"

int
main(int argc, char** argv)
{
    const int n = argc > 1 ? atoi(argv[1]) : 24;
    AcMeshInfo info;
    for (int i = 0; i < NUM_INT_PARAMS; ++i)
        info.int_params[i] = 0;
    for (int i = 0; i < NUM_REAL_PARAMS; ++i)
        info.real_params[i] = 0;
    info.int_params[AC_nx]           = n;
    info.int_params[AC_ny]           = n;
    info.int_params[AC_nz]           = n;
    info.real_params[AC_dsx]         = (AcReal)(6.283185307179586 / n);
    info.real_params[AC_dsy]         = (AcReal)(6.283185307179586 / n);
    info.real_params[AC_dsz]         = (AcReal)(6.283185307179586 / n);
    info.real_params[AC_nu_visc]     = (AcReal)5e-3;
    info.real_params[AC_eta]         = (AcReal)5e-3;
    info.real_params[AC_cs_sound]    = (AcReal)1.0;
    info.real_params[AC_cs2_sound]   = (AcReal)1.0;
    info.real_params[AC_gamma]       = (AcReal)(5.0 / 3.0);
    info.real_params[AC_zeta]        = (AcReal)0.01;
    info.real_params[AC_mu0]         = (AcReal)1.4;
    info.real_params[AC_cp_sound]    = (AcReal)1.0;
    info.real_params[AC_lnT0]        = (AcReal)1.2;
    info.real_params[AC_lnrho0]      = (AcReal)1.3;
    acHostUpdateBuiltinParams(&info);

    AcMesh mesh;
    acHostMeshCreate(info, &mesh);
    acHostMeshRandomize(&mesh);
    /* small amplitudes: lnrho, u, A ~ 1e-3 (U[0,1) - 0.5), entropy 0 */
    for (int q = 0; q < NUM_VTXBUF_HANDLES; ++q)
        for (size_t i = 0; i < acVertexBufferSize(info); ++i)
            mesh.vertex_buffer[q][i] = q == VTXBUF_ENTROPY ? 0 : (mesh.vertex_buffer[q][i] - (AcReal)0.5) * (AcReal)2e-3;

    if (acCheckDeviceAvailability() != AC_SUCCESS) {
        printf("no device\n");
        return 0;
    }
    AcReal umax, urms, lrmin, arms;
    if (argc > 2 && argv[2][0] == 'n') { /* Node interface */
        Node node;
        if (acNodeCreate(0, info, &node) != AC_SUCCESS)
            return 1;
        DeviceConfiguration dc;
        if (acNodeQueryDeviceConfiguration(node, &dc) != AC_SUCCESS || dc.num_devices != 1 ||
            dc.subgrid.n.z != n)
            return 1;
        acNodeLoadMesh(node, STREAM_DEFAULT, mesh);
        acNodePeriodicBoundconds(node, STREAM_DEFAULT);
        for (int s = 0; s < 2; ++s)
            acNodeIntegrate(node, (AcReal)1e-3);
        acNodePeriodicBoundconds(node, STREAM_DEFAULT);
        acNodeStoreMesh(node, STREAM_DEFAULT, &mesh);
        acNodeSynchronizeStream(node, STREAM_ALL);
        acNodeReduceVec(node, STREAM_DEFAULT, RTYPE_MAX, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ, &umax);
        acNodeReduceVec(node, STREAM_DEFAULT, RTYPE_RMS, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ, &urms);
        acNodeReduceScal(node, STREAM_DEFAULT, RTYPE_MIN, VTXBUF_LNRHO, &lrmin);
        acNodeReduceVecScal(node, STREAM_DEFAULT, RTYPE_ALFVEN_RMS, VTXBUF_AX, VTXBUF_AY,
                            VTXBUF_AZ, VTXBUF_LNRHO, &arms);
        acNodeDestroy(node);
    }
    else {
        if (acInit(info) != AC_SUCCESS)
            return 1;
        acLoad(mesh);
        acBoundcondStep();
        for (int s = 0; s < 2; ++s)
            acIntegrate((AcReal)1e-3);
        acBoundcondStep();
        acStore(&mesh);
        if (!acGetNode())
            return 1;
        umax  = acReduceVec(RTYPE_MAX, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ);
        urms  = acReduceVec(RTYPE_RMS, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ);
        lrmin = acReduceScal(RTYPE_MIN, VTXBUF_LNRHO);
        arms  = acReduceVecScal(RTYPE_ALFVEN_RMS, VTXBUF_AX, VTXBUF_AY, VTXBUF_AZ, VTXBUF_LNRHO);
        acQuit();
    }
    printf("%.17g %.17g %.17g %.17g %.17g\n", (double)umax, (double)urms, (double)lrmin,
           (double)arms, (double)mesh.vertex_buffer[VTXBUF_UUX][acVertexBufferIdx(5, 6, 7, info)]);
    acHostMeshDestroy(&mesh);
    return 0;
}
