// Host helpers of include/astaroth.h and the configuration loader of include/astaroth_utils.h.
//
//  acHostUpdateBuiltinParams  ref src/core/astaroth.cc:190-227
//  acHostMeshCreate/Destroy   ref src/core/astaroth.cc:229-266
//  acHostMeshRandomize        same contract (every cell of every field in [0, 1)), portable
//                             splitmix64 stream instead of the C library's rand()
//  acLoadConfig               ref src/utils/config_loader.c:49-112
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "astaroth.h"
#include "astaroth_utils.h"

extern "C" {

AcResult
acHostUpdateBuiltinParams(AcMeshInfo* config)
{
    config->int_params[AC_mx] = config->int_params[AC_nx] + STENCIL_ORDER;
    config->int_params[AC_my] = config->int_params[AC_ny] + STENCIL_ORDER;
    config->int_params[AC_mz] = config->int_params[AC_nz] + STENCIL_ORDER;

    config->int_params[AC_nx_min] = NGHOST;
    config->int_params[AC_ny_min] = NGHOST;
    config->int_params[AC_nz_min] = NGHOST;
    config->int_params[AC_nx_max] = NGHOST + config->int_params[AC_nx];
    config->int_params[AC_ny_max] = NGHOST + config->int_params[AC_ny];
    config->int_params[AC_nz_max] = NGHOST + config->int_params[AC_nz];

    config->real_params[AC_inv_dsx] = (AcReal)(1.) / config->real_params[AC_dsx];
    config->real_params[AC_inv_dsy] = (AcReal)(1.) / config->real_params[AC_dsy];
    config->real_params[AC_inv_dsz] = (AcReal)(1.) / config->real_params[AC_dsz];

    config->int_params[AC_mxy]  = config->int_params[AC_mx] * config->int_params[AC_my];
    config->int_params[AC_nxy]  = config->int_params[AC_nx] * config->int_params[AC_ny];
    config->int_params[AC_nxyz] = config->int_params[AC_nxy] * config->int_params[AC_nz];
    return AC_SUCCESS;
}

AcResult
acHostMeshCreate(const AcMeshInfo info, AcMesh* mesh)
{
    mesh->info         = info;
    const size_t bytes = acVertexBufferSizeBytes(info);
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        mesh->vertex_buffer[w] = (AcReal*)malloc(bytes);
        if (!mesh->vertex_buffer[w]) {
            fprintf(stderr, "acHostMeshCreate: out of host memory (%zu bytes)\n", bytes);
            abort();
        }
    }
    return AC_SUCCESS;
}

static uint64_t
splitmix64_next(uint64_t* state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

AcResult
acHostMeshRandomize(AcMesh* mesh)
{
    const size_t n = acVertexBufferSize(mesh->info);
    uint64_t state = 123456789ULL;
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w)
        for (size_t i = 0; i < n; ++i)
            mesh->vertex_buffer[w][i] =
                (AcReal)((double)(splitmix64_next(&state) >> 11) * (1.0 / 9007199254740992.0));
    return AC_SUCCESS;
}

AcResult
acHostMeshDestroy(AcMesh* mesh)
{
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        free(mesh->vertex_buffer[w]);
        mesh->vertex_buffer[w] = NULL;
    }
    return AC_SUCCESS;
}

static int
find_name(const char* key, const char* const* names, int n)
{
    for (int i = 0; i < n; ++i)
        if (!strcmp(key, names[i]))
            return i;
    return -1;
}

AcResult
acLoadConfig(const char* config_path, AcMeshInfo* config)
{
    if (!config_path || !config) {
        fprintf(stderr, "acLoadConfig: null argument\n");
        abort();
    }
    memset(config, 0xFF, sizeof(*config));
    FILE* fp = fopen(config_path, "r");
    if (!fp) {
        fprintf(stderr, "acLoadConfig: cannot open %s\n", config_path);
        abort();
    }
    char keyword[128], value[128];
    int matched;
    // "%s = %s" over whitespace-separated tokens: comment tokens never match a parameter name.
    while ((matched = fscanf(fp, "%127s = %127s", keyword, value)) != EOF) {
        if (matched < 2)
            continue;
        int idx;
        if ((idx = find_name(keyword, intparam_names, NUM_INT_PARAMS)) >= 0)
            config->int_params[idx] = atoi(value);
        else if ((idx = find_name(keyword, realparam_names, NUM_REAL_PARAMS)) >= 0)
            config->real_params[idx] = (AcReal)atof(value);
    }
    fclose(fp);
    acHostUpdateBuiltinParams(config);

    AcResult retval = AC_SUCCESS;
    const uint32_t* words = (const uint32_t*)config;
    for (size_t i = 0; i < sizeof(*config) / sizeof(uint32_t); ++i)
        if (words[i] == 0xFFFFFFFFu)
            retval = AC_FAILURE;
    return retval;
}

} // extern "C"

// CFL time step of the reference's standalone driver (samples/standalone/model/host_timestep.cc:
// 34-56, evaluated in long double as there):
//   dt = min(cdt dsmin / (|u_max| + sqrt(cs^2 + v_A^2)),
//            cdtv dsmin^2 / max(max(nu, eta), max(gamma, chi)))   with chi = 0,
// dsmin = AC_dsmin, or min(dsx, dsy, dsz) when AC_dsmin is unset (the standalone configuration
// loader sets it so, config_loader.cc:102).
namespace ac {
double
cfl_timestep(double umax, double vamax, const AcMeshInfo& info)
{
    typedef long double L;
    const L cdt   = info.real_params[AC_cdt];
    const L cdtv  = info.real_params[AC_cdtv];
    const L cs2   = info.real_params[AC_cs2_sound];
    const L nu    = info.real_params[AC_nu_visc];
    const L eta   = info.real_params[AC_eta];
    const L gamma = info.real_params[AC_gamma];
    const L chi   = 0;
    L dsmin       = info.real_params[AC_dsmin];
    if (!(dsmin > 0)) {
        dsmin = info.real_params[AC_dsx];
        if ((L)info.real_params[AC_dsy] < dsmin)
            dsmin = info.real_params[AC_dsy];
        if ((L)info.real_params[AC_dsz] < dsmin)
            dsmin = info.real_params[AC_dsz];
    }
    const L va     = vamax;
    const L uu_dt  = cdt * dsmin / (fabsl((L)umax) + sqrtl(cs2 + va * va));
    const L m1     = nu > eta ? nu : eta;
    const L m2     = gamma > chi ? gamma : chi;
    const L visc_dt = cdtv * dsmin * dsmin / (m1 > m2 ? m1 : m2);
    return (double)(uu_dt < visc_dt ? uu_dt : visc_dt);
}
} // namespace ac
