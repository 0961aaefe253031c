/*
 * ORACLE — test infrastructure, not product code.
 *
 * The reference's host utility API (include/astaroth_utils.h) in AcReal precision, built once
 * per precision (liboracle_fp32.so, liboracle_fp64.so) so that C test programs written against
 * the reference utilities link unchanged. Solver, reductions and fill forward to oracle.h;
 * the error metric restates src/utils/verification.c:30-168.
 */
#include "astaroth_utils.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

using namespace oracle;

namespace {

double*
real_params_as_double(const AcMeshInfo& info, double* rp)
{
    for (int i = 0; i < NUM_REAL_PARAMS; ++i)
        rp[i] = (double)info.real_params[i];
    return rp;
}

Grid<AcReal>
grid_of(const AcMeshInfo& info)
{
    return Grid<AcReal>(info.int_params[AC_nx], info.int_params[AC_ny], info.int_params[AC_nz]);
}

bool
is_valid(const AcReal a)
{
    return !isnan(a) && !isinf(a);
}

} // namespace

extern "C" {

/* memory.c:23-39 */
AcResult
acHostVertexBufferSet(const VertexBufferHandle handle, const AcReal value, AcMesh* mesh)
{
    const size_t n = acVertexBufferSize(mesh->info);
    for (size_t i = 0; i < n; ++i)
        mesh->vertex_buffer[handle][i] = value;
    return AC_SUCCESS;
}

AcResult
acHostMeshSet(const AcReal value, AcMesh* mesh)
{
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w)
        acHostVertexBufferSet((VertexBufferHandle)w, value, mesh);
    return AC_SUCCESS;
}

/* memory.c:41-104 */
AcResult
acHostMeshApplyPeriodicBounds(AcMesh* mesh)
{
    apply_periodic(grid_of(mesh->info), mesh->vertex_buffer);
    return AC_SUCCESS;
}

/* memory.c:107-111 */
AcResult
acHostMeshClear(AcMesh* mesh)
{
    return acHostMeshSet(0, mesh);
}

/* modelsolver.c:987-1041. The reference aborts when inv_ds or cs2 is not finite
 * (checkConfiguration, :981-984); here that is reported as AC_FAILURE. */
AcResult
acHostIntegrateStep(AcMesh mesh, const AcReal dt)
{
    double rp[NUM_REAL_PARAMS];
    const Params<AcReal> p = make_params<AcReal>(real_params_as_double(mesh.info, rp));
    if (!is_valid(p.inv_dsx) || !is_valid(p.inv_dsy) || !is_valid(p.inv_dsz) ||
        !is_valid(p.cs2_sound))
        return AC_FAILURE;
    integrate_step(grid_of(mesh.info), mesh.vertex_buffer, p, dt);
    return AC_SUCCESS;
}

/* modelreduce.c:76-139 */
AcReal
acHostReduceScal(const AcMesh mesh, const ReductionType rtype, const VertexBufferHandle a)
{
    const AcReal* v[1] = {mesh.vertex_buffer[a]};
    AcReal res         = NAN;
    if (reduce_vals(grid_of(mesh.info), (int)rtype, 1, v, &res)) {
        fprintf(stderr, "acHostReduceScal: unsupported rtype %d\n", (int)rtype);
        abort();
    }
    return res;
}

/* modelreduce.c:141-209 */
AcReal
acHostReduceVec(const AcMesh mesh, const ReductionType rtype, const VertexBufferHandle a,
                const VertexBufferHandle b, const VertexBufferHandle c)
{
    const AcReal* v[3] = {mesh.vertex_buffer[a], mesh.vertex_buffer[b], mesh.vertex_buffer[c]};
    AcReal res         = NAN;
    if (reduce_vals(grid_of(mesh.info), (int)rtype, 3, v, &res)) {
        fprintf(stderr, "acHostReduceVec: unsupported rtype %d\n", (int)rtype);
        abort();
    }
    return res;
}

/* verification.c:30-65. ulp = 2^(floor(log2|model|) - (p - 1)); rel_error in units of
 * 2^-p, p = 24 or 53. */
Error
acGetError(const AcReal model, const AcReal candidate)
{
    Error error;
    error.abs_error = 0;
    error.model     = model;
    error.candidate = candidate;
    if (error.model == error.candidate || fabsl((long double)model - (long double)candidate) == 0) {
        error.abs_error = 0;
        error.rel_error = 0;
        error.ulp_error = 0;
    }
    else if (!is_valid(error.model) || !is_valid(error.candidate)) {
        error.abs_error = INFINITY;
        error.rel_error = INFINITY;
        error.ulp_error = INFINITY;
    }
    else {
        const int p                       = sizeof(AcReal) == 4 ? 24 : 53;
        const long double e               = floorl(logl(fabsl((long double)model)) / logl(2));
        const long double ulp             = powl(2, e - (p - 1));
        const long double machine_epsilon = 0.5l * powl(2, -(p - 1));
        error.abs_error = fabsl((long double)model - (long double)candidate);
        error.ulp_error = error.abs_error / ulp;
        error.rel_error = fabsl(1.0l - (long double)candidate / (long double)model) /
                          machine_epsilon;
    }
    error.maximum_magnitude = error.minimum_magnitude = 0;
    return error;
}

/* verification.c:78-102. The reference also appends every verdict to ./verification.out;
 * here that happens only when AC_VERIFICATION_FILE names a file. */
bool
acEvalError(const char* label, const Error error)
{
    const long double max_ulp_error = 5;
    bool acceptable;
    if (error.ulp_error < max_ulp_error)
        acceptable = true;
    else if (error.abs_error < error.minimum_magnitude * AC_REAL_EPSILON)
        acceptable = true;
    else
        acceptable = false;

    printf("%-15s... %s ", label, acceptable ? "OK! " : "FAIL! ");
    printf("| %.3Lg (abs), %.3Lg (ulps), %.3Lg (rel). Range: [%.3g, %.3g]\n", error.abs_error,
           error.ulp_error, error.rel_error, (double)error.minimum_magnitude,
           (double)error.maximum_magnitude);
    const char* path = getenv("AC_VERIFICATION_FILE");
    if (path && path[0]) {
        FILE* file = fopen(path, "a");
        if (file) {
            fprintf(file, "%s, %Lg, %Lg, %Lg, %g, %g\n", label, error.ulp_error, error.abs_error,
                    error.rel_error, (double)error.maximum_magnitude,
                    (double)error.minimum_magnitude);
            fclose(file);
        }
    }
    return acceptable;
}

/* verification.c:104-168 */
AcResult
acVerifyMesh(const char* label, const AcMesh model, const AcMesh candidate)
{
    printf("---Test: %s---\n", label);
    printf("Errors at the point of the maximum absolute error:\n");
    const size_t n   = acVertexBufferSize(model.info);
    int errors_found = 0;
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        const AcReal* m = model.vertex_buffer[w];
        const AcReal* c = candidate.vertex_buffer[w];
        Error error;
        memset(&error, 0, sizeof(error));
        error.abs_error = -1;
        AcReal maxmag = -INFINITY, minmag = INFINITY;
        for (size_t i = 0; i < n; ++i) {
            const Error cur = acGetError(m[i], c[i]);
            if (cur.abs_error > error.abs_error)
                error = cur;
            const AcReal a = fabs(m[i]);
            maxmag         = a > maxmag ? a : maxmag;
            minmag         = a < minmag ? a : minmag;
        }
        error.maximum_magnitude = maxmag;
        error.minimum_magnitude = minmag;
        if (!acEvalError(vtxbuf_names[w], error))
            ++errors_found;
    }
    if (errors_found > 0)
        printf("Failure. Found %d errors\n", errors_found);
    fflush(stdout);
    return errors_found ? AC_FAILURE : AC_SUCCESS;
}

} // extern "C"
