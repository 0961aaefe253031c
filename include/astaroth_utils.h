/*
 * astaroth_utils.h — host-side utilities of the reference (include/astaroth_utils.h:25-77).
 *
 * acLoadConfig is implemented by the library (libastaroth_*.so). The host model solver, host
 * reductions, host periodic fill and the error metric are implemented by the oracle
 * (oracle/liboracle_*.so), a CPU restatement used as the correctness checker; the library
 * itself never calls them.
 */
#pragma once
#include "astaroth.h"

#ifdef __cplusplus
extern "C" {
#endif

#include <stdbool.h>

/* ref include/astaroth_utils.h:34-42 */
typedef struct {
    AcReal model;
    AcReal candidate;
    long double abs_error;
    long double ulp_error;
    long double rel_error;
    AcReal maximum_magnitude;
    AcReal minimum_magnitude;
} Error;

/* ref :45. Reads `key = value` lines; every parameter that is not set stays 0xFF..FF (NaN for
 * reals) and the call then returns AC_FAILURE. Builtin parameters are derived afterwards. */
AcResult acLoadConfig(const char* config_path, AcMeshInfo* config);

/* ref :48-57 (oracle) */
AcResult acHostVertexBufferSet(const VertexBufferHandle handle, const AcReal value, AcMesh* mesh);
AcResult acHostMeshSet(const AcReal value, AcMesh* mesh);
AcResult acHostMeshApplyPeriodicBounds(AcMesh* mesh);
AcResult acHostMeshClear(AcMesh* mesh);

/* ref :60 (oracle). One RK3 step of the host model solver, in AcReal. */
AcResult acHostIntegrateStep(AcMesh mesh, const AcReal dt);

/* ref :63-67 (oracle). RMS types return sqrt(sum / n^3). */
AcReal acHostReduceScal(const AcMesh mesh, const ReductionType rtype, const VertexBufferHandle a);
AcReal acHostReduceVec(const AcMesh mesh, const ReductionType rtype, const VertexBufferHandle a,
                       const VertexBufferHandle b, const VertexBufferHandle c);

/* ref :69-73 (oracle). Error at the point of maximum absolute error; accepted when below
 * 5 ulps, or below min|model| * AC_REAL_EPSILON. */
Error acGetError(const AcReal model, const AcReal candidate);
bool acEvalError(const char* label, const Error error);
AcResult acVerifyMesh(const char* label, const AcMesh model, const AcMesh candidate);

#ifdef __cplusplus
} /* extern "C" */
#endif
