/*
 * ORACLE — test infrastructure, not product code.
 *
 * Plain-pointer C entry points over the templated restatement in oracle.h, instantiated for
 * float (_f32), double (_f64) and long double (_f80), for the Python tests and the CPU
 * baseline. Fields are eight separate m^3 arrays (x fastest, 3 ghost cells per side), as in
 * the reference's AcMesh; real_params is the full AcRealParam array in double.
 */
#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

namespace {

template <typename T>
void rhs_all(const int* dims, T* const* fields, const double* rp, T* const* rates)
{
    const Grid<T> g(dims[0], dims[1], dims[2]);
    const Params<T> p = make_params<T>(rp);
#pragma omp parallel for schedule(static)
    for (int k = NGHOST_; k < NGHOST_ + g.nz; ++k)
        for (int j = NGHOST_; j < NGHOST_ + g.ny; ++j)
            for (int i = NGHOST_; i < NGHOST_ + g.nx; ++i) {
                T rate[NFIELDS];
                rhs_cell(g, fields, i, j, k, p, rate);
                for (int q = 0; q < NFIELDS; ++q)
                    rates[q][g.idx(i, j, k)] = rate[q];
            }
}

template <typename T>
void derivs_cell(const int* dims, const T* field, const double* inv_ds, int i, int j, int k, T* out)
{
    const Grid<T> g(dims[0], dims[1], dims[2]);
    Params<T> p{};
    p.inv_dsx = T(inv_ds[0]);
    p.inv_dsy = T(inv_ds[1]);
    p.inv_dsz = T(inv_ds[2]);
    const ScalarData<T> d = read_scal_data(g, field, i, j, k, p);
    out[0] = d.value;
    out[1] = d.grad[0];
    out[2] = d.grad[1];
    out[3] = d.grad[2];
    out[4] = d.hess[0][0];
    out[5] = d.hess[1][1];
    out[6] = d.hess[2][2];
    out[7] = d.hess[0][1];
    out[8] = d.hess[0][2];
    out[9] = d.hess[1][2];
}

} // namespace

#define ORACLE_INSTANTIATE(T, SFX)                                                                 \
    extern "C" void oracle_integrate_steps_##SFX(const int* dims, T* const* fields,              \
                                                 const double* rp, T dt, int nsteps)             \
    {                                                                                              \
        const Grid<T> g(dims[0], dims[1], dims[2]);                                                \
        const Params<T> p = make_params<T>(rp);                                                    \
        for (int s = 0; s < nsteps; ++s)                                                           \
            integrate_step(g, fields, p, dt);                                                      \
    }                                                                                              \
    extern "C" void oracle_integrate_steps_gpuform_##SFX(const int* dims, T* const* fields,     \
                                                         T* const* scratch, const double* rp,   \
                                                         T dt, int nsteps)                       \
    {                                                                                              \
        const Grid<T> g(dims[0], dims[1], dims[2]);                                                \
        const Params<T> p = make_params<T>(rp);                                                    \
        T* const* a = fields;                                                                      \
        T* const* b = scratch;                                                                     \
        for (int s = 0; s < nsteps; ++s)                                                           \
            if (gpu_form_step(g, a, b, p, dt) == 1) {                                              \
                T* const* t = a;                                                                   \
                a           = b;                                                                   \
                b           = t;                                                                   \
            }                                                                                      \
        if (a != fields)                                                                           \
            for (int q = 0; q < NFIELDS; ++q)                                                      \
                memcpy(fields[q], a[q], sizeof(T) * g.size());                                     \
    }                                                                                              \
    extern "C" void oracle_substep_gpuform_##SFX(const int* dims, T* const* in, T* const* out,   \
                                                 const double* rp, int step, T dt,                 \
                                                 const int* start, const int* end)                 \
    {                                                                                              \
        const Grid<T> g(dims[0], dims[1], dims[2]);                                                \
        gpu_form_substep(g, in, out, make_params<T>(rp), step, dt, start, end);                    \
    }                                                                                              \
    extern "C" void oracle_rhs_##SFX(const int* dims, T* const* fields, const double* rp,        \
                                     T* const* rates)                                              \
    {                                                                                              \
        rhs_all<T>(dims, fields, rp, rates);                                                       \
    }                                                                                              \
    extern "C" void oracle_periodic_##SFX(const int* dims, T* const* fields, int nfields)        \
    {                                                                                              \
        const Grid<T> g(dims[0], dims[1], dims[2]);                                                \
        apply_periodic(g, fields, nfields);                                                        \
    }                                                                                              \
    extern "C" int oracle_reduce_##SFX(const int* dims, int rtype, int nvals,                    \
                                       const T* const* vals, T* result)                            \
    {                                                                                              \
        const Grid<T> g(dims[0], dims[1], dims[2]);                                                \
        return reduce_vals(g, rtype, nvals, vals, result);                                         \
    }                                                                                              \
    extern "C" void oracle_derivs_##SFX(const int* dims, const T* field, const double* inv_ds,   \
                                        int i, int j, int k, T* out)                               \
    {                                                                                              \
        derivs_cell<T>(dims, field, inv_ds, i, j, k, out);                                         \
    }

ORACLE_INSTANTIATE(float, f32)
ORACLE_INSTANTIATE(double, f64)
ORACLE_INSTANTIATE(long double, f80)

extern "C" int
oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
