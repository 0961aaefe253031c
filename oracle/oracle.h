/*
 * ORACLE — test infrastructure, not product code.
 *
 * CPU restatement of the reference's host model solver, templated on the arithmetic type
 * T in {float, double, long double}. Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py load it, and only as the checker or the timed CPU baseline; the library
 * under astaroth_amd/ never links or calls it.
 *
 * Followed files (reference checkout, read-only):
 *   src/utils/modelsolver.c   FD operators, per-cell data, equations, alpha/beta passes
 *   src/utils/modelreduce.c   host reductions
 *   src/utils/memory.c        periodic ghost fill
 *   src/core/astaroth.cc      builtin parameters (m = n + 6, n_min, n_max, inv_ds)
 *   src/core/kernels/integration.cuh  two-buffer RK3 substep form of the GPU path
 *
 * Pinning status (what checks this restatement, see tests/test_oracle_*.py):
 *   FD operators          closed forms on sin(kx) and exactness on polynomials
 *   equations (RHS)       polynomial fields (FD exact) against an independent analytic
 *                         formulation (pressure as -grad(p)/rho, viscosity as
 *                         div(2 rho nu S)/rho, current as curl curl A)
 *   RK3 (host form and the GPU path's two-buffer form)
 *                         Beltrami decay against the amplification factor 1+z+z^2/2+z^3/6;
 *                         the two forms agree to rounding with each other and with long double
 *   periodic fill         numpy.pad(mode="wrap"), bitwise
 *   reductions            numpy max/min and compensated (math.fsum) sums
 *   error metric          known-answer values of ulp/abs/rel error
 * The reference's own C oracle needs generated headers (user_defines.h) and the CUDA runtime
 * header for its vector types; neither exists in this image, so it is not built here.
 */
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace oracle {

constexpr int NGHOST_ = 3;
constexpr int NFIELDS = 8;
enum Field { LNRHO = 0, UUX, UUY, UUZ, AX, AY, AZ, ENTROPY };

// Real-parameter indices used by the solver (values of the AcRealParam enum).
enum RealParamIdx {
    P_dsx = 0, P_dsy = 1, P_dsz = 2, P_inv_dsx = 3, P_inv_dsy = 4, P_inv_dsz = 5,
    P_nu_visc = 35, P_cs_sound = 36, P_eta = 37, P_mu0 = 38, P_cp_sound = 39, P_gamma = 40,
    P_lnT0 = 42, P_lnrho0 = 43, P_zeta = 44, P_cs2_sound = 74, P_NUM = 75
};

template <typename T> struct Grid {
    int nx, ny, nz, mx, my, mz;
    Grid(int nx_, int ny_, int nz_)
        : nx(nx_), ny(ny_), nz(nz_), mx(nx_ + 2 * NGHOST_), my(ny_ + 2 * NGHOST_),
          mz(nz_ + 2 * NGHOST_) {}
    size_t idx(int i, int j, int k) const
    {
        return (size_t)i + (size_t)j * mx + (size_t)k * (size_t)mx * my;
    }
    size_t size() const { return (size_t)mx * my * mz; }
};

// Parameters as the solver sees them: inv_ds recomputed from ds and cs2 = cs^2, as
// acHostIntegrateStep does on entry (modelsolver.c:993-997).
template <typename T> struct Params {
    T inv_dsx, inv_dsy, inv_dsz;
    T nu_visc, eta, mu0, cp_sound, gamma, lnT0, lnrho0, zeta, cs2_sound;
};

template <typename T> Params<T> make_params(const double* rp)
{
    Params<T> p;
    p.inv_dsx   = T(1.0) / T(rp[P_dsx]);
    p.inv_dsy   = T(1.0) / T(rp[P_dsy]);
    p.inv_dsz   = T(1.0) / T(rp[P_dsz]);
    p.cs2_sound = T(rp[P_cs_sound]) * T(rp[P_cs_sound]);
    p.nu_visc   = T(rp[P_nu_visc]);
    p.eta       = T(rp[P_eta]);
    p.mu0       = T(rp[P_mu0]);
    p.cp_sound  = T(rp[P_cp_sound]);
    p.gamma     = T(rp[P_gamma]);
    p.lnT0      = T(rp[P_lnT0]);
    p.lnrho0    = T(rp[P_lnrho0]);
    p.zeta      = T(rp[P_zeta]);
    return p;
}

// ---- FD operators, 6th order (modelsolver.c:101-207) ----
template <typename T> inline T first_derivative(const T* pencil, T inv_ds)
{
    const T c[] = {T(0), T(3.0 / 4.0), T(-3.0 / 20.0), T(1.0 / 60.0)};
    T res = 0;
    for (int i = 1; i <= 3; ++i)
        res += c[i] * (pencil[3 + i] - pencil[3 - i]);
    return res * inv_ds;
}

template <typename T> inline T second_derivative(const T* pencil, T inv_ds)
{
    const T c[] = {T(-49.0 / 18.0), T(3.0 / 2.0), T(-3.0 / 20.0), T(1.0 / 90.0)};
    T res = c[0] * pencil[3];
    for (int i = 1; i <= 3; ++i)
        res += c[i] * (pencil[3 + i] + pencil[3 - i]);
    return res * inv_ds * inv_ds;
}

template <typename T> inline T cross_derivative(const T* a, const T* b, T inv_ds_a, T inv_ds_b)
{
    const T fac = T(1. / 720.);
    const T c[] = {T(0) * fac, T(270.0) * fac, T(-27.0) * fac, T(2.0) * fac};
    T res = T(0.);
    for (int i = 1; i <= 3; ++i)
        res += c[i] * (a[3 + i] + a[3 - i] - b[3 + i] - b[3 - i]);
    return res * inv_ds_a * inv_ds_b;
}

// Pencil gathers along an axis or a diagonal (modelsolver.c:209-327). (di,dj,dk) is the step
// of the a-pencil; the b-pencil of a cross derivative flips the second axis.
template <typename T>
inline void pencil(const Grid<T>& g, const T* f, int i, int j, int k, int di, int dj, int dk,
                   T* out)
{
    for (int o = 0; o < 7; ++o) {
        const int s = o - 3;
        out[o]      = f[g.idx(i + s * di, j + s * dj, k + s * dk)];
    }
}

template <typename T> struct ScalarData {
    T value;
    T grad[3];
    T hess[3][3];
};

// value, gradient and full Hessian of one field (modelsolver.c:367-417)
template <typename T>
inline ScalarData<T> read_scal_data(const Grid<T>& g, const T* f, int i, int j, int k,
                                    const Params<T>& p)
{
    ScalarData<T> d;
    T pa[7], pb[7];
    d.value = f[g.idx(i, j, k)];
    pencil(g, f, i, j, k, 1, 0, 0, pa);
    d.grad[0]    = first_derivative(pa, p.inv_dsx);
    d.hess[0][0] = second_derivative(pa, p.inv_dsx);
    pencil(g, f, i, j, k, 0, 1, 0, pa);
    d.grad[1]    = first_derivative(pa, p.inv_dsy);
    d.hess[1][1] = second_derivative(pa, p.inv_dsy);
    pencil(g, f, i, j, k, 0, 0, 1, pa);
    d.grad[2]    = first_derivative(pa, p.inv_dsz);
    d.hess[2][2] = second_derivative(pa, p.inv_dsz);
    // xy: a along (+1,+1), b along (+1,-1) (modelsolver.c:231-247)
    pencil(g, f, i, j, k, 1, 1, 0, pa);
    pencil(g, f, i, j, k, 1, -1, 0, pb);
    d.hess[0][1] = cross_derivative(pa, pb, p.inv_dsx, p.inv_dsy);
    pencil(g, f, i, j, k, 1, 0, 1, pa);
    pencil(g, f, i, j, k, 1, 0, -1, pb);
    d.hess[0][2] = cross_derivative(pa, pb, p.inv_dsx, p.inv_dsz);
    pencil(g, f, i, j, k, 0, 1, 1, pa);
    pencil(g, f, i, j, k, 0, 1, -1, pb);
    d.hess[1][2] = cross_derivative(pa, pb, p.inv_dsy, p.inv_dsz);
    d.hess[1][0] = d.hess[0][1];
    d.hess[2][0] = d.hess[0][2];
    d.hess[2][1] = d.hess[1][2];
    return d;
}

template <typename T> struct V3 {
    T x, y, z;
};
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> struct M3 {
    V3<T> row[3];
};
template <typename T> inline V3<T> mul(const M3<T>& m, V3<T> x)
{
    return {dot(m.row[0], x), dot(m.row[1], x), dot(m.row[2], x)};
}

template <typename T> inline V3<T> gradient(const ScalarData<T>& d) { return {d.grad[0], d.grad[1], d.grad[2]}; }
template <typename T> inline T laplace(const ScalarData<T>& d) { return d.hess[0][0] + d.hess[1][1] + d.hess[2][2]; }

template <typename T> struct VectorData {
    ScalarData<T> x, y, z;
};

// Vector calculus (modelsolver.c:537-609)
template <typename T> inline V3<T> vecvalue(const VectorData<T>& v) { return {v.x.value, v.y.value, v.z.value}; }
template <typename T> inline T divergence(const VectorData<T>& v) { return v.x.grad[0] + v.y.grad[1] + v.z.grad[2]; }
template <typename T> inline V3<T> laplace_vec(const VectorData<T>& v) { return {laplace(v.x), laplace(v.y), laplace(v.z)}; }
template <typename T> inline V3<T> curl(const VectorData<T>& v)
{
    return {v.z.grad[1] - v.y.grad[2], v.x.grad[2] - v.z.grad[0], v.y.grad[0] - v.x.grad[1]};
}
template <typename T> inline V3<T> gradient_of_divergence(const VectorData<T>& v)
{
    return {v.x.hess[0][0] + v.y.hess[0][1] + v.z.hess[0][2],
            v.x.hess[1][0] + v.y.hess[1][1] + v.z.hess[1][2],
            v.x.hess[2][0] + v.y.hess[2][1] + v.z.hess[2][2]};
}
template <typename T> inline M3<T> gradients(const VectorData<T>& v)
{
    return {{gradient(v.x), gradient(v.y), gradient(v.z)}};
}
template <typename T> inline M3<T> stress_tensor(const VectorData<T>& v)
{
    M3<T> S;
    S.row[0].x = T(2. / 3.) * v.x.grad[0] - T(1. / 3.) * (v.y.grad[1] + v.z.grad[2]);
    S.row[0].y = T(1. / 2.) * (v.x.grad[1] + v.y.grad[0]);
    S.row[0].z = T(1. / 2.) * (v.x.grad[2] + v.z.grad[0]);
    S.row[1].y = T(2. / 3.) * v.y.grad[1] - T(1. / 3.) * (v.x.grad[0] + v.z.grad[2]);
    S.row[1].z = T(1. / 2.) * (v.y.grad[2] + v.z.grad[1]);
    S.row[2].z = T(2. / 3.) * v.z.grad[2] - T(1. / 3.) * (v.x.grad[0] + v.y.grad[1]);
    S.row[1].x = S.row[0].y;
    S.row[2].x = S.row[0].z;
    S.row[2].y = S.row[1].z;
    return S;
}
template <typename T> inline T contract(const M3<T>& m)
{
    T res = 0;
    for (int i = 0; i < 3; ++i)
        res += dot(m.row[i], m.row[i]);
    return res;
}

// ---- Equations (modelsolver.c:628-777; LENTROPY = LMAGNETIC = 1, LUPWD = LFORCING = 0) ----
template <typename T> inline T continuity(const VectorData<T>& uu, const ScalarData<T>& lnrho)
{
    return -dot(vecvalue(uu), gradient(lnrho)) - divergence(uu);
}

template <typename T>
inline V3<T> momentum(const VectorData<T>& uu, const ScalarData<T>& lnrho, const ScalarData<T>& ss,
                      const VectorData<T>& aa, const Params<T>& p)
{
    using std::exp;
    const M3<T> S = stress_tensor(uu);
    const T cs2   = p.cs2_sound *
                  exp(p.gamma * ss.value / p.cp_sound + (p.gamma - 1) * (lnrho.value - p.lnrho0));
    const V3<T> j      = (T(1.) / p.mu0) * (gradient_of_divergence(aa) - laplace_vec(aa));
    const V3<T> B      = curl(aa);
    const T inv_rho    = T(1.) / exp(lnrho.value);
    const V3<T> mom = -mul(gradients(uu), vecvalue(uu)) -
                      cs2 * ((T(1.) / p.cp_sound) * gradient(ss) + gradient(lnrho)) +
                      inv_rho * cross(j, B) +
                      p.nu_visc * (laplace_vec(uu) + T(1. / 3.) * gradient_of_divergence(uu) +
                                   T(2.) * mul(S, gradient(lnrho))) +
                      p.zeta * gradient_of_divergence(uu);
    return mom;
}

template <typename T> inline V3<T> induction(const VectorData<T>& uu, const VectorData<T>& aa, const Params<T>& p)
{
    const V3<T> B   = curl(aa);
    const V3<T> lap = laplace_vec(aa);
    return cross(vecvalue(uu), B) + p.eta * lap;
}

template <typename T> inline T lnT(const ScalarData<T>& ss, const ScalarData<T>& lnrho, const Params<T>& p)
{
    return p.lnT0 + p.gamma * ss.value / p.cp_sound + (p.gamma - T(1.)) * (lnrho.value - p.lnrho0);
}

template <typename T>
inline T heat_conduction(const ScalarData<T>& ss, const ScalarData<T>& lnrho, const Params<T>& p)
{
    using std::exp;
    const T K_THERMAL    = T(0.001); // modelsolver.c:43
    const T inv_cp_sound = T(1.) / p.cp_sound;
    const V3<T> grad_ln_chi = -gradient(lnrho);
    const T first_term = p.gamma * inv_cp_sound * laplace(ss) + (p.gamma - T(1.)) * laplace(lnrho);
    const V3<T> second_term = p.gamma * inv_cp_sound * gradient(ss) + (p.gamma - T(1.)) * gradient(lnrho);
    const V3<T> third_term  = p.gamma * (inv_cp_sound * gradient(ss) + gradient(lnrho)) + grad_ln_chi;
    const T chi = K_THERMAL / (exp(lnrho.value) * p.cp_sound);
    return p.cp_sound * chi * (first_term + dot(second_term, third_term));
}

template <typename T>
inline T entropy(const ScalarData<T>& ss, const VectorData<T>& uu, const ScalarData<T>& lnrho,
                 const VectorData<T>& aa, const Params<T>& p)
{
    using std::exp;
    const M3<T> S   = stress_tensor(uu);
    const T inv_pT  = T(1.) / (exp(lnrho.value) * exp(lnT(ss, lnrho, p)));
    const V3<T> j   = (T(1.) / p.mu0) * (gradient_of_divergence(aa) - laplace_vec(aa));
    const T H_CONST = 0, C_CONST = 0;
    const T RHS = H_CONST - C_CONST + p.eta * p.mu0 * dot(j, j) +
                  T(2.) * exp(lnrho.value) * p.nu_visc * contract(S) +
                  p.zeta * exp(lnrho.value) * divergence(uu) * divergence(uu);
    return -dot(vecvalue(uu), gradient(ss)) + inv_pT * RHS + heat_conduction(ss, lnrho, p);
}

// Rate of change of all eight fields at one interior cell (modelsolver.c:880-912).
template <typename T>
inline void rhs_cell(const Grid<T>& g, T* const* f, int i, int j, int k, const Params<T>& p,
                     T rate[NFIELDS])
{
    const ScalarData<T> lnrho = read_scal_data(g, f[LNRHO], i, j, k, p);
    VectorData<T> uu, aa;
    uu.x                  = read_scal_data(g, f[UUX], i, j, k, p);
    uu.y                  = read_scal_data(g, f[UUY], i, j, k, p);
    uu.z                  = read_scal_data(g, f[UUZ], i, j, k, p);
    aa.x                  = read_scal_data(g, f[AX], i, j, k, p);
    aa.y                  = read_scal_data(g, f[AY], i, j, k, p);
    aa.z                  = read_scal_data(g, f[AZ], i, j, k, p);
    const ScalarData<T> ss = read_scal_data(g, f[ENTROPY], i, j, k, p);

    rate[LNRHO]       = continuity(uu, lnrho);
    const V3<T> ind   = induction(uu, aa, p);
    rate[AX]          = ind.x;
    rate[AY]          = ind.y;
    rate[AZ]          = ind.z;
    const V3<T> mom   = momentum(uu, lnrho, ss, aa, p);
    rate[UUX]         = mom.x;
    rate[UUY]         = mom.y;
    rate[UUZ]         = mom.z;
    rate[ENTROPY]     = entropy(ss, uu, lnrho, aa, p);
}

// ---- Periodic ghost fill (memory.c:41-104) ----
template <typename T> void apply_periodic(const Grid<T>& g, T* const* f, int nfields = NFIELDS)
{
    const int nx_min = NGHOST_, ny_min = NGHOST_, nz_min = NGHOST_;
    const int nx_max = nx_min + g.nx, ny_max = ny_min + g.ny, nz_max = nz_min + g.nz;
    for (int w = 0; w < nfields; ++w) {
#pragma omp parallel for schedule(static)
        for (int k_dst = 0; k_dst < g.mz; ++k_dst)
            for (int j_dst = 0; j_dst < g.my; ++j_dst)
                for (int i_dst = 0; i_dst < g.mx; ++i_dst) {
                    if (i_dst >= nx_min && i_dst < nx_max && j_dst >= ny_min && j_dst < ny_max &&
                        k_dst >= nz_min && k_dst < nz_max)
                        continue;
                    const int i_src = (i_dst - nx_min + g.nx) % g.nx + nx_min;
                    const int j_src = (j_dst - ny_min + g.ny) % g.ny + ny_min;
                    const int k_src = (k_dst - nz_min + g.nz) % g.nz + nz_min;
                    f[w][g.idx(i_dst, j_dst, k_dst)] = f[w][g.idx(i_src, j_src, k_src)];
                }
    }
}

// ---- One RK3 step (modelsolver.c:879-948, 987-1041) ----
// Host form of Williamson's 2N scheme: w <- alpha_s w + dt F(u); u <- u + beta_s w, with a
// periodic fill of u before each substep. OpenMP over k-planes (each cell is independent, so
// the result does not depend on the thread count).
template <typename T>
void integrate_step(const Grid<T>& g, T* const* f, const Params<T>& p, T dt)
{
    const T alpha[] = {T(.0), T(-5. / 9.), T(-153. / 128.)};
    const T beta[]  = {T(1. / 3.), T(15. / 16.), T(8. / 15.)};
    std::vector<T> wbuf((size_t)NFIELDS * g.size(), T(0));
    T* w[NFIELDS];
    for (int q = 0; q < NFIELDS; ++q)
        w[q] = wbuf.data() + (size_t)q * g.size();

    for (int s = 0; s < 3; ++s) {
        apply_periodic(g, f);
#pragma omp parallel for schedule(static)
        for (int k = NGHOST_; k < NGHOST_ + g.nz; ++k)
            for (int j = NGHOST_; j < NGHOST_ + g.ny; ++j)
                for (int i = NGHOST_; i < NGHOST_ + g.nx; ++i) {
                    T rate[NFIELDS];
                    rhs_cell(g, f, i, j, k, p, rate);
                    const size_t idx = g.idx(i, j, k);
                    for (int q = 0; q < NFIELDS; ++q) {
                        if (s == 0)
                            w[q][idx] = rate[q] * dt;
                        else
                            w[q][idx] = alpha[s] * w[q][idx] + rate[q] * dt;
                    }
                }
#pragma omp parallel for schedule(static)
        for (int k = NGHOST_; k < NGHOST_ + g.nz; ++k)
            for (int j = NGHOST_; j < NGHOST_ + g.ny; ++j)
                for (int i = NGHOST_; i < NGHOST_ + g.nx; ++i) {
                    const size_t idx = g.idx(i, j, k);
                    for (int q = 0; q < NFIELDS; ++q)
                        f[q][idx] += beta[s] * w[q][idx];
                }
    }
}

// ---- Two-buffer RK3 substep of the reference's GPU path (src/core/kernels/integration.cuh:14-37,
// launched over [start, end) by integration.cuh:217-249 and device.cc:427-434) ----
// out = in + beta_0 F dt                                       (step 0)
// out = in + beta_s (alpha_s (1 / beta_{s-1}) (in - out) + F dt)  (steps 1, 2)
// `out` holds the previous stage on entry; ghost cells of `in` must be filled by the caller.
template <typename T>
void gpu_form_substep(const Grid<T>& g, T* const* in, T* const* out, const Params<T>& p, int step,
                      T dt, const int* start, const int* end)
{
    const T alpha[] = {T(0), T(.0), T(-5. / 9.), T(-153. / 128.)};
    const T beta[]  = {T(0), T(1. / 3.), T(15. / 16.), T(8. / 15.)};
#pragma omp parallel for schedule(static)
    for (int k = start[2]; k < end[2]; ++k)
        for (int j = start[1]; j < end[1]; ++j)
            for (int i = start[0]; i < end[0]; ++i) {
                T rate[NFIELDS];
                rhs_cell(g, in, i, j, k, p, rate);
                const size_t idx = g.idx(i, j, k);
                for (int q = 0; q < NFIELDS; ++q) {
                    const T cur = in[q][idx];
                    if (step == 0)
                        out[q][idx] = cur + beta[step + 1] * rate[q] * dt;
                    else
                        out[q][idx] = cur + beta[step + 1] * (alpha[step + 1] * (T(1.) / beta[step]) *
                                                                  (cur - out[q][idx]) +
                                                              rate[q] * dt);
                }
            }
}

// One RK3 step driven the way a single-device caller drives the GPU path:
// 3 x {periodic fill of in, substep over the interior, swap}. Returns the index (0 or 1) of the
// buffer set holding the new state.
template <typename T>
int gpu_form_step(const Grid<T>& g, T* const* a, T* const* b, const Params<T>& p, T dt)
{
    const int start[3] = {NGHOST_, NGHOST_, NGHOST_};
    const int end[3]   = {NGHOST_ + g.nx, NGHOST_ + g.ny, NGHOST_ + g.nz};
    T* const* in       = a;
    T* const* out      = b;
    for (int s = 0; s < 3; ++s) {
        apply_periodic(g, in);
        gpu_form_substep(g, in, out, p, s, dt, start, end);
        T* const* t = in;
        in          = out;
        out         = t;
    }
    return in == a ? 0 : 1;
}

// ---- Host reductions (modelreduce.c:76-209) ----
// rtype values: 0 MAX, 1 MIN, 2 RMS, 3 RMS_EXP, 7 SUM. MAX/MIN start from the first interior
// cell; RMS types return sqrt(sum / nxyz). Accumulation is serial, in T, in k-j-i order.
template <typename T> inline T red_max(T a, T b) { return a > b ? a : b; }
template <typename T> inline T red_min(T a, T b) { return a < b ? a : b; }

template <typename T>
int reduce_vals(const Grid<T>& g, int rtype, int nvals, const T* const* v, T* result)
{
    using std::exp;
    using std::sqrt;
    auto initial = [&](size_t idx) -> T {
        if (nvals == 1) {
            const T a = v[0][idx];
            switch (rtype) {
            case 0: case 1: case 7: return a;
            case 2: return a * a;
            case 3: return exp(a) * exp(a);
            }
        }
        else {
            const T a = v[0][idx], b = v[1][idx], c = v[2][idx];
            switch (rtype) {
            case 0: case 1: case 7: return sqrt(a * a + b * b + c * c);
            case 2: return a * a + b * b + c * c;
            case 3: return exp(a) * exp(a) + exp(b) * exp(b) + exp(c) * exp(c);
            }
        }
        return T(NAN);
    };
    if (!(rtype == 0 || rtype == 1 || rtype == 2 || rtype == 3 || rtype == 7))
        return 1;
    T res = (rtype == 0 || rtype == 1) ? initial(g.idx(NGHOST_, NGHOST_, NGHOST_)) : T(0);
    for (int k = NGHOST_; k < NGHOST_ + g.nz; ++k)
        for (int j = NGHOST_; j < NGHOST_ + g.ny; ++j)
            for (int i = NGHOST_; i < NGHOST_ + g.nx; ++i) {
                const T cur = initial(g.idx(i, j, k));
                if (rtype == 0)
                    res = red_max(res, cur);
                else if (rtype == 1)
                    res = red_min(res, cur);
                else
                    res = res + cur;
            }
    if (rtype == 2 || rtype == 3) {
        const T inv_n = T(1.0) / T((long long)g.nx * g.ny * g.nz);
        res           = sqrt(inv_n * res);
    }
    *result = res;
    return 0;
}

} // namespace oracle
