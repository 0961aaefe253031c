// Per-cell right-hand side of the 8-field compressible resistive MHD system and the low-storage
// RK3 update, shared by every solve-kernel variant.
//
// Semantics follow the reference's DSL module: continuity, momentum, induction, lnT,
// heat_conduction and entropy of acc/mhd_solver/stencil_kernel.ac:355-528 (LENTROPY =
// LMAGNETIC = 1, LUPWD = LFORCING = LSINK = 0), the vector calculus of
// acc/stdlib/stdderiv.h:266-335, and rk3_integrate of src/core/kernels/integration.cuh:14-37.
// The arithmetic is rearranged for the GPU: reciprocal constants are precomputed, the grid
// spacing is folded into the stencil weights, and exp(lnT) reuses the exponential already
// needed for cs^2 (exp(lnT) = exp(lnT0) * exp(gamma s / cp + (gamma - 1)(lnrho - lnrho0))).
#pragma once
#include "ac_internal.h"

namespace ac {

// Value and the derivatives the equations use for one field at one cell.
template <typename T> struct Fd {
    T v;
    T x, y, z;
    T xx, yy, zz;
    T xy, xz, yz;
};

template <typename T> __device__ __forceinline__ T fast_exp(T a);
template <> __device__ __forceinline__ float fast_exp<float>(float a) { return __expf(a); }
template <> __device__ __forceinline__ double fast_exp<double>(double a) { return exp(a); }

template <typename T> __device__ __forceinline__ T fast_rcp(T a);
template <> __device__ __forceinline__ float fast_rcp<float>(float a)
{
    return __builtin_amdgcn_rcpf(a);
}
template <> __device__ __forceinline__ double fast_rcp<double>(double a) { return 1.0 / a; }

// rate[q] = dF_q/dt for q in VertexBufferHandle order.
__device__ __forceinline__ void
mhd_rates(const Fd<AcReal>& lr, const Fd<AcReal>& ux, const Fd<AcReal>& uy, const Fd<AcReal>& uz,
          const Fd<AcReal>& ax, const Fd<AcReal>& ay, const Fd<AcReal>& az, const Fd<AcReal>& ss,
          const PhysParams& p, AcReal rate[NF])
{
    typedef AcReal T;
    // continuity (stencil_kernel.ac:355-367)
    const T divu = ux.x + uy.y + uz.z;
    rate[VTXBUF_LNRHO] = -(ux.v * lr.x + uy.v * lr.y + uz.v * lr.z) - divu;

    // traceless rate-of-strain tensor (stdderiv.h:300-323)
    const T S00 = T(2. / 3.) * ux.x - T(1. / 3.) * (uy.y + uz.z);
    const T S11 = T(2. / 3.) * uy.y - T(1. / 3.) * (ux.x + uz.z);
    const T S22 = T(2. / 3.) * uz.z - T(1. / 3.) * (ux.x + uy.y);
    const T S01 = T(0.5) * (ux.y + uy.x);
    const T S02 = T(0.5) * (ux.z + uz.x);
    const T S12 = T(0.5) * (uy.z + uz.y);

    // grad(div u), lap u, grad(div A), lap A (stdderiv.h:266-297)
    const T gdu_x = ux.xx + uy.xy + uz.xz;
    const T gdu_y = ux.xy + uy.yy + uz.yz;
    const T gdu_z = ux.xz + uy.yz + uz.zz;
    const T lapu_x = ux.xx + ux.yy + ux.zz;
    const T lapu_y = uy.xx + uy.yy + uy.zz;
    const T lapu_z = uz.xx + uz.yy + uz.zz;
    const T gda_x = ax.xx + ay.xy + az.xz;
    const T gda_y = ax.xy + ay.yy + az.yz;
    const T gda_z = ax.xz + ay.yz + az.zz;
    const T lapa_x = ax.xx + ax.yy + ax.zz;
    const T lapa_y = ay.xx + ay.yy + ay.zz;
    const T lapa_z = az.xx + az.yy + az.zz;

    // current density j = (grad div A - lap A) / mu0 and B = curl A
    const T jx = p.inv_mu0 * (gda_x - lapa_x);
    const T jy = p.inv_mu0 * (gda_y - lapa_y);
    const T jz = p.inv_mu0 * (gda_z - lapa_z);
    const T Bx = az.y - ay.z;
    const T By = ax.z - az.x;
    const T Bz = ay.x - ax.y;

    // thermodynamics: E = exp(gamma s / cp + (gamma - 1)(lnrho - lnrho0))
    const T rho     = fast_exp<T>(lr.v);
    const T inv_rho = fast_rcp<T>(rho);
    const T E       = fast_exp<T>(p.gamma_inv_cp * ss.v + p.gamma_m1 * (lr.v - p.lnrho0));
    const T cs2     = p.cs2_sound * E;

    // momentum (stencil_kernel.ac:370-403)
    {
        const T jxB_x = jy * Bz - jz * By;
        const T jxB_y = jz * Bx - jx * Bz;
        const T jxB_z = jx * By - jy * Bx;
        const T Sgl_x = S00 * lr.x + S01 * lr.y + S02 * lr.z;
        const T Sgl_y = S01 * lr.x + S11 * lr.y + S12 * lr.z;
        const T Sgl_z = S02 * lr.x + S12 * lr.y + S22 * lr.z;
        rate[VTXBUF_UUX] = -(ux.x * ux.v + ux.y * uy.v + ux.z * uz.v) -
                           cs2 * (p.inv_cp * ss.x + lr.x) + inv_rho * jxB_x +
                           p.nu_visc * (lapu_x + T(1. / 3.) * gdu_x + T(2.) * Sgl_x) +
                           p.zeta * gdu_x;
        rate[VTXBUF_UUY] = -(uy.x * ux.v + uy.y * uy.v + uy.z * uz.v) -
                           cs2 * (p.inv_cp * ss.y + lr.y) + inv_rho * jxB_y +
                           p.nu_visc * (lapu_y + T(1. / 3.) * gdu_y + T(2.) * Sgl_y) +
                           p.zeta * gdu_y;
        rate[VTXBUF_UUZ] = -(uz.x * ux.v + uz.y * uy.v + uz.z * uz.v) -
                           cs2 * (p.inv_cp * ss.z + lr.z) + inv_rho * jxB_z +
                           p.nu_visc * (lapu_z + T(1. / 3.) * gdu_z + T(2.) * Sgl_z) +
                           p.zeta * gdu_z;
    }

    // induction, Weyl gauge without the grad(div A) term (stencil_kernel.ac:463-481)
    rate[VTXBUF_AX] = (uy.v * Bz - uz.v * By) + p.eta * lapa_x;
    rate[VTXBUF_AY] = (uz.v * Bx - ux.v * Bz) + p.eta * lapa_y;
    rate[VTXBUF_AZ] = (ux.v * By - uy.v * Bx) + p.eta * lapa_z;

    // entropy with heat conduction (stencil_kernel.ac:484-528)
    {
        const T inv_pT   = inv_rho * fast_rcp<T>(p.exp_lnT0 * E);
        const T contract = S00 * S00 + S11 * S11 + S22 * S22 +
                           T(2.) * (S01 * S01 + S02 * S02 + S12 * S12);
        const T heating = p.eta_mu0 * (jx * jx + jy * jy + jz * jz) +
                          T(2.) * rho * p.nu_visc * contract + p.zeta * rho * divu * divu;
        const T lap_s  = ss.xx + ss.yy + ss.zz;
        const T lap_lr = lr.xx + lr.yy + lr.zz;
        // grad lnT = (gamma/cp) grad s + (gamma - 1) grad lnrho; the reference's second and
        // third terms of heat_conduction are both this vector.
        const T glT_x = p.gamma_inv_cp * ss.x + p.gamma_m1 * lr.x;
        const T glT_y = p.gamma_inv_cp * ss.y + p.gamma_m1 * lr.y;
        const T glT_z = p.gamma_inv_cp * ss.z + p.gamma_m1 * lr.z;
        const T first = p.gamma_inv_cp * lap_s + p.gamma_m1 * lap_lr;
        // cp * chi = cp * K / (rho cp) = K / rho
        const T heat = p.k_thermal * inv_rho * (first + glT_x * glT_x + glT_y * glT_y + glT_z * glT_z);
        rate[VTXBUF_ENTROPY] = -(ux.v * ss.x + uy.v * ss.y + uz.v * ss.z) + inv_pT * heating + heat;
    }
}

// ---- Streaming form of the same right-hand side ----
// The equations use each field only through a few combinations of its derivatives. A kernel
// that folds every field's derivatives into these aggregates as soon as they are formed keeps
// about 37 values live instead of all 8 x 10 derivatives:
//   lnrho, s : value, gradient, Laplacian
//   u        : value, gradients (ug[r][c] = d u_r / d x_c), Laplacian, grad(div u)
//   A        : Laplacian (induction), grad(div A) - lap A (current), curl A
// The arithmetic is that of mhd_rates up to the order of additions inside these sums.
struct MhdAgg {
    AcReal lr_v, lr_x, lr_y, lr_z, lr_lap;
    AcReal s_v, s_x, s_y, s_z, s_lap;
    AcReal uv[3], ug[3][3], ulap[3], gdu[3];
    AcReal alap[3], jj[3], B[3];
};

__device__ __forceinline__ void
mhd_rates_agg(const MhdAgg& g, const PhysParams& p, AcReal rate[NF])
{
    typedef AcReal T;
    const T divu = g.ug[0][0] + g.ug[1][1] + g.ug[2][2];
    rate[VTXBUF_LNRHO] = -(g.uv[0] * g.lr_x + g.uv[1] * g.lr_y + g.uv[2] * g.lr_z) - divu;

    const T S00 = T(2. / 3.) * g.ug[0][0] - T(1. / 3.) * (g.ug[1][1] + g.ug[2][2]);
    const T S11 = T(2. / 3.) * g.ug[1][1] - T(1. / 3.) * (g.ug[0][0] + g.ug[2][2]);
    const T S22 = T(2. / 3.) * g.ug[2][2] - T(1. / 3.) * (g.ug[0][0] + g.ug[1][1]);
    const T S01 = T(0.5) * (g.ug[0][1] + g.ug[1][0]);
    const T S02 = T(0.5) * (g.ug[0][2] + g.ug[2][0]);
    const T S12 = T(0.5) * (g.ug[1][2] + g.ug[2][1]);

    const T jx = p.inv_mu0 * g.jj[0], jy = p.inv_mu0 * g.jj[1], jz = p.inv_mu0 * g.jj[2];
    const T Bx = g.B[0], By = g.B[1], Bz = g.B[2];

    const T rho     = fast_exp<T>(g.lr_v);
    const T inv_rho = fast_rcp<T>(rho);
    const T E       = fast_exp<T>(p.gamma_inv_cp * g.s_v + p.gamma_m1 * (g.lr_v - p.lnrho0));
    const T cs2     = p.cs2_sound * E;
    {
        const T jxB[3] = {jy * Bz - jz * By, jz * Bx - jx * Bz, jx * By - jy * Bx};
        const T sg[3]  = {g.s_x, g.s_y, g.s_z};
        const T lg[3]  = {g.lr_x, g.lr_y, g.lr_z};
        const T Sgl[3] = {S00 * g.lr_x + S01 * g.lr_y + S02 * g.lr_z,
                          S01 * g.lr_x + S11 * g.lr_y + S12 * g.lr_z,
                          S02 * g.lr_x + S12 * g.lr_y + S22 * g.lr_z};
#pragma unroll
        for (int r = 0; r < 3; ++r)
            rate[VTXBUF_UUX + r] =
                -(g.ug[r][0] * g.uv[0] + g.ug[r][1] * g.uv[1] + g.ug[r][2] * g.uv[2]) -
                cs2 * (p.inv_cp * sg[r] + lg[r]) + inv_rho * jxB[r] +
                p.nu_visc * (g.ulap[r] + T(1. / 3.) * g.gdu[r] + T(2.) * Sgl[r]) + p.zeta * g.gdu[r];
    }
    rate[VTXBUF_AX] = (g.uv[1] * Bz - g.uv[2] * By) + p.eta * g.alap[0];
    rate[VTXBUF_AY] = (g.uv[2] * Bx - g.uv[0] * Bz) + p.eta * g.alap[1];
    rate[VTXBUF_AZ] = (g.uv[0] * By - g.uv[1] * Bx) + p.eta * g.alap[2];
    {
        const T inv_pT   = inv_rho * fast_rcp<T>(p.exp_lnT0 * E);
        const T contract = S00 * S00 + S11 * S11 + S22 * S22 +
                           T(2.) * (S01 * S01 + S02 * S02 + S12 * S12);
        const T heating = p.eta_mu0 * (jx * jx + jy * jy + jz * jz) +
                          T(2.) * rho * p.nu_visc * contract + p.zeta * rho * divu * divu;
        const T glT_x = p.gamma_inv_cp * g.s_x + p.gamma_m1 * g.lr_x;
        const T glT_y = p.gamma_inv_cp * g.s_y + p.gamma_m1 * g.lr_y;
        const T glT_z = p.gamma_inv_cp * g.s_z + p.gamma_m1 * g.lr_z;
        const T first = p.gamma_inv_cp * g.s_lap + p.gamma_m1 * g.lr_lap;
        const T heat  = p.k_thermal * inv_rho * (first + glT_x * glT_x + glT_y * glT_y + glT_z * glT_z);
        rate[VTXBUF_ENTROPY] =
            -(g.uv[0] * g.s_x + g.uv[1] * g.s_y + g.uv[2] * g.s_z) + inv_pT * heating + heat;
    }
}

// Low-storage RK3 update of one value (integration.cuh:14-37).
__device__ __forceinline__ AcReal
rk3_update(const RK3Coeffs& rk, AcReal prev, AcReal cur, AcReal rate)
{
    if (rk.step == 0)
        return cur + rk.beta_dt * rate;
    return cur + rk.beta * (rk.ab * (cur - prev) + rk.dt * rate);
}

} // namespace ac
