// Fused MHD solve + low-storage RK3 substep kernels for gfx950.
//
// Replaces the reference's generated solve<step> kernel (acc/mhd_solver/stencil_kernel.ac:693-734,
// launched by src/core/kernels/integration.cuh:217-249) and its autotuner
// (integration.cuh:128-215). The per-cell equations live in physics.h; this file decides how the
// stencil data reaches a thread:
//
//  variant 0 ("direct"): one thread per cell, every stencil point loaded from global memory.
//      Simple and used as the in-library cross-check of variant 1.
//
//  variant 1 ("z-march"): a block owns a TX x TY column of cells and marches it along z over
//      z_depth planes. Each new plane p of the six vector components (u, A) is staged in LDS with
//      a 3-cell x/y halo in a 4-slot ring, so the in-plane (x, y, xy) stencils of the output
//      plane z = p - 3 are read from LDS; lnrho and entropy, which need no cross derivatives,
//      get a one-slot LDS plane for the output plane only. z stencils come from a per-thread
//      register queue of the 7 planes z-3..z+3. The x-z and y-z cross derivatives are built by
//      scattering each plane's x (or y) differences into 7 pending per-output accumulators, so
//      no diagonal pencil is ever gathered: the contribution of plane p to output z = p - o is
//      sign(o) c_|o| (f(x+|o|, p) - f(x-|o|, p)).
//      Global loads of plane p+1 are issued before plane p is processed. Blocks are remapped so
//      that each XCD (round-robin dispatch over 8 XCDs) works on a contiguous range of tiles and
//      neighbouring tiles share the XCD's L2.
#include "physics.h"

namespace ac {

namespace {

constexpr int NXCD = 8;

__device__ __forceinline__ int
xcd_remap(int b, int nb)
{
    const int q = nb / NXCD, r = nb % NXCD;
    const int xcd = b % NXCD, local = b / NXCD;
    return xcd * q + (xcd < r ? xcd : r) + local;
}

// ------------------------------------------------------------------------------------------
// variant 0: direct global-memory stencils
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void
axis_direct(const AcReal* __restrict__ f, long idx, long s, const AcReal* d1, const AcReal* d2,
            AcReal c, AcReal& first, AcReal& second)
{
    const AcReal p1 = f[idx + s], m1 = f[idx - s];
    const AcReal p2 = f[idx + 2 * s], m2 = f[idx - 2 * s];
    const AcReal p3 = f[idx + 3 * s], m3 = f[idx - 3 * s];
    first  = d1[1] * (p1 - m1) + d1[2] * (p2 - m2) + d1[3] * (p3 - m3);
    second = d2[0] * c + d2[1] * (p1 + m1) + d2[2] * (p2 + m2) + d2[3] * (p3 + m3);
}

__device__ __forceinline__ AcReal
cross_direct(const AcReal* __restrict__ f, long idx, long sa, long sb, const AcReal* c)
{
    AcReal res = 0;
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
        const long pp = i * (sa + sb), pm = i * (sa - sb);
        res += c[i] * ((f[idx + pp] + f[idx - pp]) - (f[idx + pm] + f[idx - pm]));
    }
    return res;
}

template <bool XY, bool XZ, bool YZ>
__device__ __forceinline__ Fd<AcReal>
field_direct(const AcReal* __restrict__ f, long idx, long sy, long sz, const PhysParams& p)
{
    Fd<AcReal> d;
    d.v = f[idx];
    axis_direct(f, idx, 1, p.d1[0], p.d2[0], d.v, d.x, d.xx);
    axis_direct(f, idx, sy, p.d1[1], p.d2[1], d.v, d.y, d.yy);
    axis_direct(f, idx, sz, p.d1[2], p.d2[2], d.v, d.z, d.zz);
    d.xy = XY ? cross_direct(f, idx, 1, sy, p.cxy) : AcReal(0);
    d.xz = XZ ? cross_direct(f, idx, 1, sz, p.cxz) : AcReal(0);
    d.yz = YZ ? cross_direct(f, idx, sy, sz, p.cyz) : AcReal(0);
    return d;
}

template <int TX, int TY>
__global__ void __launch_bounds__(TX* TY)
solve_direct(const SubstepArgs a)
{
    const int i = a.start.x + (int)threadIdx.x + (int)blockIdx.x * TX;
    const int j = a.start.y + (int)threadIdx.y + (int)blockIdx.y * TY;
    const int k = a.start.z + (int)blockIdx.z;
    if (i >= a.end.x || j >= a.end.y)
        return;
    const long sy = a.mx, sz = a.mxy;
    const long idx = i + j * sy + k * sz;
    const PhysParams& p = a.phys;

    const Fd<AcReal> lr = field_direct<false, false, false>(a.in[VTXBUF_LNRHO], idx, sy, sz, p);
    const Fd<AcReal> ux = field_direct<true, true, false>(a.in[VTXBUF_UUX], idx, sy, sz, p);
    const Fd<AcReal> uy = field_direct<true, false, true>(a.in[VTXBUF_UUY], idx, sy, sz, p);
    const Fd<AcReal> uz = field_direct<false, true, true>(a.in[VTXBUF_UUZ], idx, sy, sz, p);
    const Fd<AcReal> ax = field_direct<true, true, false>(a.in[VTXBUF_AX], idx, sy, sz, p);
    const Fd<AcReal> ay = field_direct<true, false, true>(a.in[VTXBUF_AY], idx, sy, sz, p);
    const Fd<AcReal> az = field_direct<false, true, true>(a.in[VTXBUF_AZ], idx, sy, sz, p);
    const Fd<AcReal> ss = field_direct<false, false, false>(a.in[VTXBUF_ENTROPY], idx, sy, sz, p);

    AcReal rate[NF];
    mhd_rates(lr, ux, uy, uz, ax, ay, az, ss, p, rate);
    const AcReal cur[NF] = {lr.v, ux.v, uy.v, uz.v, ax.v, ay.v, az.v, ss.v};
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        const AcReal prev = a.rk.step > 0 ? a.out[q][idx] : AcReal(0);
        a.out[q][idx]     = rk3_update(a.rk, prev, cur[q], rate[q]);
    }
}

// ------------------------------------------------------------------------------------------
// variant 1: LDS tiles + z-march
// ------------------------------------------------------------------------------------------
// Ring index of the six vector components (handles 1..6).
constexpr int R_UX = 0, R_UY = 1, R_UZ = 2, R_AX = 3, R_AY = 4, R_AZ = 5;
// Components that need an x-z cross derivative, and a y-z one (ring indices).
constexpr int XZ_FIELDS[4] = {R_UX, R_UZ, R_AX, R_AZ};
constexpr int YZ_FIELDS[4] = {R_UY, R_UZ, R_AY, R_AZ};

template <int LX>
__device__ __forceinline__ void
inplane_lds(const AcReal* __restrict__ s, int c, const PhysParams& p, AcReal v, AcReal& dx,
            AcReal& dy, AcReal& dxx, AcReal& dyy)
{
    const AcReal xp1 = s[c + 1], xm1 = s[c - 1], xp2 = s[c + 2], xm2 = s[c - 2];
    const AcReal xp3 = s[c + 3], xm3 = s[c - 3];
    const AcReal yp1 = s[c + LX], ym1 = s[c - LX], yp2 = s[c + 2 * LX], ym2 = s[c - 2 * LX];
    const AcReal yp3 = s[c + 3 * LX], ym3 = s[c - 3 * LX];
    dx  = p.d1[0][1] * (xp1 - xm1) + p.d1[0][2] * (xp2 - xm2) + p.d1[0][3] * (xp3 - xm3);
    dxx = p.d2[0][0] * v + p.d2[0][1] * (xp1 + xm1) + p.d2[0][2] * (xp2 + xm2) +
          p.d2[0][3] * (xp3 + xm3);
    dy  = p.d1[1][1] * (yp1 - ym1) + p.d1[1][2] * (yp2 - ym2) + p.d1[1][3] * (yp3 - ym3);
    dyy = p.d2[1][0] * v + p.d2[1][1] * (yp1 + ym1) + p.d2[1][2] * (yp2 + ym2) +
          p.d2[1][3] * (yp3 + ym3);
}

template <int LX>
__device__ __forceinline__ AcReal
cross_xy_lds(const AcReal* __restrict__ s, int c, const AcReal* cxy)
{
    AcReal res = 0;
#pragma unroll
    for (int i = 1; i <= 3; ++i) {
        const int pp = i * (1 + LX), pm = i * (1 - LX);
        res += cxy[i] * ((s[c + pp] + s[c - pp]) - (s[c + pm] + s[c - pm]));
    }
    return res;
}

__device__ __forceinline__ void
axis_queue(const AcReal* q, const AcReal* d1, const AcReal* d2, AcReal& first, AcReal& second)
{
    first  = d1[1] * (q[4] - q[2]) + d1[2] * (q[5] - q[1]) + d1[3] * (q[6] - q[0]);
    second = d2[0] * q[3] + d2[1] * (q[4] + q[2]) + d2[2] * (q[5] + q[1]) + d2[3] * (q[6] + q[0]);
}

template <int TX, int TY>
__global__ void __launch_bounds__(TX* TY)
solve_zmarch(const SubstepArgs a, const int zdepth, const int ntx, const int nty)
{
    constexpr int NT = TX * TY;
    constexpr int LX = TX + 6, LY = TY + 6, LN = LX * LY;
    constexpr int R  = (LN + NT - 1) / NT;

    __shared__ AcReal ring[4][6][LN];
    __shared__ AcReal scal[2][LN];

    const PhysParams& P = a.phys;
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int b   = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int btx = b % ntx;
    const int bty = (b / ntx) % nty;
    const int btz = b / (ntx * nty);
    const int x0 = a.start.x + btx * TX, y0 = a.start.y + bty * TY;
    const int zs = a.start.z + btz * zdepth;
    const int ze = min(zs + zdepth, a.end.z);
    const int gx = x0 + tx, gy = y0 + ty;
    const bool active = gx < a.end.x && gy < a.end.y;
    // Own column; inactive threads of partial tiles read a valid column and write nothing.
    const long own = active ? (long)gx + (long)gy * a.mx : (long)x0 + (long)y0 * a.mx;

    // This thread's share of staging one (TX+6) x (TY+6) plane.
    int goff[R];
    bool lval[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e  = tid + r * NT;
        const int ex = e % LX, ey = e / LX;
        const int xx = x0 - 3 + ex, yy = y0 - 3 + ey;
        lval[r]      = e < LN && xx < a.mx && yy < a.my;
        goff[r]      = lval[r] ? xx + yy * a.mx : 0;
    }
    const int cofs = (ty + 3) * LX + (tx + 3);

    AcReal q[NF][7];        // z queue, q[f][6] = newest plane
    AcReal axz[4][7];       // x-z accumulators, [k] <-> output plane p - 3 + k
    AcReal ayz[4][7];       // y-z accumulators
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < 7; ++k)
            q[f][k] = 0;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int k = 0; k < 7; ++k)
            axz[f][k] = ayz[f][k] = 0;

    AcReal pv[6][R], ps[2][R], pc[2];
    const int pbeg = zs - 3, pend = ze + 3;

    auto issue_loads = [&](int p) {
        const long base = (long)p * a.mxy;
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int r = 0; r < R; ++r)
                pv[f][r] = lval[r] ? a.in[1 + f][base + goff[r]] : AcReal(0);
        const int zo = p - 3;
        if (zo >= zs) {
            const long base2 = (long)zo * a.mxy;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                ps[0][r] = lval[r] ? a.in[VTXBUF_LNRHO][base2 + goff[r]] : AcReal(0);
                ps[1][r] = lval[r] ? a.in[VTXBUF_ENTROPY][base2 + goff[r]] : AcReal(0);
            }
        }
        pc[0] = a.in[VTXBUF_LNRHO][base + own];
        pc[1] = a.in[VTXBUF_ENTROPY][base + own];
    };

    issue_loads(pbeg);
    for (int p = pbeg; p < pend; ++p) {
        const int slot = p & 3;
        const int z    = p - 3;
        const bool out_plane = z >= zs;

        __syncthreads(); // every thread is done with slot `slot` (plane p - 4) and with `scal`
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r * NT + tid < LN)
                    ring[slot][f][r * NT + tid] = pv[f][r];
        if (out_plane) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r * NT + tid < LN) {
                    scal[0][r * NT + tid] = ps[0][r];
                    scal[1][r * NT + tid] = ps[1][r];
                }
        }
        const AcReal c_lr = pc[0], c_ss = pc[1];
        __syncthreads();

        if (p + 1 < pend)
            issue_loads(p + 1);

        // Previous RK stage of the output cell.
        AcReal prev[NF];
        const long oidx = own + (long)z * a.mxy;
        if (out_plane && active && a.rk.step > 0) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                prev[f] = a.out[f][oidx];
        }
        else {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                prev[f] = 0;
        }

        // ---- ingest plane p: push the z queues, scatter cross contributions ----
        const AcReal* __restrict__ rs = &ring[slot][0][0];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                q[f][k] = q[f][k + 1];
        q[VTXBUF_LNRHO][6] = c_lr;
        q[VTXBUF_ENTROPY][6] = c_ss;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            q[1 + f][6] = rs[f * LN + cofs];

#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const AcReal* __restrict__ s = rs + XZ_FIELDS[n] * LN;
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                const AcReal d = P.cxz[i] * (s[cofs + i] - s[cofs - i]);
                axz[n][3 - i] += d;
                axz[n][3 + i] -= d;
            }
        }
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const AcReal* __restrict__ s = rs + YZ_FIELDS[n] * LN;
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                const AcReal d = P.cyz[i] * (s[cofs + i * LX] - s[cofs - i * LX]);
                ayz[n][3 - i] += d;
                ayz[n][3 + i] -= d;
            }
        }

        // ---- output plane z = p - 3 ----
        if (out_plane && active) {
            const AcReal* __restrict__ ro = &ring[z & 3][0][0];
            Fd<AcReal> d[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                d[f].v = q[f][3];
                axis_queue(q[f], P.d1[2], P.d2[2], d[f].z, d[f].zz);
                d[f].xy = d[f].xz = d[f].yz = 0;
            }
            inplane_lds<LX>(&scal[0][0], cofs, P, d[VTXBUF_LNRHO].v, d[VTXBUF_LNRHO].x,
                            d[VTXBUF_LNRHO].y, d[VTXBUF_LNRHO].xx, d[VTXBUF_LNRHO].yy);
            inplane_lds<LX>(&scal[1][0], cofs, P, d[VTXBUF_ENTROPY].v, d[VTXBUF_ENTROPY].x,
                            d[VTXBUF_ENTROPY].y, d[VTXBUF_ENTROPY].xx, d[VTXBUF_ENTROPY].yy);
#pragma unroll
            for (int f = 0; f < 6; ++f)
                inplane_lds<LX>(ro + f * LN, cofs, P, d[1 + f].v, d[1 + f].x, d[1 + f].y,
                                d[1 + f].xx, d[1 + f].yy);
            d[VTXBUF_UUX].xy = cross_xy_lds<LX>(ro + R_UX * LN, cofs, P.cxy);
            d[VTXBUF_UUY].xy = cross_xy_lds<LX>(ro + R_UY * LN, cofs, P.cxy);
            d[VTXBUF_AX].xy  = cross_xy_lds<LX>(ro + R_AX * LN, cofs, P.cxy);
            d[VTXBUF_AY].xy  = cross_xy_lds<LX>(ro + R_AY * LN, cofs, P.cxy);
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                d[1 + XZ_FIELDS[n]].xz = axz[n][0];
                d[1 + YZ_FIELDS[n]].yz = ayz[n][0];
            }
            AcReal rate[NF];
            mhd_rates(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], P, rate);
#pragma unroll
            for (int f = 0; f < NF; ++f)
                a.out[f][oidx] = rk3_update(a.rk, prev[f], d[f].v, rate[f]);
        }

        // ---- retire output z, open output p + 4 ----
#pragma unroll
        for (int n = 0; n < 4; ++n) {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                axz[n][k] = axz[n][k + 1];
                ayz[n][k] = ayz[n][k + 1];
            }
            axz[n][6] = 0;
            ayz[n][6] = 0;
        }
    }
}

// ------------------------------------------------------------------------------------------
// variant 2: the z-march of variant 1 with the plane loop unrolled by the queue length
// ------------------------------------------------------------------------------------------
// * Vector-field staging loads are unconditional: an element outside the grid (partial tiles)
//   reads the plane's first cell instead, and the staging buffers are padded to whole rounds of
//   the block, so no exec-mask branch surrounds those loads or their LDS stores. Values staged
//   outside the grid are only read by threads of partial tiles, which compute but do not store.
// * The z loop advances 7 planes per trip with the phase PH = (p - pbeg) mod 7 known at compile
//   time: plane p' lives in queue slot (p' - pbeg) mod 7 and output z' accumulates in slot
//   (z' - pbeg) mod 7, so pushing a plane and retiring an output move no registers.
template <int TX, int TY> struct ZMarch2 {
    static constexpr int NT = TX * TY;
    static constexpr int LX = TX + 6, LY = TY + 6, LN = LX * LY;
    static constexpr int R  = (LN + NT - 1) / NT;
    static constexpr int LNP = R * NT; // padded plane: every thread stores every round

    struct Shared {
        AcReal ring[4][6][LNP];
        AcReal scal[2][LNP];
    };

    struct State {
        AcReal q[NF][7];
        AcReal axz[4][7];
        AcReal ayz[4][7];
        AcReal pv[6][R];
        AcReal ps[2][R];
        AcReal pc[2];
    };

    const SubstepArgs& a;
    Shared& sh;
    int tid, cofs, zs, ze, pend;
    long own;
    bool active;
    int goff[R];

    __device__ __forceinline__ ZMarch2(const SubstepArgs& args, Shared& shared) : a(args), sh(shared) {}

    __device__ __forceinline__ void issue_loads(State& st, int p) const
    {
        const long base = (long)p * a.mxy;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const AcReal* __restrict__ src = a.in[1 + f] + base;
#pragma unroll
            for (int r = 0; r < R; ++r)
                st.pv[f][r] = src[goff[r]];
        }
        const int zo = p - 3;
        if (zo >= zs) {
            const long base2 = (long)zo * a.mxy;
            const AcReal* __restrict__ l = a.in[VTXBUF_LNRHO] + base2;
            const AcReal* __restrict__ e = a.in[VTXBUF_ENTROPY] + base2;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                st.ps[0][r] = l[goff[r]];
                st.ps[1][r] = e[goff[r]];
            }
        }
        st.pc[0] = a.in[VTXBUF_LNRHO][base + own];
        st.pc[1] = a.in[VTXBUF_ENTROPY][base + own];
    }

    template <int PH> __device__ __forceinline__ void step(State& st, int p) const
    {
        const PhysParams& P = a.phys;
        const int slot       = p & 3;
        const int z          = p - 3;
        const bool out_plane = z >= zs;

        __syncthreads(); // all threads are done with slot `slot` (plane p - 4) and with `scal`
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int r = 0; r < R; ++r)
                sh.ring[slot][f][r * NT + tid] = st.pv[f][r];
        if (out_plane) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                sh.scal[0][r * NT + tid] = st.ps[0][r];
                sh.scal[1][r * NT + tid] = st.ps[1][r];
            }
        }
        const AcReal c_lr = st.pc[0], c_ss = st.pc[1];
        __syncthreads();

        if (p + 1 < pend)
            issue_loads(st, p + 1);

        AcReal prev[NF];
        const long oidx = own + (long)z * a.mxy;
        const bool need_prev = out_plane && a.rk.step > 0;
#pragma unroll
        for (int f = 0; f < NF; ++f)
            prev[f] = need_prev ? a.out[f][oidx] : AcReal(0);

        // ---- ingest plane p ----
        const AcReal* __restrict__ rs = &sh.ring[slot][0][0];
        st.q[VTXBUF_LNRHO][PH]   = c_lr;
        st.q[VTXBUF_ENTROPY][PH] = c_ss;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            st.q[1 + f][PH] = rs[f * LNP + cofs];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const AcReal* __restrict__ sx = rs + XZ_FIELDS[n] * LNP;
            const AcReal* __restrict__ sy = rs + YZ_FIELDS[n] * LNP;
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                const AcReal dx = P.cxz[i] * (sx[cofs + i] - sx[cofs - i]);
                st.axz[n][(PH - i + 7) % 7] += dx;
                st.axz[n][(PH + i) % 7] -= dx;
                const AcReal dy = P.cyz[i] * (sy[cofs + i * LX] - sy[cofs - i * LX]);
                st.ayz[n][(PH - i + 7) % 7] += dy;
                st.ayz[n][(PH + i) % 7] -= dy;
            }
        }

        // ---- output plane z = p - 3 ----
        constexpr int OS = (PH + 4) % 7; // accumulator slot of output p - 3
        if (out_plane) {
            const AcReal* __restrict__ ro = &sh.ring[z & 3][0][0];
            Fd<AcReal> d[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                AcReal zq[7];
#pragma unroll
                for (int k = 0; k < 7; ++k)
                    zq[k] = st.q[f][(PH + 1 + k) % 7];
                d[f].v = zq[3];
                axis_queue(zq, P.d1[2], P.d2[2], d[f].z, d[f].zz);
                d[f].xy = d[f].xz = d[f].yz = 0;
            }
            inplane_lds<LX>(&sh.scal[0][0], cofs, P, d[VTXBUF_LNRHO].v, d[VTXBUF_LNRHO].x,
                            d[VTXBUF_LNRHO].y, d[VTXBUF_LNRHO].xx, d[VTXBUF_LNRHO].yy);
            inplane_lds<LX>(&sh.scal[1][0], cofs, P, d[VTXBUF_ENTROPY].v, d[VTXBUF_ENTROPY].x,
                            d[VTXBUF_ENTROPY].y, d[VTXBUF_ENTROPY].xx, d[VTXBUF_ENTROPY].yy);
#pragma unroll
            for (int f = 0; f < 6; ++f)
                inplane_lds<LX>(ro + f * LNP, cofs, P, d[1 + f].v, d[1 + f].x, d[1 + f].y,
                                d[1 + f].xx, d[1 + f].yy);
            d[VTXBUF_UUX].xy = cross_xy_lds<LX>(ro + R_UX * LNP, cofs, P.cxy);
            d[VTXBUF_UUY].xy = cross_xy_lds<LX>(ro + R_UY * LNP, cofs, P.cxy);
            d[VTXBUF_AX].xy  = cross_xy_lds<LX>(ro + R_AX * LNP, cofs, P.cxy);
            d[VTXBUF_AY].xy  = cross_xy_lds<LX>(ro + R_AY * LNP, cofs, P.cxy);
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                d[1 + XZ_FIELDS[n]].xz = st.axz[n][OS];
                d[1 + YZ_FIELDS[n]].yz = st.ayz[n][OS];
            }
            AcReal rate[NF];
            mhd_rates(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], P, rate);
            if (active) {
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    a.out[f][oidx] = rk3_update(a.rk, prev[f], d[f].v, rate[f]);
            }
        }
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            st.axz[n][OS] = 0;
            st.ayz[n][OS] = 0;
        }
    }
};

template <int TX, int TY>
__global__ void __launch_bounds__(TX* TY)
solve_zmarch2(const SubstepArgs a, const int zdepth, const int ntx, const int nty)
{
    typedef ZMarch2<TX, TY> K;
    __shared__ typename K::Shared sh;
    K k(a, sh);
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int b   = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int btx = b % ntx;
    const int bty = (b / ntx) % nty;
    const int btz = b / (ntx * nty);
    const int x0 = a.start.x + btx * TX, y0 = a.start.y + bty * TY;
    k.tid        = tid;
    k.zs         = a.start.z + btz * zdepth;
    k.ze         = min(k.zs + zdepth, a.end.z);
    const int gx = x0 + tx, gy = y0 + ty;
    k.active     = gx < a.end.x && gy < a.end.y;
    k.own        = k.active ? (long)gx + (long)gy * a.mx : (long)x0 + (long)y0 * a.mx;
#pragma unroll
    for (int r = 0; r < K::R; ++r) {
        const int e  = tid + r * K::NT;
        const int ex = e % K::LX, ey = e / K::LX;
        const int xx = x0 - 3 + ex, yy = y0 - 3 + ey;
        k.goff[r]    = (e < K::LN && xx < a.mx && yy < a.my) ? xx + yy * a.mx : 0;
    }
    k.cofs = (ty + 3) * K::LX + (tx + 3);

    typename K::State st;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int j = 0; j < 7; ++j)
            st.q[f][j] = 0;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int j = 0; j < 7; ++j)
            st.axz[f][j] = st.ayz[f][j] = 0;

    const int pbeg = k.zs - 3;
    k.pend         = k.ze + 3;
    k.issue_loads(st, pbeg);
    for (int p = pbeg; p < k.pend; p += 7) {
        k.template step<0>(st, p);
        if (p + 1 < k.pend) k.template step<1>(st, p + 1);
        if (p + 2 < k.pend) k.template step<2>(st, p + 2);
        if (p + 3 < k.pend) k.template step<3>(st, p + 3);
        if (p + 4 < k.pend) k.template step<4>(st, p + 4);
        if (p + 5 < k.pend) k.template step<5>(st, p + 5);
        if (p + 6 < k.pend) k.template step<6>(st, p + 6);
    }
}

// ------------------------------------------------------------------------------------------
// configuration table
// ------------------------------------------------------------------------------------------
struct ConfigEntry {
    int variant, tx, ty;
};

constexpr ConfigEntry kConfigs[] = {
    {0, 64, 4},
    {0, 32, 4},
    {0, 64, 2},
    {1, 64, 4},
    {1, 32, 4},
    {1, 64, 2},
    {1, 32, 8},
#if AC_DOUBLE_PRECISION == 0
    {1, 128, 2},
    {1, 64, 8},
#endif
    {2, 64, 4},
    {2, 32, 4},
    {2, 32, 8},
    {2, 64, 2},
    {2, 16, 4},
    {2, 8, 8},
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

// Grid of z-march blocks: tiles of TX x TY columns, each marching z_depth planes (0 or more
// than the box: the whole box depth).
typedef void (*ZMarchKernel)(const SubstepArgs, const int, const int, const int);

template <int TX, int TY>
hipError_t
launch_zmarch_grid(ZMarchKernel kernel, const SubstepArgs& a, int zdepth, hipStream_t stream)
{
    const int nx = a.end.x - a.start.x, ny = a.end.y - a.start.y, nz = a.end.z - a.start.z;
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    if (zdepth <= 0 || zdepth > nz)
        zdepth = nz;
    const int ntz = (nz + zdepth - 1) / zdepth;
    const long nblocks = (long)ntx * nty * ntz;
    if (nblocks > 0x7fffffffL)
        return hipErrorInvalidConfiguration;
    kernel<<<dim3((unsigned)nblocks), dim3(TX * TY), 0, stream>>>(a, zdepth, ntx, nty);
    return hipGetLastError();
}

template <int TX, int TY>
hipError_t
launch_direct(const SubstepArgs& a, hipStream_t stream)
{
    const int nx = a.end.x - a.start.x, ny = a.end.y - a.start.y, nz = a.end.z - a.start.z;
    const dim3 grid((nx + TX - 1) / TX, (ny + TY - 1) / TY, nz);
    solve_direct<TX, TY><<<grid, dim3(TX, TY), 0, stream>>>(a);
    return hipGetLastError();
}

} // namespace

bool
kernel_config_supported(const AcKernelConfig& c)
{
    for (int i = 0; i < kNumConfigs; ++i)
        if (kConfigs[i].variant == c.variant && kConfigs[i].tx == c.tile_x &&
            kConfigs[i].ty == c.tile_y)
            return c.z_depth >= 0;
    return false;
}

int
kernel_config_count()
{
    return kNumConfigs;
}

AcKernelConfig
kernel_config_at(int index)
{
    AcKernelConfig c{};
    if (index >= 0 && index < kNumConfigs) {
        c.variant = kConfigs[index].variant;
        c.tile_x  = kConfigs[index].tx;
        c.tile_y  = kConfigs[index].ty;
        c.z_depth = 0;
    }
    return c;
}

AcKernelConfig
default_kernel_config(int nx, int ny, int nz)
{
    (void)nx;
    (void)ny;
    AcKernelConfig c;
    c.variant = 1;
    c.tile_x  = 64;
    c.tile_y  = 4;
    c.z_depth = nz < 64 ? nz : 64;
    return c;
}

hipError_t
launch_substep(const AcKernelConfig& c, const SubstepArgs& a, hipStream_t stream)
{
    if (a.end.x <= a.start.x || a.end.y <= a.start.y || a.end.z <= a.start.z)
        return hipSuccess;
#define AC_DIRECT(TX, TY)                                                                          \
    if (c.variant == 0 && c.tile_x == TX && c.tile_y == TY)                                        \
        return launch_direct<TX, TY>(a, stream);
#define AC_ZMARCH(TX, TY)                                                                          \
    if (c.variant == 1 && c.tile_x == TX && c.tile_y == TY)                                        \
        return launch_zmarch_grid<TX, TY>(solve_zmarch<TX, TY>, a, c.z_depth, stream);
    AC_DIRECT(64, 4)
    AC_DIRECT(32, 4)
    AC_DIRECT(64, 2)
    AC_ZMARCH(64, 4)
    AC_ZMARCH(32, 4)
    AC_ZMARCH(64, 2)
    AC_ZMARCH(32, 8)
#if AC_DOUBLE_PRECISION == 0
    AC_ZMARCH(128, 2)
    AC_ZMARCH(64, 8)
#endif
#define AC_ZMARCH2(TX, TY)                                                                         \
    if (c.variant == 2 && c.tile_x == TX && c.tile_y == TY)                                        \
        return launch_zmarch_grid<TX, TY>(solve_zmarch2<TX, TY>, a, c.z_depth, stream);
    AC_ZMARCH2(64, 4)
    AC_ZMARCH2(32, 4)
    AC_ZMARCH2(32, 8)
    AC_ZMARCH2(64, 2)
    AC_ZMARCH2(16, 4)
    AC_ZMARCH2(8, 8)
#undef AC_DIRECT
#undef AC_ZMARCH
#undef AC_ZMARCH2
    return hipErrorInvalidConfiguration;
}

} // namespace ac
