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
//
//  variant 2: variant 1 with the plane loop unrolled by the 7-plane queue length, so queue and
//      accumulator slots are compile-time registers and no register moves rotate them.
//
//  variants 3, 4 ("lean z-march"): all eight fields in one 4-slot LDS ring, a 3-register z
//      queue per field, six combined cross-derivative sums instead of eight, a branch-free plane
//      loop and a right-hand side formed field by field. About 25 % fewer issued instructions
//      per cell than variant 2; variant 4 is the same kernel compiled for two waves per SIMD.
//  variants 7-12: the lean z-march with a 5-slot ring (7: one barrier per plane), or loading
//      the previous stage before the next plane and storing each output plane at the start of
//      the next step (8: 4 slots, the FP64 default; 9: 5 slots), and its two-waves-per-SIMD
//      builds on 16 x 16 tiles (10: deferred stores; 11: previous stage first; 12: both).
//  variant 13: variant 12 with an 8-phase plane loop (compile-time ring slots); 16: variant 12
//      with a 5-slot ring.
//  variants 17, 19, 22: variants 16, 12 and 10 with pinned LDS windows (every stencil read of a
//      field's plane through one base register); 19 is the FP32 default.
//  variant 23: variant 17 loading each plane straight into its ring slot (direct-to-LDS loads).
//  variants 5, 6: the lean z-march with a cell-major LDS ring (FP32; kept for comparison).
//  The defaults also have builds that read the periodic images of the ghost cells (WRAP), used by
//  the whole-step entry points of a periodic box where they are the faster form.
#include "physics.h"

#include <algorithm>
#include <atomic>
#include <cmath>

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

template <int LX, typename Ptr>
__device__ __forceinline__ void
inplane_lds(Ptr s, int c, const PhysParams& p, AcReal v, AcReal& dx, AcReal& dy, AcReal& dxx,
            AcReal& dyy)
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

template <int LX, typename Ptr>
__device__ __forceinline__ AcReal
cross_xy_lds(Ptr s, int c, const AcReal* cxy)
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
// variants 3 and 4: lean z-march with all eight fields in one LDS ring
// ------------------------------------------------------------------------------------------
// The data flow of variant 2 with the per-thread state cut down and the plane loop free of
// branches (variant 3 is compiled for one wave per SIMD, variant 4 for two):
// * Every field's planes z..z+3 sit in a 4-slot LDS ring (a haloed TX x TY tile each) and only
//   z-3..z-1 in a 3-register queue, so the z stencils read four of their seven values from LDS.
// * The x-z and y-z cross derivatives are only used in six sums (grad div u and
//   grad div A - lap A), so six pending accumulators per output plane are kept instead of eight:
//     X0 = uz_xz, X1 = uz_yz, X2 = ux_xz + uy_yz, X3 = az_xz, X4 = az_yz, X5 = ax_xz + ay_yz.
//   The first contribution to an output plane (from plane z - 3) is assigned, so no slot is
//   ever cleared.
// * The march is a prologue of six planes that only fill the queues, then whole rounds of 7
//   planes that each retire one output plane. Planes past the grid are read clamped to its
//   last plane; the outputs that depend on them (z >= ze) are not stored. Stores are the only
//   guarded code: by z < ze (uniform) and by the thread's column lying in the box.
// * Global loads and stores use a uniform per-plane base and a 32-bit per-thread byte offset.
// * The right-hand side is formed field by field into the aggregates of MhdAgg, which keeps the
//   live set small.
// NSLOT = 4: the ring holds planes p-3..p and a step stages plane p into the slot just freed,
// between two barriers. NSLOT = 5: plane p+1 is staged into a fifth slot while planes p-3..p are
// read, so one barrier per step suffices (the staging registers then hold plane p+1 across the
// step instead of plane p).
// NQ: period of the cross-sum slots and of the unrolled plane loop. 7 is the minimum (outputs
// p-3..p+3 are pending); with NQ = 8 and NSLOT = 4 the ring slot of every plane a step touches
// is a compile-time function of the phase, so no LDS address is recomputed per step.
// WIN: every stencil read of one field in one ring plane goes through one LDS base register
// (the window's first cell, pinned in a VGPR), so that all its reads pair into ds_read2 with
// immediate offsets; otherwise the compiler forms a new base (one VALU add) for most pairs.
// WRAP: the tile halos and the planes above and below the box are read at their periodic images
// inside the interior, so no ghost cell is read and the ghosts need not be filled first (the
// values read are the same as after a periodic fill).
// DMA (5-slot ring only): plane p + 1 is loaded straight into its ring slot with direct-to-LDS
// loads (global_load_lds_dword: lane l of a wave writes the 4 bytes it loads to the wave's LDS
// base + 4 l), issued at the start of step p into the slot that plane p - 4 left; no staging
// registers and no LDS writes. Each wave waits for its own loads (vmcnt 0) before the step's
// barrier, so after it the whole plane is in LDS. In FP64 a round of 64 doubles is two such
// loads, lane l fetching the low or high half (l & 1) of double l / 2 of its half-round.
template <int TX, int TY, bool COMP = false, int NSLOT = 4, bool DEFER = false,
          bool PREV_FIRST = false, int NQ = 7, bool WIN = false, bool WRAP = false,
          bool DMA = false>
struct ZMarchLean {
    static_assert(NQ == 7 || (NQ == 8 && NSLOT == 4), "8-phase loop needs the 4-slot ring");
    static_assert(!DMA || NSLOT == 5, "direct-to-LDS staging needs the 5-slot ring");
    static constexpr int NT  = TX * TY;
    static_assert(NT % 64 == 0, "whole waves");
    static constexpr int LX  = TX + 6, LY = TY + 6, LN = LX * LY;
    static constexpr int R   = (LN + NT - 1) / NT;
    static constexpr int LNP = R * NT;
    // loads per field and round: the direct-to-LDS load moves 4 bytes per lane
    static constexpr int NH = DMA ? (int)(sizeof(AcReal) / 4) : 1;

    struct Shared {
        AcReal ring[NSLOT][NF][LNP];
    };

    struct State {
        AcReal qv[NF][3]; // planes z-3..z-1 of every field
        AcReal x[6][NQ];  // cross-derivative sums, slot (z - pbeg) mod NQ
        AcReal pv[DMA ? 1 : NF][DMA ? 1 : R]; // staging registers of the next plane
        AcReal res[NF];   // updated values of plane res_z, stored during the next step
        int res_z;        // -1: nothing pending
    };

    // Stores the output plane held in `res`. With DEFER it is issued at the start of the next
    // step, before that step's loads: memory operations complete in issue order, so the wait for
    // the next plane's loads does not also wait for the stores' acknowledgements.
    __device__ __forceinline__ void flush(State& st) const
    {
        if (st.res_z >= 0 && active) {
            const unsigned long oz = (unsigned long)st.res_z * a.mxy;
            const unsigned ow      = pin(own);
#pragma unroll
            for (int f = 0; f < NF; ++f)
                st_global(a.out[f] + oz, ow, st.res[f]);
        }
        st.res_z = -1;
    }

    const SubstepArgs& a;
    Shared& sh;
    int tid, cofs, ze, zmax;
    int wave; // wave index within the block (uniform)
    unsigned own; // byte offset of the thread's column within a plane
    bool active;
    unsigned goff[R][NH]; // byte offsets within a plane of what this lane loads per round

    __device__ __forceinline__ ZMarchLean(const SubstepArgs& args, Shared& shared) : a(args), sh(shared) {}

    static __device__ __forceinline__ unsigned pin(unsigned off)
    {
        asm volatile("" : "+v"(off));
        return off;
    }

    // Field memory is global memory. The batched solve reads its pointers from a device array,
    // whose address space the compiler cannot tell: without the cast its loads and stores are
    // flat operations, which also count in the LDS counter, so every wait for a ring read would
    // also wait for the next plane's loads in flight.
    typedef __attribute__((address_space(1))) AcReal GReal;
    typedef __attribute__((address_space(1))) const char GByte;

    static __device__ __forceinline__ AcReal ld(const AcReal* base, unsigned byte_off)
    {
        return *(const GReal*)((GByte*)base + byte_off);
    }

    static __device__ __forceinline__ void st_global(AcReal* base, unsigned byte_off, AcReal v)
    {
        *(GReal*)((GByte*)base + byte_off) = v;
    }

    // Ring slot of plane q (q >= 0).
    static __device__ __forceinline__ int ring_slot(int q) { return NSLOT == 4 ? (q & 3) : q % NSLOT; }

    __device__ __forceinline__ void stage(const State& st, int slot) const
    {
        static_assert(!DMA, "direct-to-LDS loads need no staging");
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < R; ++r)
                sh.ring[slot][f][r * NT + tid] = st.pv[f][r];
    }

    __device__ __forceinline__ void issue_loads(State& st, int p) const
    {
        int q = min(p, zmax);
        if (WRAP) { // periodic image of a ghost plane (planes run 0 .. zmax = nz + 5)
            const int nz = zmax - 5;
            q = q < NGHOST ? q + nz : q >= nz + NGHOST ? q - nz : q;
        }
        const long pl = (long)q * a.mxy;
        unsigned go[R][NH];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int h = 0; h < NH; ++h)
                go[r][h] = pin(goff[r][h]);
        if constexpr (DMA) {
            const int slot = ring_slot(p);
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const char* src = (const char*)(a.in[f] + pl);
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int h = 0; h < NH; ++h) {
                        AcReal* dst = &sh.ring[slot][f][r * NT + wave * 64 + h * (64 / NH)];
                        __builtin_amdgcn_global_load_lds(
                            src + go[r][h], (__attribute__((address_space(3))) void*)dst, 4, 0, 0);
                    }
            }
        }
        else {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const AcReal* src = a.in[f] + pl;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    st.pv[f][r] = ld(src, go[r][0]);
            }
        }
    }

    // Wait for this wave's outstanding memory operations (vmcnt 0; expcnt and lgkmcnt free).
    static __device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

    __device__ __forceinline__ void push_queue(State& st, const AcReal* ro) const
    {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            st.qv[f][0] = st.qv[f][1];
            st.qv[f][1] = st.qv[f][2];
            st.qv[f][2] = ro[f * LNP + cofs];
        }
    }

    // Value and z derivatives of field f at plane z: qv (z-3..z-1) and ring slots z..z+3.
    // w: field f's cell in plane z; c1, c2, c3: this thread's cell of planes z+1..z+3 (field 0)
    template <typename PW, typename PC>
    __device__ __forceinline__ void zderivs(const State& st, int f, PW w, PC c1, PC c2, PC c3,
                                            AcReal& v, AcReal& dz, AcReal& dzz) const
    {
        const int o        = f * LNP;
        const AcReal zq[7] = {st.qv[f][0], st.qv[f][1], st.qv[f][2], w[0], c1[o], c2[o], c3[o]};
        v                  = zq[3];
        axis_queue(zq, a.phys.d1[2], a.phys.d2[2], dz, dzz);
    }

    typedef const AcReal __attribute__((address_space(3))) LdsReal;
    static constexpr int W0 = 3 * LX + 3; // window corner (-3, -3) to its centre

    // LDS address laundered through a VGPR: the compiler cannot rebase reads made through it.
    static __device__ __forceinline__ const LdsReal* lds_pin(const AcReal* q)
    {
        unsigned addr = (unsigned)(size_t)(const LdsReal*)q;
        asm volatile("" : "+v"(addr));
        return (const LdsReal*)(size_t)addr;
    }

    // This thread's cell of field plane pf (pf: first cell of a field in a ring plane).
    __device__ __forceinline__ auto centre(const AcReal* pf) const
    {
        if constexpr (WIN)
            return lds_pin(pf + cofs - W0) + W0;
        else
            return pf + cofs;
    }
    // This thread's cell of field 0 of ring plane rp (other fields at multiples of LNP).
    __device__ __forceinline__ auto plane_cell(const AcReal* rp) const
    {
        if constexpr (WIN)
            return lds_pin(rp + cofs);
        else
            return rp + cofs;
    }

    template <int PH, bool OUT> __device__ __forceinline__ void step(State& st, int p) const
    {
        const PhysParams& P = a.phys;
        // ring slot of plane p (and below, of planes z = p - 3, z + 1, z + 2)
        const int slot = NQ == 8 ? (PH & 3) : ring_slot(p);

        if constexpr (NSLOT == 4) {
            __syncthreads(); // every thread is done with ring slot `slot` (plane p - 4)
            stage(st, slot);
            __syncthreads();
        }
        else {
            // plane p (staged during step p - 1) is visible, and every thread is done with
            // plane p - 4, whose slot now receives plane p + 1
            if constexpr (DMA)
                wait_vm(); // this wave's part of plane p has landed
            __syncthreads();
            if constexpr (!DMA)
                stage(st, ring_slot(p + 1));
        }

        if (DEFER)
            flush(st);
        // With PREV_FIRST the previous stage's values are loaded before the next plane: memory
        // operations complete in issue order, so waiting for them at the update then does not
        // also wait for the next plane's loads (measured faster in FP64, slower in FP32).
        // next plane to load: into registers, p + 1 (4 slots) or p + 2 (5 slots, staged at the
        // next step); straight into its ring slot, p + 1
        constexpr int AHEAD = (NSLOT == 4 || DMA) ? 1 : 2;
        if (!PREV_FIRST)
            issue_loads(st, p + AHEAD);
        const int z             = p - 3;
        const unsigned long oz  = (unsigned long)min(z, ze - 1) * a.mxy;
        const unsigned ow       = pin(own);
        // In FP32 the output cells are read on every substep, also on substep 0, which has no
        // previous stage: a 16-wide tile row covers part of a cache line, and a store that misses
        // the cache waits for the line to be filled; with the line already loaded the store hits
        // (measured: substep 0 at 512^3 5.45 -> 4.6 ms). In FP64 the extra reads cost more than
        // they save (9.7 -> 10.9 ms).
        constexpr bool TOUCH = sizeof(AcReal) == 4;
        AcReal prev[NF];
        if (OUT && (TOUCH || a.rk.step > 0)) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                prev[f] = ld(a.out[f] + oz, ow);
        }
        // compensated integration: the rounding-error terms of the output cell, loaded with the
        // previous stage so that their latency is hidden behind the step's arithmetic
        AcReal cin[COMP ? NF : 1], cprev[COMP ? NF : 1];
        if (COMP && OUT) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                cin[f]   = ld(a.cin[f] + oz, ow);
                cprev[f] = a.rk.step > 0 ? ld(a.cout[f] + oz, ow) : AcReal(0);
            }
        }
        if (PREV_FIRST)
            issue_loads(st, p + AHEAD);

        // ---- ingest plane p: x-z and y-z differences into the pending sums ----
        const AcReal* __restrict__ rs = &sh.ring[slot][0][0];
#pragma unroll
        for (int i = 1; i <= 3; ++i) {
            const int lo = (PH - i + NQ) % NQ, hi = (PH + i) % NQ; // outputs p - i and p + i
            const AcReal cx = P.cxz[i], cy = P.cyz[i];
            const auto ux = centre(rs + VTXBUF_UUX * LNP);
            const auto uy = centre(rs + VTXBUF_UUY * LNP);
            const auto uz = centre(rs + VTXBUF_UUZ * LNP);
            const auto ax = centre(rs + VTXBUF_AX * LNP);
            const auto ay = centre(rs + VTXBUF_AY * LNP);
            const auto az = centre(rs + VTXBUF_AZ * LNP);
            const AcReal d0 = cx * (uz[i] - uz[-i]);
            const AcReal d1 = cy * (uz[i * LX] - uz[-i * LX]);
            const AcReal d2 = cx * (ux[i] - ux[-i]) + cy * (uy[i * LX] - uy[-i * LX]);
            const AcReal d3 = cx * (az[i] - az[-i]);
            const AcReal d4 = cy * (az[i * LX] - az[-i * LX]);
            const AcReal d5 = cx * (ax[i] - ax[-i]) + cy * (ay[i * LX] - ay[-i * LX]);
            st.x[0][lo] += d0;
            st.x[1][lo] += d1;
            st.x[2][lo] += d2;
            st.x[3][lo] += d3;
            st.x[4][lo] += d4;
            st.x[5][lo] += d5;
            if (i == 3) { // first contribution to output p + 3
                st.x[0][hi] = -d0;
                st.x[1][hi] = -d1;
                st.x[2][hi] = -d2;
                st.x[3][hi] = -d3;
                st.x[4][hi] = -d4;
                st.x[5][hi] = -d5;
            }
            else {
                st.x[0][hi] -= d0;
                st.x[1][hi] -= d1;
                st.x[2][hi] -= d2;
                st.x[3][hi] -= d3;
                st.x[4][hi] -= d4;
                st.x[5][hi] -= d5;
            }
        }
        const AcReal* __restrict__ ro =
            &sh.ring[NQ == 8 ? ((PH + 1) & 3) : ring_slot(p + NSLOT - 3)][0][0]; // plane z
        if (!OUT) {
            push_queue(st, ro);
            return;
        }

        // ---- output plane z = p - 3 ----
        constexpr int OS = (PH + NQ - 3) % NQ; // slot of output z = p - 3
        const AcReal* __restrict__ r1 =
            &sh.ring[NQ == 8 ? ((PH + 2) & 3) : ring_slot(p + NSLOT - 2)][0][0]; // plane z + 1
        const AcReal* __restrict__ r2 =
            &sh.ring[NQ == 8 ? ((PH + 3) & 3) : ring_slot(p + NSLOT - 1)][0][0]; // plane z + 2
        MhdAgg g;
        AcReal v, dx, dy, dz, dxx, dyy, dzz;
        const auto c1 = plane_cell(r1), c2 = plane_cell(r2), c3 = plane_cell(rs);

        const auto wlnrho = centre(ro + VTXBUF_LNRHO * LNP);
        zderivs(st, VTXBUF_LNRHO, wlnrho, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(wlnrho, 0, P, v, dx, dy, dxx, dyy);
        g.lr_v = v, g.lr_x = dx, g.lr_y = dy, g.lr_z = dz, g.lr_lap = dxx + dyy + dzz;

        const auto wentropy = centre(ro + VTXBUF_ENTROPY * LNP);
        zderivs(st, VTXBUF_ENTROPY, wentropy, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(wentropy, 0, P, v, dx, dy, dxx, dyy);
        g.s_v = v, g.s_x = dx, g.s_y = dy, g.s_z = dz, g.s_lap = dxx + dyy + dzz;

        // grad(div u)_r = sum_c d_r d_c u_c; the x-z / y-z parts come from x[0..2]
        const auto wuux = centre(ro + VTXBUF_UUX * LNP);
        zderivs(st, VTXBUF_UUX, wuux, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(wuux, 0, P, v, dx, dy, dxx, dyy);
        const AcReal ux_xy = cross_xy_lds<LX>(wuux, 0, P.cxy);
        g.uv[0] = v, g.ug[0][0] = dx, g.ug[0][1] = dy, g.ug[0][2] = dz;
        g.ulap[0] = dxx + dyy + dzz;
        g.gdu[0]  = dxx + st.x[0][OS];
        g.gdu[1]  = ux_xy + st.x[1][OS];

        const auto wuuy = centre(ro + VTXBUF_UUY * LNP);
        zderivs(st, VTXBUF_UUY, wuuy, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(wuuy, 0, P, v, dx, dy, dxx, dyy);
        const AcReal uy_xy = cross_xy_lds<LX>(wuuy, 0, P.cxy);
        g.uv[1] = v, g.ug[1][0] = dx, g.ug[1][1] = dy, g.ug[1][2] = dz;
        g.ulap[1] = dxx + dyy + dzz;
        g.gdu[0] += uy_xy;
        g.gdu[1] += dyy;

        const auto wuuz = centre(ro + VTXBUF_UUZ * LNP);
        zderivs(st, VTXBUF_UUZ, wuuz, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(wuuz, 0, P, v, dx, dy, dxx, dyy);
        g.uv[2] = v, g.ug[2][0] = dx, g.ug[2][1] = dy, g.ug[2][2] = dz;
        g.ulap[2] = dxx + dyy + dzz;
        g.gdu[2]  = st.x[2][OS] + dzz;

        // grad(div A) - lap A; the own-axis second derivative cancels
        const auto wax = centre(ro + VTXBUF_AX * LNP);
        zderivs(st, VTXBUF_AX, wax, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(wax, 0, P, v, dx, dy, dxx, dyy);
        const AcReal ax_xy = cross_xy_lds<LX>(wax, 0, P.cxy);
        g.alap[0] = dxx + dyy + dzz;
        g.jj[0]   = st.x[3][OS] - (dyy + dzz);
        g.jj[1]   = ax_xy + st.x[4][OS];
        g.B[1]    = dz;
        g.B[2]    = -dy;

        const auto way = centre(ro + VTXBUF_AY * LNP);
        zderivs(st, VTXBUF_AY, way, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(way, 0, P, v, dx, dy, dxx, dyy);
        const AcReal ay_xy = cross_xy_lds<LX>(way, 0, P.cxy);
        g.alap[1] = dxx + dyy + dzz;
        g.jj[0] += ay_xy;
        g.jj[1] -= dxx + dzz;
        g.B[0] = -dz;
        g.B[2] += dx;

        const auto waz = centre(ro + VTXBUF_AZ * LNP);
        zderivs(st, VTXBUF_AZ, waz, c1, c2, c3, v, dz, dzz);
        inplane_lds<LX>(waz, 0, P, v, dx, dy, dxx, dyy);
        g.alap[2] = dxx + dyy + dzz;
        g.jj[2]   = st.x[5][OS] - (dxx + dyy);
        g.B[0] += dy;
        g.B[1] -= dx;

        AcReal rate[NF];
        mhd_rates_agg(g, P, rate);
        push_queue(st, ro); // qv[f][2] is now plane z
        if (COMP) {
            store_compensated(st, rate, prev, cin, cprev, oz, ow, z);
            return;
        }
        if (a.rk.step == 0) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                st.res[f] = st.qv[f][2] + a.rk.beta_dt * rate[f];
                if (TOUCH)
                    asm volatile("" ::"v"(prev[f])); // the line-touching read is not dead
            }
        }
        else {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const AcReal cur = st.qv[f][2];
                st.res[f] = cur + a.rk.beta * (a.rk.ab * (cur - prev[f]) + a.rk.dt * rate[f]);
            }
        }
        st.res_z = z < ze ? z : -1;
        if (!DEFER)
            flush(st);
    }

    // Compensated update: the state is the pair (u, c) with u + c carrying the rounding error
    // of every previous update (u in `in`/`out`, c in `cin`/`cout`). The increment is formed
    // from the pairs and added with an error-free two-sum (Knuth's TwoSum), whose error term
    // becomes the new compensation:
    //   s = 0: d = beta_dt F + c_in
    //   s > 0: d = beta (ab ((u_in - u_prev) + (c_in - c_prev)) + dt F) + c_in
    //   u_out + c_out = u_in + d exactly.
    // The stencils read u only.
    __device__ __forceinline__ void store_compensated(const State& st, const AcReal* rate,
                                                      const AcReal* prev, const AcReal* cin,
                                                      const AcReal* cprev, unsigned long oz,
                                                      unsigned ow, int z) const
    {
        if (!(z < ze && active))
            return;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const AcReal u  = st.qv[f][2];
            const AcReal ci = cin[f];
            AcReal d;
            if (a.rk.step == 0) {
                d = a.rk.beta_dt * rate[f] + ci;
            }
            else {
                const AcReal cp = cprev[f];
                d = a.rk.beta * (a.rk.ab * ((u - prev[f]) + (ci - cp)) + a.rk.dt * rate[f]) + ci;
            }
            const AcReal sum = u + d;
            const AcReal bb  = sum - u;
            const AcReal err = (u - (sum - bb)) + (d - bb);
            st_global(a.out[f] + oz, ow, sum);
            st_global(a.cout[f] + oz, ow, err);
        }
    }
};

#if AC_DOUBLE_PRECISION == 0
// ------------------------------------------------------------------------------------------
// variants 5 and 6 (FP32): the lean z-march with a cell-major LDS ring
// ------------------------------------------------------------------------------------------
// The ring stores each cell's eight values as two float4 records (in two rings), so one
// ds_read_b128 brings four fields of a stencil point and all fields of a plane share two LDS
// base addresses per ring slot. This halves the LDS instructions of variant 3 and removes its
// per-field LDS address arithmetic. Group A = (u_x, A_x, u_y, A_y): the x-y cross derivatives
// need all of it, the x-z sums its low half and the y-z sums its high half; group
// B = (u_z, A_z, lnrho, s): the cross sums need its low half.
struct AosCell {
    float4 A; // ux, ax, uy, ay
    float4 B; // uz, az, lnrho, s
};
// LDS position of each VertexBufferHandle (lnrho, ux, uy, uz, ax, ay, az, s)
constexpr int AOS_POS[NF] = {6, 0, 2, 4, 1, 3, 5, 7};

__device__ __forceinline__ float
cell_get(const AosCell& c, int k)
{
    switch (k) {
    case 0: return c.A.x;
    case 1: return c.A.y;
    case 2: return c.A.z;
    case 3: return c.A.w;
    case 4: return c.B.x;
    case 5: return c.B.y;
    case 6: return c.B.z;
    default: return c.B.w;
    }
}

template <int TX, int TY> struct ZMarchAos {
    static constexpr int NT  = TX * TY;
    static constexpr int LX  = TX + 6, LY = TY + 6, LN = LX * LY;
    static constexpr int R   = (LN + NT - 1) / NT;
    static constexpr int LNP = R * NT;

    // Two rings of float4 (group A, group B): consecutive lanes read consecutive 16-byte
    // records, which keeps ds_read_b128 free of bank conflicts.
    struct Shared {
        float4 ra[4][LNP];
        float4 rb[4][LNP];
    };

    struct CellPtr {
        const float4* __restrict__ a;
        const float4* __restrict__ b;
        __device__ __forceinline__ AosCell operator[](int d) const
        {
            AosCell c;
            c.A = a[d];
            c.B = b[d];
            return c;
        }
    };

    __device__ __forceinline__ CellPtr cells(int slot) const
    {
        CellPtr c;
        c.a = &sh.ra[slot][cofs];
        c.b = &sh.rb[slot][cofs];
        return c;
    }

    struct State {
        float qv[NF][3]; // planes z-3..z-1, in LDS order
        float x[6][7];   // cross-derivative sums, slot (z - pbeg) mod 7
        float pv[NF][R]; // staging registers of the next plane (VertexBufferHandle order)
    };

    const SubstepArgs& a;
    Shared& sh;
    int tid, cofs, ze, zmax;
    unsigned own;
    bool active;
    unsigned goff[R];

    __device__ __forceinline__ ZMarchAos(const SubstepArgs& args, Shared& shared) : a(args), sh(shared) {}

    static __device__ __forceinline__ unsigned pin(unsigned off)
    {
        asm volatile("" : "+v"(off));
        return off;
    }

    static __device__ __forceinline__ float ld(const float* base, unsigned byte_off)
    {
        return *(const float*)((const char*)base + byte_off);
    }

    __device__ __forceinline__ void issue_loads(State& st, int p) const
    {
        const long pl = (long)min(p, zmax) * a.mxy;
        unsigned go[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            go[r] = pin(goff[r]);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const float* src = a.in[f] + pl;
#pragma unroll
            for (int r = 0; r < R; ++r)
                st.pv[f][r] = ld(src, go[r]);
        }
    }

    __device__ __forceinline__ void push_queue(State& st, const AosCell& c) const
    {
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            st.qv[k][0] = st.qv[k][1];
            st.qv[k][1] = st.qv[k][2];
            st.qv[k][2] = cell_get(c, k);
        }
    }

    template <int PH, bool OUT> __device__ __forceinline__ void step(State& st, int p) const
    {
        const PhysParams& P = a.phys;
        const int slot      = p & 3;

        __syncthreads(); // every thread is done with ring slot `slot` (plane p - 4)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sh.ra[slot][r * NT + tid] = make_float4(st.pv[VTXBUF_UUX][r], st.pv[VTXBUF_AX][r],
                                                    st.pv[VTXBUF_UUY][r], st.pv[VTXBUF_AY][r]);
            sh.rb[slot][r * NT + tid] = make_float4(st.pv[VTXBUF_UUZ][r], st.pv[VTXBUF_AZ][r],
                                                    st.pv[VTXBUF_LNRHO][r],
                                                    st.pv[VTXBUF_ENTROPY][r]);
        }
        __syncthreads();
        issue_loads(st, p + 1);

        const int z            = p - 3;
        const unsigned long oz = (unsigned long)min(z, ze - 1) * a.mxy;
        const unsigned ow      = pin(own);
        float prev[NF];
        if (OUT && a.rk.step > 0) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                prev[f] = ld(a.out[f] + oz, ow);
        }

        // ---- ingest plane p: x-z and y-z differences into the pending sums ----
        const CellPtr c3 = cells(slot);
#pragma unroll
        for (int i = 1; i <= 3; ++i) {
            const int lo = (PH - i + 7) % 7, hi = (PH + i) % 7; // outputs p - i and p + i
            const float cx = P.cxz[i], cy = P.cyz[i];
            const float2 xpA = *(const float2*)&c3.a[i];            // ux, ax at x + i
            const float2 xmA = *(const float2*)&c3.a[-i];
            const float2 xpB = *(const float2*)&c3.b[i];            // uz, az at x + i
            const float2 xmB = *(const float2*)&c3.b[-i];
            const float2 ypA = *((const float2*)&c3.a[i * LX] + 1);  // uy, ay at y + i
            const float2 ymA = *((const float2*)&c3.a[-i * LX] + 1);
            const float2 ypB = *(const float2*)&c3.b[i * LX];       // uz, az at y + i
            const float2 ymB = *(const float2*)&c3.b[-i * LX];
            const float d0 = cx * (xpB.x - xmB.x);
            const float d1 = cy * (ypB.x - ymB.x);
            const float d2 = cx * (xpA.x - xmA.x) + cy * (ypA.x - ymA.x);
            const float d3 = cx * (xpB.y - xmB.y);
            const float d4 = cy * (ypB.y - ymB.y);
            const float d5 = cx * (xpA.y - xmA.y) + cy * (ypA.y - ymA.y);
            st.x[0][lo] += d0;
            st.x[1][lo] += d1;
            st.x[2][lo] += d2;
            st.x[3][lo] += d3;
            st.x[4][lo] += d4;
            st.x[5][lo] += d5;
            if (i == 3) { // first contribution to output p + 3
                st.x[0][hi] = -d0;
                st.x[1][hi] = -d1;
                st.x[2][hi] = -d2;
                st.x[3][hi] = -d3;
                st.x[4][hi] = -d4;
                st.x[5][hi] = -d5;
            }
            else {
                st.x[0][hi] -= d0;
                st.x[1][hi] -= d1;
                st.x[2][hi] -= d2;
                st.x[3][hi] -= d3;
                st.x[4][hi] -= d4;
                st.x[5][hi] -= d5;
            }
        }
        const CellPtr c0 = cells((p + 1) & 3); // plane z
        if (!OUT) {
            push_queue(st, c0[0]);
            return;
        }

        // ---- output plane z = p - 3 ----
        constexpr int OS              = (PH + 4) % 7;
        const CellPtr c1 = cells((p + 2) & 3); // plane z + 1
        const CellPtr c2 = cells((p + 3) & 3); // plane z + 2
        float v[NF], dx[NF], dy[NF], dxx[NF], dyy[NF], dz[NF], dzz[NF];
        {
            const AosCell m0 = c0[0], m1 = c1[0], m2 = c2[0], m3 = c3[0];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const float zq[7] = {st.qv[k][0], st.qv[k][1], st.qv[k][2], cell_get(m0, k),
                                     cell_get(m1, k), cell_get(m2, k), cell_get(m3, k)};
                v[k] = zq[3];
                axis_queue(zq, P.d1[2], P.d2[2], dz[k], dzz[k]);
            }
        }
        // in-plane first and second derivatives of all eight fields, stencil point by point
        {
            AosCell xp[3], xm[3], yp[3], ym[3];
#pragma unroll
            for (int i = 1; i <= 3; ++i) {
                xp[i - 1] = c0[i];
                xm[i - 1] = c0[-i];
                yp[i - 1] = c0[i * LX];
                ym[i - 1] = c0[-i * LX];
            }
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const float p1 = cell_get(xp[0], k), m1 = cell_get(xm[0], k);
                const float p2 = cell_get(xp[1], k), m2 = cell_get(xm[1], k);
                const float p3 = cell_get(xp[2], k), m3 = cell_get(xm[2], k);
                dx[k]  = P.d1[0][1] * (p1 - m1) + P.d1[0][2] * (p2 - m2) + P.d1[0][3] * (p3 - m3);
                dxx[k] = P.d2[0][0] * v[k] + P.d2[0][1] * (p1 + m1) + P.d2[0][2] * (p2 + m2) +
                         P.d2[0][3] * (p3 + m3);
                const float q1 = cell_get(yp[0], k), n1 = cell_get(ym[0], k);
                const float q2 = cell_get(yp[1], k), n2 = cell_get(ym[1], k);
                const float q3 = cell_get(yp[2], k), n3 = cell_get(ym[2], k);
                dy[k]  = P.d1[1][1] * (q1 - n1) + P.d1[1][2] * (q2 - n2) + P.d1[1][3] * (q3 - n3);
                dyy[k] = P.d2[1][0] * v[k] + P.d2[1][1] * (q1 + n1) + P.d2[1][2] * (q2 + n2) +
                         P.d2[1][3] * (q3 + n3);
            }
        }
        // x-y cross derivatives of ux, ax, uy, ay (group A)
        float4 xy = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int i = 1; i <= 3; ++i) {
            const int pp = i * (1 + LX), pm = i * (1 - LX);
            const float4 a1 = c0.a[pp], a2 = c0.a[-pp], a3 = c0.a[pm], a4 = c0.a[-pm];
            const float cc = P.cxy[i];
            xy.x += cc * ((a1.x + a2.x) - (a3.x + a4.x));
            xy.y += cc * ((a1.y + a2.y) - (a3.y + a4.y));
            xy.z += cc * ((a1.z + a2.z) - (a3.z + a4.z));
            xy.w += cc * ((a1.w + a2.w) - (a3.w + a4.w));
        }
        // LDS order: 0 ux, 1 ax, 2 uy, 3 ay, 4 uz, 5 az, 6 lnrho, 7 s
        MhdAgg g;
        g.lr_v = v[6], g.lr_x = dx[6], g.lr_y = dy[6], g.lr_z = dz[6];
        g.lr_lap = dxx[6] + dyy[6] + dzz[6];
        g.s_v = v[7], g.s_x = dx[7], g.s_y = dy[7], g.s_z = dz[7];
        g.s_lap = dxx[7] + dyy[7] + dzz[7];
        g.uv[0] = v[0], g.ug[0][0] = dx[0], g.ug[0][1] = dy[0], g.ug[0][2] = dz[0];
        g.uv[1] = v[2], g.ug[1][0] = dx[2], g.ug[1][1] = dy[2], g.ug[1][2] = dz[2];
        g.uv[2] = v[4], g.ug[2][0] = dx[4], g.ug[2][1] = dy[4], g.ug[2][2] = dz[4];
        g.ulap[0] = dxx[0] + dyy[0] + dzz[0];
        g.ulap[1] = dxx[2] + dyy[2] + dzz[2];
        g.ulap[2] = dxx[4] + dyy[4] + dzz[4];
        g.gdu[0] = dxx[0] + st.x[0][OS] + xy.z; // ux_xx + uz_xz + uy_xy
        g.gdu[1] = xy.x + st.x[1][OS] + dyy[2]; // ux_xy + uz_yz + uy_yy
        g.gdu[2] = st.x[2][OS] + dzz[4];        // ux_xz + uy_yz + uz_zz
        g.alap[0] = dxx[1] + dyy[1] + dzz[1];
        g.alap[1] = dxx[3] + dyy[3] + dzz[3];
        g.alap[2] = dxx[5] + dyy[5] + dzz[5];
        g.jj[0] = (st.x[3][OS] - (dyy[1] + dzz[1])) + xy.w; // az_xz - (ax_yy + ax_zz) + ay_xy
        g.jj[1] = (xy.y + st.x[4][OS]) - (dxx[3] + dzz[3]);  // ax_xy + az_yz - (ay_xx + ay_zz)
        g.jj[2] = st.x[5][OS] - (dxx[5] + dyy[5]);          // ax_xz + ay_yz - (az_xx + az_yy)
        g.B[0] = -dz[3] + dy[5];                            // az_y - ay_z
        g.B[1] = dz[1] - dx[5];                             // ax_z - az_x
        g.B[2] = -dy[1] + dx[3];                            // ay_x - ax_y

        float rate[NF];
        mhd_rates_agg(g, P, rate);
        push_queue(st, c0[0]); // qv[k][2] is now plane z
        if (z < ze && active) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const float cur = st.qv[AOS_POS[f]][2];
                float outv;
                if (a.rk.step == 0)
                    outv = cur + a.rk.beta_dt * rate[f];
                else
                    outv = cur + a.rk.beta * (a.rk.ab * (cur - prev[f]) + a.rk.dt * rate[f]);
                *(float*)((char*)(a.out[f] + oz) + ow) = outv;
            }
        }
    }
};

template <int TX, int TY, int W>
__global__ void __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(W)))
solve_zaos(const SubstepArgs a, const int zdepth, const int ntx, const int nty)
{
    typedef ZMarchAos<TX, TY> K;
    __shared__ typename K::Shared sh;
    const int b   = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int btx = b % ntx;
    const int bty = (b / ntx) % nty;
    const int btz = b / (ntx * nty);
    K k(a, sh);
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int x0 = a.start.x + btx * TX, y0 = a.start.y + bty * TY;
    const int zs = a.start.z + btz * zdepth;
    k.tid        = tid;
    k.ze         = min(zs + zdepth, a.end.z);
    k.zmax       = a.mz - 1;
    const int gx = x0 + tx, gy = y0 + ty;
    k.active     = gx < a.end.x && gy < a.end.y;
    k.own        = (unsigned)(k.active ? gx + gy * a.mx : x0 + y0 * a.mx) * sizeof(float);
#pragma unroll
    for (int r = 0; r < K::R; ++r) {
        const int e  = tid + r * K::NT;
        const int ex = e % K::LX, ey = e / K::LX;
        const int xx = x0 - 3 + ex, yy = y0 - 3 + ey;
        k.goff[r] = (unsigned)((e < K::LN && xx < a.mx && yy < a.my) ? xx + yy * a.mx : 0) *
                    sizeof(float);
    }
    k.cofs = (ty + 3) * K::LX + (tx + 3);

    typename K::State st;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            st.qv[f][j] = 0;
#pragma unroll
    for (int n = 0; n < 6; ++n)
#pragma unroll
        for (int j = 0; j < 7; ++j)
            st.x[n][j] = 0;

    const int pbeg = zs - 3;
    k.issue_loads(st, pbeg);
    k.template step<0, false>(st, pbeg);
    k.template step<1, false>(st, pbeg + 1);
    k.template step<2, false>(st, pbeg + 2);
    k.template step<3, false>(st, pbeg + 3);
    k.template step<4, false>(st, pbeg + 4);
    k.template step<5, false>(st, pbeg + 5);
    for (int p = pbeg + 6; p < k.ze + 3; p += 7) {
        k.template step<6, true>(st, p);
        k.template step<0, true>(st, p + 1);
        k.template step<1, true>(st, p + 2);
        k.template step<2, true>(st, p + 3);
        k.template step<3, true>(st, p + 4);
        k.template step<4, true>(st, p + 5);
        k.template step<5, true>(st, p + 6);
    }
}
#endif // AC_DOUBLE_PRECISION == 0

// Tile row of thread `tid` in the lean z-march. LDS serves a 32-bit access 32 lanes at a time
// over 32 banks. On a 16-wide tile the natural order puts rows y and y + 1 in one 32-lane group;
// their LDS rows start 22 words (the padded row pitch) apart, so 6 of the 16 banks overlap and
// every stencil read takes two passes. Here the two 16-lane halves of each group hold rows 8
// apart instead: 8 * 22 = 176 = 16 mod 32, so each group covers all 32 banks once, for every
// stencil offset. Row(tid) = 2 (tid / 64) + (tid / 32 mod 2) + 8 (tid / 16 mod 2).
template <int TX, int TY>
__device__ __forceinline__ int
lean_tile_row(const int tid)
{
    if (TX == 16 && TY == 16)
        return 2 * (tid >> 6) + ((tid >> 5) & 1) + 8 * ((tid >> 4) & 1);
    return tid / TX;
}

// Tile of logical block b (after the XCD remap) for the lean z-march. Boxes up to 32 tiles wide
// keep plain row order; wider ones are ordered in strips 16 tiles wide, row after row inside a
// strip. The blocks that run at the same time on one XCD (about 64 consecutive logical indices)
// then cover four rows of a strip rather than one row of the box, and the y halos they share
// are read from that XCD's L2 instead of HBM: at 1024^3 (64 tiles of 16 per row) a substep
// takes 37.5 ms instead of 43.0 ms; at 512^3 row order and strips measured alike.
__device__ __forceinline__ void
strip_tile(const int b, const int ntx, const int nty, int& btx, int& bty, int& btz)
{
    const int STRIP = ntx <= 32 ? ntx : 16;
    const int per_z = ntx * nty;
    btz             = b / per_z;
    const int l     = b - btz * per_z;
    const int s     = l / (STRIP * nty);          // strip (the last one may be narrower)
    const int loc   = l - s * STRIP * nty;
    const int w     = min(STRIP, ntx - s * STRIP); // strip width in tiles
    bty             = loc / w;
    btx             = s * STRIP + (loc - bty * w);
}

// One z-march column: tile (btx, bty) of box `a`, output planes [zs, min(zs + zdepth, end.z)).
template <int TX, int TY, bool COMP = false, int NSLOT = 4, bool DEFER = false,
          bool PREV_FIRST = false, int NQ = 7, bool WIN = false, bool WRAP = false,
          bool DMA = false>
__device__ __forceinline__ void
zlean_column(const SubstepArgs& a,
             typename ZMarchLean<TX, TY, COMP, NSLOT, DEFER, PREV_FIRST, NQ, WIN, WRAP, DMA>::Shared& sh,
             const int btx, const int bty, const int zs, const int zdepth)
{
    typedef ZMarchLean<TX, TY, COMP, NSLOT, DEFER, PREV_FIRST, NQ, WIN, WRAP, DMA> K;
    K k(a, sh);
    const int tid = threadIdx.x;
    k.wave        = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tx = tid % TX, ty = lean_tile_row<TX, TY>(tid);
    const int x0 = a.start.x + btx * TX, y0 = a.start.y + bty * TY;
    k.tid        = tid;
    k.ze         = min(zs + zdepth, a.end.z);
    k.zmax       = a.mz - 1;
    const int gx = x0 + tx, gy = y0 + ty;
    k.active     = gx < a.end.x && gy < a.end.y;
    k.own        = (unsigned)(k.active ? gx + gy * a.mx : x0 + y0 * a.mx) * sizeof(AcReal);
#pragma unroll
    for (int r = 0; r < K::R; ++r)
#pragma unroll
        for (int h = 0; h < K::NH; ++h) {
            // the tile cell (haloed LX x LY tile, row-major) this lane loads in round r: its own,
            // or with two loads per double the one whose half it fetches
            const int e  = K::NH == 1 ? tid + r * K::NT
                                      : r * K::NT + (tid & ~63) + h * 32 + ((tid & 63) >> 1);
            const int ex = e % K::LX, ey = e / K::LX;
            int xx = x0 - 3 + ex, yy = y0 - 3 + ey;
            const bool in_box = e < K::LN && xx < a.mx && yy < a.my;
            if (WRAP) { // periodic images of the x and y ghost columns
                const int nx = a.mx - 2 * NGHOST, ny = a.my - 2 * NGHOST;
                xx = xx < NGHOST ? xx + nx : xx >= nx + NGHOST ? xx - nx : xx;
                yy = yy < NGHOST ? yy + ny : yy >= ny + NGHOST ? yy - ny : yy;
            }
            k.goff[r][h] = (unsigned)(in_box ? xx + yy * a.mx : 0) * sizeof(AcReal) +
                           (K::NH == 1 ? 0u : (unsigned)(tid & 1) * 4u);
        }
    k.cofs = (ty + 3) * K::LX + (tx + 3);

    typename K::State st;
    st.res_z = -1;
    // Queue entries of planes below the march and sums of outputs below zs never reach a
    // stored output; zero them so no undefined value enters the arithmetic.
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            st.qv[f][j] = 0;
#pragma unroll
    for (int n = 0; n < 6; ++n)
#pragma unroll
        for (int j = 0; j < NQ; ++j)
            st.x[n][j] = 0;

    const int pbeg = zs - 3;
    if constexpr (DMA) {
        __syncthreads(); // the previous column of this block is done with the ring
        k.issue_loads(st, pbeg);
    }
    else {
        k.issue_loads(st, pbeg);
        if (NSLOT == 5) {
            __syncthreads(); // the previous column of this block is done with the ring
            k.stage(st, K::ring_slot(pbeg));
            k.issue_loads(st, pbeg + 1);
        }
    }
    k.template step<0, false>(st, pbeg);
    k.template step<1, false>(st, pbeg + 1);
    k.template step<2, false>(st, pbeg + 2);
    k.template step<3, false>(st, pbeg + 3);
    k.template step<4, false>(st, pbeg + 4);
    k.template step<5, false>(st, pbeg + 5);
    if (NQ == 7) {
        // Output planes z = p - 3 for p in [pbeg + 6, ze + 3): whole rounds of the 7 phases,
        // then the phases of a last, partial round up to the last output plane, so that a column
        // whose depth is not a multiple of 7 computes no plane past its end (the branches are
        // uniform; the partial round is a second copy of the phases' code, outside the loop).
        const int pend = k.ze + 3;
        int p          = pbeg + 6;
        for (; p + 7 <= pend; p += 7) {
            k.template step<6, true>(st, p);
            k.template step<0, true>(st, p + 1);
            k.template step<1, true>(st, p + 2);
            k.template step<2, true>(st, p + 3);
            k.template step<3, true>(st, p + 4);
            k.template step<4, true>(st, p + 5);
            k.template step<5, true>(st, p + 6);
        }
        if (p < pend) {
            k.template step<6, true>(st, p);
            if (p + 1 < pend) {
                k.template step<0, true>(st, p + 1);
                if (p + 2 < pend) {
                    k.template step<1, true>(st, p + 2);
                    if (p + 3 < pend) {
                        k.template step<2, true>(st, p + 3);
                        if (p + 4 < pend) {
                            k.template step<3, true>(st, p + 4);
                            if (p + 5 < pend)
                                k.template step<4, true>(st, p + 5);
                        }
                    }
                }
            }
        }
    }
    else {
        for (int p = pbeg + 6; p < k.ze + 3; p += 8) {
            k.template step<6, true>(st, p);
            k.template step<7 % NQ, true>(st, p + 1);
            k.template step<0, true>(st, p + 2);
            k.template step<1, true>(st, p + 3);
            k.template step<2, true>(st, p + 4);
            k.template step<3, true>(st, p + 5);
            k.template step<4, true>(st, p + 6);
            k.template step<5, true>(st, p + 7);
        }
    }
    if (DEFER)
        k.flush(st); // the last output plane
    if constexpr (DMA)
        K::wait_vm(); // no load into the ring is left in flight when the block moves on or ends
}

template <int TX, int TY, int W, bool COMP = false, int NSLOT = 4, bool DEFER = false,
          bool PREV_FIRST = false, int NQ = 7, bool WIN = false, bool WRAP = false,
          bool DMA = false>
__global__ void __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(W)))
solve_zlean(const SubstepArgs a, const int zdepth, const int ntx, const int nty)
{
    __shared__
        typename ZMarchLean<TX, TY, COMP, NSLOT, DEFER, PREV_FIRST, NQ, WIN, WRAP, DMA>::Shared sh;
    int btx, bty, btz;
    strip_tile(xcd_remap((int)blockIdx.x, (int)gridDim.x), ntx, nty, btx, bty, btz);
    zlean_column<TX, TY, COMP, NSLOT, DEFER, PREV_FIRST, NQ, WIN, WRAP, DMA>(
        a, sh, btx, bty, a.start.z + btz * zdepth, zdepth);
}

// Persistent form for ensembles of independent boxes: a grid sized to the resident block slots
// walks a (box, tile column, z range) work list; each box brings its own arguments (pointers,
// sizes, physical parameters), read from global memory.
// The 16 x 16 FP32 build has the single-box default's options (4-slot ring, previous stage
// first, deferred stores, pinned LDS windows, two waves per SIMD): 7.93 ms per 96-box step
// against 8.87 ms for a 5-slot ring without pinned windows, which spills 20 registers to
// scratch (interleaved on one box, work items dealt to XCDs in contiguous ranges).
template <int TX, int TY, int W = 1, int NSLOT = 4, bool DEFER = false, bool PREV_FIRST = false,
          bool WIN = false>
__global__ void __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(W)))
solve_zlean_batch(const SubstepArgs* __restrict__ boxes, const BatchItem* __restrict__ items,
                  const int nitems)
{
    // periodic images read directly: the ensemble step fills no ghosts between its substeps
    __shared__
        typename ZMarchLean<TX, TY, false, NSLOT, DEFER, PREV_FIRST, 7, WIN, true>::Shared sh;
    // Round r takes items [r G, (r + 1) G) (G = gridDim.x), each XCD a contiguous range of them
    // (xcd_remap): consecutive items are mostly neighbouring tile columns of one box, whose halo
    // lines are then shared in one L2.
    const int slot = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    for (int it = slot; it < nitems; it += gridDim.x) {
        const BatchItem w = items[it];
        zlean_column<TX, TY, false, NSLOT, DEFER, PREV_FIRST, 7, WIN, true>(
            boxes[w.box], sh, w.btx, w.bty, w.zs, w.zdepth);
    }
}

#if AC_DOUBLE_PRECISION == 0
#define AC_BATCH16 solve_zlean_batch<16, 16, 2, 4, true, true, true>
#else
#define AC_BATCH16 solve_zlean_batch<16, 16>
#endif

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
    {3, 16, 16},
#if AC_DOUBLE_PRECISION == 0
    {3, 32, 4},
    {3, 16, 8},
    {4, 16, 16},
    {5, 32, 4},
    {5, 16, 8},
    {5, 16, 16},
    {6, 16, 16},
    {6, 32, 4},
    {7, 32, 4},
    {7, 16, 8},
    {8, 32, 4},
    {9, 32, 4},
    {10, 16, 16},
    {11, 16, 16},
    {12, 16, 16},
    {13, 16, 16},
    {16, 16, 16},
    {17, 16, 16},
    {19, 16, 16},
    {22, 16, 16},
    {23, 16, 16},
#endif
    {8, 16, 16},
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

bool
kernel_config_reads_images(const AcKernelConfig& c)
{
    // the defaults (FP32 variant 19, FP64 variant 8), compiled a second time with WRAP
#if AC_DOUBLE_PRECISION == 1
    return c.variant == 8 && c.tile_x == 16 && c.tile_y == 16 && c.z_depth >= 0;
#else
    return c.variant == 19 && c.tile_x == 16 && c.tile_y == 16 && c.z_depth >= 0;
#endif
}

bool
default_reads_images(int nx, int ny, int nz)
{
    // The WRAP builds save the three fills of a step, but their launches slow down with the box
    // (the far-side halo and plane reads: at 1024^3 FP32 a launch takes 51 ms against 36 ms
    // with fills). Per step, fill + launch against images (ms, interleaved on one box):
    //   FP64  384^3 13.15 / 13.02   512^3 29.37 / 29.53   768^3 100.0 / 105.1   1024^3 230 / 240
    //   FP32  256^3 1.67 / 1.58     384^3 6.31 / 6.36     448^3 10.24 / 10.18   512^3 12.78 / 13.44
    // so images below 512^3 cells in FP64 and below 384^3 in FP32.
#if AC_DOUBLE_PRECISION == 1
    return (double)nx * ny * nz < 512.0 * 512.0 * 512.0;
#else
    return (double)nx * ny * nz < 384.0 * 384.0 * 384.0;
#endif
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

// Compute units of the current GPU (queried once per device; 256 if the query fails).
static int
current_cu_count()
{
    constexpr int MAXDEV = 64;
    static std::atomic<int> known[MAXDEV]; // 0: not queried yet
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return 256;
    if (dev < MAXDEV) {
        const int c = known[dev].load(std::memory_order_relaxed);
        if (c > 0)
            return c;
    }
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    if (dev < MAXDEV)
        known[dev].store(n, std::memory_order_relaxed);
    return n;
}

int
balanced_z_depth(int nx, int ny, int nz, int tile_x, int tile_y, int blocks_per_cu)
{
    // A z-march block costs its depth plus the 6 planes that fill its queues before the first
    // output. Blocks are dispatched in rounds of `slots` (the resident block slots of the GPU),
    // so a launch costs about rounds x (depth + 6). Splits of the box into k equal z-chunks of at
    // least 8 planes are considered up to four times the k that fills every slot in the first
    // round, and the cheapest is taken (the smallest k among equals). Equal chunks keep each
    // XCD's contiguous block range equally loaded: an uneven last chunk leaves the XCDs that hold
    // its blocks idle for the rest of the launch.
    const long cols  = (long)((nx + tile_x - 1) / tile_x) * ((ny + tile_y - 1) / tile_y);
    const long slots = (long)blocks_per_cu * current_cu_count();
    const long kfill = std::max(1L, (slots + cols - 1) / cols);
    const int kmax   = (int)std::max(1L, std::min(4 * kfill, (long)nz / 8));
    int best_zd      = nz;
    double best_cost = INFINITY;
    for (int k = 1; k <= kmax; ++k) {
        const int zd       = (nz + k - 1) / k;
        const long blocks  = cols * ((nz + zd - 1) / zd);
        const long rounds  = (blocks + slots - 1) / slots;
        const double cost  = (double)rounds * (zd + 6);
        if (cost < best_cost) {
            best_cost = cost;
            best_zd   = zd;
        }
    }
    return best_zd >= nz ? 0 : best_zd;
}

AcKernelConfig
default_kernel_config(int nx, int ny, int nz)
{
    // Fastest at 512^3: the lean z-march on 16 x 16 tiles with a 4-slot ring, loading the
    // previous stage first and deferring its stores; in FP32 compiled for two waves per SIMD with
    // pinned LDS windows (variant 19: two blocks of 4 waves per CU, 64 KiB of LDS each), in FP64
    // for one wave per SIMD (variant 8: its 128 KiB ring leaves room for one block per CU). The
    // z extent of a block is cut into equal chunks when the tile columns alone leave resident
    // block slots empty (balanced_z_depth).
    AcKernelConfig c;
    c.variant = AC_DOUBLE_PRECISION ? 8 : 19;
    c.tile_x  = 16;
    c.tile_y  = 16;
    c.z_depth = balanced_z_depth(nx, ny, nz, c.tile_x, c.tile_y, AC_DOUBLE_PRECISION ? 1 : 2);
    return c;
}

hipError_t
launch_substep(const AcKernelConfig& c, const SubstepArgs& a, hipStream_t stream)
{
    if (a.end.x <= a.start.x || a.end.y <= a.start.y || a.end.z <= a.start.z)
        return hipSuccess;
    if (a.wrap) { // ghost cells read at their periodic images (the WRAP build)
        if (a.cin[0] != nullptr || !kernel_config_reads_images(c))
            return hipErrorInvalidConfiguration;
#if AC_DOUBLE_PRECISION == 1
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 1, false, 4, true, true, 7, false, true>,
                                          a, c.z_depth, stream);
#else
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, true, true, 7, true, true>,
                                          a, c.z_depth, stream);
#endif
    }
    if (a.cin[0] != nullptr) {
        // compensated integration: the lean z-march in its default tile, z-depth from `c`
        const int zd = c.variant >= 1 ? c.z_depth : 0;
#if AC_DOUBLE_PRECISION == 0
        return launch_zmarch_grid<32, 4>(solve_zlean<32, 4, 1, true, 4, false, true>, a, zd,
                                         stream);
#else
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 1, true>, a, zd, stream);
#endif
    }
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
    if (c.variant == 3 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 1>, a, c.z_depth, stream);
#if AC_DOUBLE_PRECISION == 0
    if (c.variant == 3 && c.tile_x == 32 && c.tile_y == 4)
        return launch_zmarch_grid<32, 4>(solve_zlean<32, 4, 1>, a, c.z_depth, stream);
    if (c.variant == 3 && c.tile_x == 16 && c.tile_y == 8)
        return launch_zmarch_grid<16, 8>(solve_zlean<16, 8, 1>, a, c.z_depth, stream);
    if (c.variant == 4 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2>, a, c.z_depth, stream);
    if (c.variant == 5 && c.tile_x == 32 && c.tile_y == 4)
        return launch_zmarch_grid<32, 4>(solve_zaos<32, 4, 1>, a, c.z_depth, stream);
    if (c.variant == 5 && c.tile_x == 16 && c.tile_y == 8)
        return launch_zmarch_grid<16, 8>(solve_zaos<16, 8, 1>, a, c.z_depth, stream);
    if (c.variant == 5 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zaos<16, 16, 1>, a, c.z_depth, stream);
    if (c.variant == 6 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zaos<16, 16, 2>, a, c.z_depth, stream);
    if (c.variant == 6 && c.tile_x == 32 && c.tile_y == 4)
        return launch_zmarch_grid<32, 4>(solve_zaos<32, 4, 2>, a, c.z_depth, stream);
    if (c.variant == 7 && c.tile_x == 32 && c.tile_y == 4)
        return launch_zmarch_grid<32, 4>(solve_zlean<32, 4, 1, false, 5>, a, c.z_depth, stream);
    if (c.variant == 7 && c.tile_x == 16 && c.tile_y == 8)
        return launch_zmarch_grid<16, 8>(solve_zlean<16, 8, 1, false, 5>, a, c.z_depth, stream);
    if (c.variant == 8 && c.tile_x == 32 && c.tile_y == 4)
        return launch_zmarch_grid<32, 4>(solve_zlean<32, 4, 1, false, 4, true, true>, a, c.z_depth,
                                         stream);
    if (c.variant == 9 && c.tile_x == 32 && c.tile_y == 4)
        return launch_zmarch_grid<32, 4>(solve_zlean<32, 4, 1, false, 5, true, true>, a, c.z_depth,
                                         stream);
    // two waves per SIMD (16 x 16 is the only tile whose 4-slot ring leaves room for two blocks)
    if (c.variant == 10 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, true, false>, a,
                                          c.z_depth, stream);
    if (c.variant == 11 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, false, true>, a,
                                          c.z_depth, stream);
    if (c.variant == 12 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, true, true>, a,
                                          c.z_depth, stream);
    // 8-phase plane loop (compile-time ring slots)
    if (c.variant == 13 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, true, true, 8>, a,
                                          c.z_depth, stream);
    // 5-slot ring (one barrier per plane), two waves per SIMD: 80 KiB of LDS per block
    if (c.variant == 16 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 5, true, true>, a,
                                          c.z_depth, stream);
    // pinned LDS windows (WIN) on the two-wave builds: 5-slot ring; 4-slot ring with deferred
    // stores and previous stage first; 4-slot ring with deferred stores
    if (c.variant == 17 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 5, true, true, 7, true>,
                                          a, c.z_depth, stream);
    if (c.variant == 19 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, true, true, 7, true>,
                                          a, c.z_depth, stream);
    if (c.variant == 22 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 2, false, 4, true, false, 7, true>,
                                          a, c.z_depth, stream);
    // direct-to-LDS plane loads into a 5-slot ring, pinned LDS windows
    if (c.variant == 23 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(
            solve_zlean<16, 16, 2, false, 5, true, true, 7, true, false, true>, a, c.z_depth,
            stream);
#endif
    if (c.variant == 8 && c.tile_x == 16 && c.tile_y == 16)
        return launch_zmarch_grid<16, 16>(solve_zlean<16, 16, 1, false, 4, true, true>, a,
                                          c.z_depth, stream);
#undef AC_DIRECT
#undef AC_ZMARCH
#undef AC_ZMARCH2
    return hipErrorInvalidConfiguration;
}

bool
batch_tile_supported(int tx, int ty)
{
    return (tx == 32 && ty == 4) || (tx == 16 && ty == 8) || (tx == 16 && ty == 16);
}

hipError_t
launch_substep_batch(int tx, int ty, const SubstepArgs* boxes, const BatchItem* items, int nitems,
                     int nblocks, hipStream_t stream)
{
    if (nitems <= 0)
        return hipSuccess;
    if (nblocks <= 0 || nblocks > nitems)
        nblocks = nitems;
    if (tx == 32 && ty == 4)
        solve_zlean_batch<32, 4><<<dim3(nblocks), dim3(128), 0, stream>>>(boxes, items, nitems);
    else if (tx == 16 && ty == 8)
        solve_zlean_batch<16, 8><<<dim3(nblocks), dim3(128), 0, stream>>>(boxes, items, nitems);
    else if (tx == 16 && ty == 16)
        AC_BATCH16<<<dim3(nblocks), dim3(256), 0, stream>>>(boxes, items, nitems);
    else
        return hipErrorInvalidConfiguration;
    return hipGetLastError();
}

int
batch_resident_blocks(int tx, int ty)
{
    int per_cu = 0, cus = 0, dev = 0;
    hipError_t e = hipErrorInvalidConfiguration;
    if (tx == 32 && ty == 4)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, solve_zlean_batch<32, 4>, 128, 0);
    else if (tx == 16 && ty == 8)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, solve_zlean_batch<16, 8>, 128, 0);
    else if (tx == 16 && ty == 16)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, AC_BATCH16, 256, 0);
    if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    return per_cu * cus;
}
#undef AC_BATCH16

} // namespace ac
