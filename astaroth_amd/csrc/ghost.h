// Enumeration of the ghost shell of an n_x x n_y x n_z box with 3 ghost layers, shared by the
// periodic fill of one box and of an ensemble of boxes. Cell t of the shell is
//   Z slabs: k in [0,3) u [n_z+3, m_z), all (i, j)            6 m_x m_y cells
//   Y slabs: k interior, j in [0,3) u [n_y+3, m_y), all i     6 n_z m_x cells
//   X slabs: k, j interior, i in [0,3) u [n_x+3, m_x)         6 n_z n_y cells
// so consecutive t of the Z and Y slabs walk contiguous x rows.
#pragma once
#include "ac_internal.h"

namespace ac {

struct GhostShell {
    int nx, ny, nz, mx, my;
    long cz, cy, cx; // cell counts of the three slab families

    __host__ __device__ static GhostShell of(int nx, int ny, int nz)
    {
        GhostShell g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.mx = nx + 6;
        g.my = ny + 6;
        g.cz = 6L * g.mx * g.my;
        g.cy = 6L * nz * g.mx;
        g.cx = 6L * nz * ny;
        return g;
    }
    __host__ __device__ long cells() const { return cz + cy + cx; }

    // Ghost cell t (0 <= t < cells()) -> (i, j, k).
    __device__ __forceinline__ void cell(long t, int& i, int& j, int& k) const
    {
        if (t < cz) {
            const long plane = (long)mx * my;
            const int pz     = (int)(t / plane);
            const long rem   = t - (long)pz * plane;
            k                = pz < 3 ? pz : nz + pz;
            j                = (int)(rem / mx);
            i                = (int)(rem - (long)j * mx);
        }
        else if (t < cz + cy) {
            const long u    = t - cz;
            const long slab = 6L * mx;
            const int kk    = (int)(u / slab);
            const long rem  = u - (long)kk * slab;
            const int jr    = (int)(rem / mx);
            k               = 3 + kk;
            j               = jr < 3 ? jr : ny + jr;
            i               = (int)(rem - (long)jr * mx);
        }
        else {
            const long u    = t - cz - cy;
            const long slab = 6L * ny;
            const int kk    = (int)(u / slab);
            const int rem   = (int)(u - (long)kk * slab);
            const int ir    = rem % 6;
            k               = 3 + kk;
            j               = 3 + rem / 6;
            i               = ir < 3 ? ir : nx + ir;
        }
    }

    // Linear indices of ghost cell (i, j, k) and of its periodic interior image.
    __device__ __forceinline__ void indices(int i, int j, int k, long& dst, long& src) const
    {
        const int is = (i - 3 + nx) % nx + 3;
        const int js = (j - 3 + ny) % ny + 3;
        const int ks = (k - 3 + nz) % nz + 3;
        const long mxy = (long)mx * my;
        dst            = i + (long)j * mx + (long)k * mxy;
        src            = is + (long)js * mx + (long)ks * mxy;
    }
};

} // namespace ac
