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

    // Shells whose cell count fits 32 bits (every box up to about 10^4 cells per axis) are
    // enumerated in 32-bit arithmetic: a 64-bit division costs several times a 32-bit one, and
    // the fill does little else per cell.
    __host__ __device__ bool fits_32bit() const { return cells() < 0xFFFFFFFFL; }

    // Ghost cell t (0 <= t < cells() < 2^32) -> (i, j, k).
    __device__ __forceinline__ void cell(unsigned t, int& i, int& j, int& k) const
    {
        const unsigned ucz = (unsigned)cz, ucy = (unsigned)cy, umx = (unsigned)mx;
        if (t < ucz) {
            const unsigned plane = umx * (unsigned)my;
            const unsigned pz    = t / plane;
            const unsigned rem   = t - pz * plane;
            const unsigned jj    = rem / umx;
            k                    = pz < 3 ? (int)pz : nz + (int)pz;
            j                    = (int)jj;
            i                    = (int)(rem - jj * umx);
        }
        else if (t < ucz + ucy) {
            const unsigned u    = t - ucz;
            const unsigned slab = 6u * umx;
            const unsigned kk   = u / slab;
            const unsigned rem  = u - kk * slab;
            const unsigned jr   = rem / umx;
            k                   = 3 + (int)kk;
            j                   = jr < 3 ? (int)jr : ny + (int)jr;
            i                   = (int)(rem - jr * umx);
        }
        else {
            const unsigned u    = t - ucz - ucy;
            const unsigned slab = 6u * (unsigned)ny;
            const unsigned kk   = u / slab;
            const unsigned rem  = u - kk * slab;
            const unsigned row  = rem / 6u;
            const unsigned ir   = rem - 6u * row;
            k                   = 3 + (int)kk;
            j                   = 3 + (int)row;
            i                   = ir < 3 ? (int)ir : nx + (int)ir;
        }
    }

    // Interior image of coordinate c in [0, n + 6) on an axis of n interior cells: one shift by n
    // when n >= 3, else as many as it takes.
    static __device__ __forceinline__ int image(int c, int n)
    {
        if (n >= 3)
            return c < 3 ? c + n : c >= n + 3 ? c - n : c;
        const int r = (c - 3) % n;
        return (r < 0 ? r + n : r) + 3;
    }

    // Linear indices of ghost cell (i, j, k) (each coordinate in [0, n + 6)) and of its periodic
    // interior image.
    __device__ __forceinline__ void indices(int i, int j, int k, long& dst, long& src) const
    {
        const int is = image(i, nx), js = image(j, ny), ks = image(k, nz);
        const long mxy = (long)mx * my;
        dst            = i + (long)j * mx + (long)k * mxy;
        src            = is + (long)js * mx + (long)ks * mxy;
    }
};

} // namespace ac
