// Gathers of the potential around a particle, shared by the forward
// interpolation and the adjoint gather.
#pragma once

#include "pm_common.h"

namespace pm {

// Loads of 4 / 2 consecutive mesh values along z from an element-aligned
// address (one global_load_dwordx4 / dwordx2 for FP32).
typedef float pm_f4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float pm_f2 __attribute__((ext_vector_type(2), aligned(4)));
typedef double pm_d2 __attribute__((ext_vector_type(2), aligned(8)));

__device__ __forceinline__ void load_run4(const float* p, float (&v)[4]) {
    const pm_f4 q = *reinterpret_cast<const pm_f4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void load_run2(const float* p, float (&v)[2]) {
    const pm_f2 q = *reinterpret_cast<const pm_f2*>(p);
    v[0] = q.x; v[1] = q.y;
}
__device__ __forceinline__ void load_run4(const double* p, double (&v)[4]) {
    const pm_d2 q0 = *reinterpret_cast<const pm_d2*>(p), q1 = *reinterpret_cast<const pm_d2*>(p + 2);
    v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
}
__device__ __forceinline__ void load_run2(const double* p, double (&v)[2]) {
    const pm_d2 q = *reinterpret_cast<const pm_d2*>(p);
    v[0] = q.x; v[1] = q.y;
}

// Mesh element address: 32-bit byte offsets from the uniform base for FP32
// (ng <= 1024 keeps them below 4 GiB), 64-bit for FP64.
template <typename Rl>
__device__ __forceinline__ const Rl* mesh_at(const Rl* phi, uint32_t e) {
    if constexpr (sizeof(Rl) == 4)
        return reinterpret_cast<const Rl*>(reinterpret_cast<const char*>(phi) + (e << 2));
    else
        return phi + e;
}

// Cell of a particle: periodic neighbours -1, 0, +1, +2 on x and y as
// element offsets, the z cell and the fractional offsets.
template <typename Rl> struct CellIdx {
    uint32_t X[4], Y[4];
    int iz;
    Rl tx, ty, tz;
};

// From an integer cell and the offsets inside it.
template <typename Rl>
__device__ __forceinline__ CellIdx<Rl> cell_at(int ix, int iy, int iz, Rl tx, Rl ty, Rl tz, int N) {
    CellIdx<Rl> c;
    c.tx = tx;
    c.ty = ty;
    c.tz = tz;
    const uint32_t n = (uint32_t)N;
    c.X[0] = (ix == 0 ? n - 1 : ix - 1) * n * n;
    c.X[1] = (uint32_t)ix * n * n;
    c.X[2] = (ix + 1 == N ? 0u : ix + 1) * n * n;
    c.X[3] = (ix + 2 >= N ? ix + 2 - N : ix + 2) * n * n;
    c.Y[0] = (iy == 0 ? n - 1 : iy - 1) * n;
    c.Y[1] = (uint32_t)iy * n;
    c.Y[2] = (iy + 1 == N ? 0u : iy + 1) * n;
    c.Y[3] = (iy + 2 >= N ? iy + 2 - N : iy + 2) * n;
    c.iz = iz;
    return c;
}

// From an FP64 position (cell by truncation, utils.cpp:79-88).
template <typename Rl>
__device__ __forceinline__ CellIdx<Rl> cell_of(double x, double y, double z, int N) {
    const int ix = (int)x, iy = (int)y, iz = (int)z;
    return cell_at<Rl>(ix, iy, iz, (Rl)(x - ix), (Rl)(y - iy), (Rl)(z - iz), N);
}

// The 32 potential values that the centred differences at the 8 corners of
// a particle's cell need sit on 12 z-runs: 4 runs of 4 (z-1..z+2) through
// the central x/y cells and 8 runs of 2 (z, z+1) for the x and y arms; away
// from the periodic z seam each run is a single vector load.
//   cl[a][b][c]: x offset a, y offset b (0, +1), z offset c-1
//   xr[s][b][c]: x offset -1 (s=0) or +2 (s=1), y offset b, z offset c
//   yr[s][a][c]: y offset -1 (s=0) or +2 (s=1), x offset a, z offset c
// The gather is split in two so that a thread can have the loads of several
// particles in flight at once: gather_runs_main issues the 12 vector loads
// with their z starts clamped into the mesh (always in bounds, no branch);
// gather_runs_seam then reloads value by value, with wrapped indices, the
// runs of a particle whose cell touches the seam (iz == 0 or iz + 2 >= N),
// for which the clamped runs hold the wrong values.  N >= 8.
template <typename Rl>
__device__ __forceinline__ void gather_runs_main(const Rl* __restrict__ phi, int N, const CellIdx<Rl>& c,
                                                 Rl (&cl)[2][2][4], Rl (&xr)[2][2][2], Rl (&yr)[2][2][2]) {
    const int z4 = min(max(c.iz - 1, 0), N - 4), z2 = min(c.iz, N - 2);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            load_run4(mesh_at(phi, c.X[a + 1] + c.Y[b + 1] + z4), cl[a][b]);
            load_run2(mesh_at(phi, c.X[a * 3] + c.Y[b + 1] + z2), xr[a][b]);
            load_run2(mesh_at(phi, c.X[b + 1] + c.Y[a * 3] + z2), yr[a][b]);
        }
}

// Seam reload, two forms with the same values.  All 32 values, one by one
// at wrapped z indices:
template <typename Rl>
__device__ __forceinline__ void seam_reload_all(const Rl* __restrict__ phi, int N, const CellIdx<Rl>& c,
                                                Rl (&cl)[2][2][4], Rl (&xr)[2][2][2], Rl (&yr)[2][2][2]) {
    const int iz = c.iz;
    if (iz >= 1 && iz + 2 < N) return;
    const uint32_t Z[4] = {(uint32_t)(iz == 0 ? N - 1 : iz - 1), (uint32_t)iz, (uint32_t)(iz + 1 == N ? 0 : iz + 1),
                           (uint32_t)(iz + 2 >= N ? iz + 2 - N : iz + 2)};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
            for (int k = 0; k < 4; ++k) cl[a][b][k] = *mesh_at(phi, c.X[a + 1] + c.Y[b + 1] + Z[k]);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                xr[a][b][k] = *mesh_at(phi, c.X[a * 3] + c.Y[b + 1] + Z[k + 1]);
                yr[a][b][k] = *mesh_at(phi, c.X[b + 1] + c.Y[a * 3] + Z[k + 1]);
            }
        }
}

// Only the missing ones: the clamped runs already hold all but one or two
// of a seam cell's values, shifted, so they are moved into place and the
// rest loaded (iz == 0: z = N-1 on the 4-runs; iz == N-2: z = 0 on the
// 4-runs; iz == N-1: z = 0 and 1 on the 4-runs, z = 0 on the 2-runs).  A
// wave issues 4 loads at the start of a z line and 16 at its end, instead
// of reloading 32 values.
template <typename Rl>
__device__ __forceinline__ void seam_reload_missing(const Rl* __restrict__ phi, int N, const CellIdx<Rl>& c,
                                                    Rl (&cl)[2][2][4], Rl (&xr)[2][2][2], Rl (&yr)[2][2][2]) {
    const int iz = c.iz;
    const bool s0 = iz == 0, s2 = iz == N - 2, s1 = iz == N - 1;
    if (!__ballot(s0 || s2 || s1)) return;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const Rl r0 = cl[a][b][0], r1 = cl[a][b][1], r2 = cl[a][b][2], r3 = cl[a][b][3];
            cl[a][b][0] = s2 ? r1 : s1 ? r2 : r0;
            cl[a][b][1] = s0 ? r0 : s2 ? r2 : s1 ? r3 : r1;
            cl[a][b][2] = s0 ? r1 : s2 ? r3 : r2;
            cl[a][b][3] = s0 ? r2 : r3;
            if (s1) {
                xr[a][b][0] = xr[a][b][1];
                yr[a][b][0] = yr[a][b][1];
            }
        }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const uint32_t col = c.X[a + 1] + c.Y[b + 1];
            if (s0) cl[a][b][0] = *mesh_at(phi, col + (uint32_t)(N - 1));
            if (s2 || s1) cl[a][b][3] = *mesh_at(phi, col + (s1 ? 1u : 0u));
            if (s1) {
                cl[a][b][2] = *mesh_at(phi, col);
                xr[a][b][1] = *mesh_at(phi, c.X[a * 3] + c.Y[b + 1]);
                yr[a][b][1] = *mesh_at(phi, c.X[b + 1] + c.Y[a * 3]);
            }
        }
}

// An FP32 mesh reloads all values (the shorter reload measured 2% slower
// there: 3.19 -> 3.25 ms for the 512^3 interpolation), an FP64 mesh only
// the missing ones (4.31 -> 4.23 ms); the results are the same bit for bit.
#ifndef PM_SEAM_MIN32
#define PM_SEAM_MIN32 0
#endif
template <typename Rl>
__device__ __forceinline__ void gather_runs_seam(const Rl* __restrict__ phi, int N, const CellIdx<Rl>& c,
                                                 Rl (&cl)[2][2][4], Rl (&xr)[2][2][2], Rl (&yr)[2][2][2]) {
    if constexpr (sizeof(Rl) == 8 || PM_SEAM_MIN32)
        seam_reload_missing(phi, N, c, cl, xr, yr);
    else
        seam_reload_all(phi, N, c, cl, xr, yr);
}

template <typename Rl>
__device__ __forceinline__ void gather_runs(const Rl* __restrict__ phi, int N, const CellIdx<Rl>& c,
                                            Rl (&cl)[2][2][4], Rl (&xr)[2][2][2], Rl (&yr)[2][2][2]) {
    gather_runs_main(phi, N, c, cl, xr, yr);
    gather_runs_seam(phi, N, c, cl, xr, yr);
}


// Values of a mesh at the 8 corners of the cell (4 z-runs of 2), in the
// same two parts as gather_runs: the runs at z starts clamped into the mesh,
// then the values of a cell on the seam (iz + 1 == N) reloaded one by one.
template <typename Rl>
__device__ __forceinline__ void gather_corners_main(const Rl* __restrict__ m, int N, const CellIdx<Rl>& c,
                                                    Rl (&f)[2][2][2]) {
    const int z2 = min(c.iz, N - 2);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) load_run2(mesh_at(m, c.X[a + 1] + c.Y[b + 1] + z2), f[a][b]);
}

template <typename Rl>
__device__ __forceinline__ void gather_corners_seam(const Rl* __restrict__ m, int N, const CellIdx<Rl>& c,
                                                    Rl (&f)[2][2][2]) {
    const int iz = c.iz;
    if (iz + 1 < N) return;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            f[a][b][0] = *mesh_at(m, c.X[a + 1] + c.Y[b + 1] + iz);
            f[a][b][1] = *mesh_at(m, c.X[a + 1] + c.Y[b + 1] + 0u);
        }
}

template <typename Rl>
__device__ __forceinline__ void gather_corners(const Rl* __restrict__ m, int N, const CellIdx<Rl>& c,
                                               Rl (&f)[2][2][2]) {
    gather_corners_main(m, N, c, f);
    gather_corners_seam(m, N, c, f);
}

// Centred differences g = (phi[i-1] - phi[i+1]) / 2 (centered_derivatives,
// utils.cpp:130-132) at the 8 corners, from the runs of gather_runs.
template <typename Rl>
__device__ __forceinline__ void corner_differences(const Rl (&cl)[2][2][4], const Rl (&xr)[2][2][2],
                                                   const Rl (&yr)[2][2][2], Rl (&gx)[2][2][2], Rl (&gy)[2][2][2],
                                                   Rl (&gz)[2][2][2]) {
    const Rl h = Rl(0.5);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const Rl xm = a == 0 ? xr[0][b][cc] : cl[0][b][cc + 1];
                const Rl xp = a == 0 ? cl[1][b][cc + 1] : xr[1][b][cc];
                gx[a][b][cc] = (xm - xp) * h;
                const Rl ym = b == 0 ? yr[0][a][cc] : cl[a][0][cc + 1];
                const Rl yp = b == 0 ? cl[a][1][cc + 1] : yr[1][a][cc];
                gy[a][b][cc] = (ym - yp) * h;
                gz[a][b][cc] = (cl[a][b][cc] - cl[a][b][cc + 2]) * h;
            }
}

// trilerp1 (utils_inline.h:55-70) of corner values f at offsets (tx, ty, tz)
template <typename Rl>
__device__ __forceinline__ Rl trilerp(const Rl (&f)[2][2][2], Rl tx, Rl ty, Rl tz) {
    const Rl mx = Rl(1) - tx, my = Rl(1) - ty, mz = Rl(1) - tz;
    const Rl f00 = f[0][0][0] * mx + f[1][0][0] * tx;
    const Rl f10 = f[0][1][0] * mx + f[1][1][0] * tx;
    const Rl fz0 = f00 * my + f10 * ty;
    const Rl g00 = f[0][0][1] * mx + f[1][0][1] * tx;
    const Rl g10 = f[0][1][1] * mx + f[1][1][1] * tx;
    const Rl fz1 = g00 * my + g10 * ty;
    return fz0 * mz + fz1 * tz;
}

// Centred-difference force of the periodic potential, trilinearly
// interpolated at a particle (centered_derivatives + trilerp1).
template <typename Rl>
__device__ __forceinline__ void force_from_runs(const Rl (&cl)[2][2][4], const Rl (&xr)[2][2][2],
                                                const Rl (&yr)[2][2][2], Rl tx, Rl ty, Rl tz, Rl (&F)[3]) {
    Rl gx[2][2][2], gy[2][2][2], gz[2][2][2];
    corner_differences(cl, xr, yr, gx, gy, gz);
    F[0] = trilerp(gx, tx, ty, tz);
    F[1] = trilerp(gy, tx, ty, tz);
    F[2] = trilerp(gz, tx, ty, tz);
}

template <typename Rl>
__device__ __forceinline__ void force_at(const Rl* __restrict__ phi, int N, const CellIdx<Rl>& c, Rl (&F)[3]) {
    Rl cl[2][2][4], xr[2][2][2], yr[2][2][2];
    gather_runs(phi, N, c, cl, xr, yr);
    force_from_runs(cl, xr, yr, c.tx, c.ty, c.tz, F);
}

template <typename Rl>
__device__ __forceinline__ void cic_force(const Rl* __restrict__ phi, int N, double x, double y, double z,
                                          double (&F)[3]) {
    Rl f[3];
    force_at(phi, N, cell_of<Rl>(x, y, z, N), f);
    F[0] = (double)f[0];
    F[1] = (double)f[1];
    F[2] = (double)f[2];
}

}  // namespace pm
