// Particle kernels of the fused one-box step: CIC deposit (atomic variant)
// and interpolation of the centred-difference force fused with kick and
// drift.  Positions and velocities are FP64; the mesh is FP32 or FP64.
#include "pm_internal.h"

namespace pm {

constexpr int kPartThreads = 256;

int launch_fill(int prec, void* p, size_t n, double v, hipStream_t st) {
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (prec == PM_FP32)
        hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(256), 0, st, (float*)p, n, (float)v);
    else
        hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, st, (double*)p, n, v);
    return check_launch("fill");
}

__device__ __forceinline__ bool in_box(double x, double box) { return x >= 0.0 && x < box; }

// deposit_particles + cic_deposit (utils.cpp:70-102, utils_inline.h:25-49)
// on the periodic mesh: node ix+1 == ng wraps to 0, which is the ghost fold
// of reorder_slab.cpp:43-67 done in place.  Weights are formed in FP64 with
// the reference's products and rounded once to the mesh precision.
template <typename Rl>
__global__ __launch_bounds__(kPartThreads) void k_deposit_atomic(const double* __restrict__ px,
                                                                 const double* __restrict__ py,
                                                                 const double* __restrict__ pz,
                                                                 int64_t np, int ng, double mass,
                                                                 Rl* __restrict__ rho,
                                                                 unsigned long long* __restrict__ err) {
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kPartThreads + threadIdx.x;
    if (i >= np) return;
    const double x = px[i], y = py[i], z = pz[i];
    const double box = (double)ng;
    if (!(in_box(x, box) && in_box(y, box) && in_box(z, box))) {
        atomicAdd(err, 1ULL);
        return;
    }
    const int ix = (int)x, iy = (int)y, iz = (int)z;
    const double tx = x - ix, ty = y - iy, tz = z - iz;
    const double sx0 = 1.0 - tx, sy0 = 1.0 - ty;
    const double sxy00 = sx0 * sy0, sxy10 = tx * sy0, sxy01 = sx0 * ty, sxy11 = tx * ty;
    const double m0 = mass * (1.0 - tz), m1 = mass * tz;
    const size_t N = (size_t)ng;
    const size_t x0 = (size_t)ix * N * N, x1 = (size_t)(ix + 1 == ng ? 0 : ix + 1) * N * N;
    const size_t y0 = (size_t)iy * N, y1 = (size_t)(iy + 1 == ng ? 0 : iy + 1) * N;
    const size_t z0 = (size_t)iz, z1 = (size_t)(iz + 1 == ng ? 0 : iz + 1);
    unsafeAtomicAdd(rho + x0 + y0 + z0, (Rl)(sxy00 * m0));
    unsafeAtomicAdd(rho + x1 + y0 + z0, (Rl)(sxy10 * m0));
    unsafeAtomicAdd(rho + x0 + y1 + z0, (Rl)(sxy01 * m0));
    unsafeAtomicAdd(rho + x1 + y1 + z0, (Rl)(sxy11 * m0));
    unsafeAtomicAdd(rho + x0 + y0 + z1, (Rl)(sxy00 * m1));
    unsafeAtomicAdd(rho + x1 + y0 + z1, (Rl)(sxy10 * m1));
    unsafeAtomicAdd(rho + x0 + y1 + z1, (Rl)(sxy01 * m1));
    unsafeAtomicAdd(rho + x1 + y1 + z1, (Rl)(sxy11 * m1));
}

int launch_deposit_atomic(int prec, const double* px, const double* py, const double* pz, int64_t np,
                          int ng, double mass, void* rho, unsigned long long* err, hipStream_t st) {
    if (np <= 0) return PM_OK;
    const unsigned nb = (unsigned)((np + kPartThreads - 1) / kPartThreads);
    if (prec == PM_FP32)
        hipLaunchKernelGGL(k_deposit_atomic<float>, dim3(nb), dim3(kPartThreads), 0, st, px, py, pz, np, ng,
                           mass, (float*)rho, err);
    else
        hipLaunchKernelGGL(k_deposit_atomic<double>, dim3(nb), dim3(kPartThreads), 0, st, px, py, pz, np,
                           ng, mass, (double*)rho, err);
    return check_launch("deposit_atomic");
}

__device__ __forceinline__ double wrap_pos(double x, double box) {
    if (x < 0.0) x += box;
    if (x >= box) x -= box;
    return x;
}

// Centred-difference force (centered_derivatives, utils.cpp:130-132) of the
// periodic potential, trilinearly interpolated (trilerp1, utils_inline.h:55-70)
// at a particle in cell (ix, iy, iz) with offsets (tx, ty, tz).
template <typename Rl>
__device__ __forceinline__ void cic_force(const Rl* __restrict__ phi, int N, double x, double y, double z,
                                          double (&F)[3]) {
    const int ix = (int)x, iy = (int)y, iz = (int)z;
    const Rl tx = (Rl)(x - ix), ty = (Rl)(y - iy), tz = (Rl)(z - iz);
    const Rl mx = Rl(1) - tx, my = Rl(1) - ty, mz = Rl(1) - tz;
    // periodic neighbours -1, 0, +1, +2 on each axis
    const int xa[4] = {ix == 0 ? N - 1 : ix - 1, ix, ix + 1 == N ? 0 : ix + 1,
                       ix + 2 >= N ? ix + 2 - N : ix + 2};
    const int ya[4] = {iy == 0 ? N - 1 : iy - 1, iy, iy + 1 == N ? 0 : iy + 1,
                       iy + 2 >= N ? iy + 2 - N : iy + 2};
    const int za[4] = {iz == 0 ? N - 1 : iz - 1, iz, iz + 1 == N ? 0 : iz + 1,
                       iz + 2 >= N ? iz + 2 - N : iz + 2};
    auto P = [&](int a, int b, int c) -> Rl {  // a, b, c in [0, 4): offsets -1..+2
        return phi[((size_t)xa[a] * N + ya[b]) * N + za[c]];
    };
    // central 2x2x2 block and the six arms
    Rl c[2][2][2], ax[2][2][2], ay[2][2][2], az[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) c[a][b][cc] = P(a + 1, b + 1, cc + 1);
    // ax[s][b][c]: s=0 -> x offset -1, s=1 -> x offset +2
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                ax[s][b][cc] = P(s * 3, b + 1, cc + 1);
                ay[s][b][cc] = P(b + 1, s * 3, cc + 1);
                az[s][b][cc] = P(b + 1, cc + 1, s * 3);
            }
    // gradient at the 8 corners: g = (phi[i-1] - phi[i+1]) / 2
    Rl gx[2][2][2], gy[2][2][2], gz[2][2][2];
    const Rl h = Rl(0.5);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const Rl xm = a == 0 ? ax[0][b][cc] : c[0][b][cc];
                const Rl xp = a == 0 ? c[1][b][cc] : ax[1][b][cc];
                gx[a][b][cc] = (xm - xp) * h;
                const Rl ym = b == 0 ? ay[0][a][cc] : c[a][0][cc];
                const Rl yp = b == 0 ? c[a][1][cc] : ay[1][a][cc];
                gy[a][b][cc] = (ym - yp) * h;
                const Rl zm = cc == 0 ? az[0][a][b] : c[a][b][0];
                const Rl zp = cc == 0 ? c[a][b][1] : az[1][a][b];
                gz[a][b][cc] = (zm - zp) * h;
            }
    auto tri = [&](const Rl (&f)[2][2][2]) -> Rl {  // trilerp1 order
        const Rl f00 = f[0][0][0] * mx + f[1][0][0] * tx;
        const Rl f10 = f[0][1][0] * mx + f[1][1][0] * tx;
        const Rl fz0 = f00 * my + f10 * ty;
        const Rl g00 = f[0][0][1] * mx + f[1][0][1] * tx;
        const Rl g10 = f[0][1][1] * mx + f[1][1][1] * tx;
        const Rl fz1 = g00 * my + g10 * ty;
        return fz0 * mz + fz1 * tz;
    };
    F[0] = (double)tri(gx);
    F[1] = (double)tri(gy);
    F[2] = (double)tri(gz);
}

// update_particles (utils.cpp:141-170) with the force of cic_force, then the
// drift that the reference leaves commented out (utils.cpp:145-146).
// MODE 0: v += g*kick, x += v*drift.  MODE 1: write g to (vx, vy, vz).
// MODE 2: as MODE 0, and bin the drifted position for the next step: its
// tile key goes to key[i], the tile counts to cnt (aggregated per workgroup
// in an LDS hash; a key that finds the hash full adds itself directly).
template <typename Rl, int MODE>
__global__ __launch_bounds__(kPartThreads) void k_interp(double* __restrict__ px, double* __restrict__ py,
                                                         double* __restrict__ pz, double* __restrict__ vx,
                                                         double* __restrict__ vy, double* __restrict__ vz,
                                                         int64_t np, int ng, const Rl* __restrict__ phi,
                                                         double kick, double drift, int ts,
                                                         uint32_t* __restrict__ cnt, uint32_t* __restrict__ key,
                                                         unsigned long long* __restrict__ err) {
    constexpr int HS = MODE == 2 ? 128 : 1;
    __shared__ uint32_t hkey[HS], hcnt[HS];
    if constexpr (MODE == 2) {
        for (int s = threadIdx.x; s < HS; s += kPartThreads) {
            hkey[s] = 0xFFFFFFFFu;
            hcnt[s] = 0u;
        }
        __syncthreads();
    }
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kPartThreads + threadIdx.x;
    const double box = (double)ng;
    bool ok = i < np;
    double x = 0.0, y = 0.0, z = 0.0;
    if (ok) {
        x = px[i];
        y = py[i];
        z = pz[i];
        if (!(in_box(x, box) && in_box(y, box) && in_box(z, box))) {
            atomicAdd(err, 1ULL);
            ok = false;
            if constexpr (MODE == 2) key[i] = 0xFFFFFFFFu;
        }
    }
    if (ok) {
        double F[3];
        cic_force(phi, ng, x, y, z, F);
        if constexpr (MODE == 1) {
            vx[i] = F[0];
            vy[i] = F[1];
            vz[i] = F[2];
        } else {
            const double nvx = vx[i] + F[0] * kick, nvy = vy[i] + F[1] * kick, nvz = vz[i] + F[2] * kick;
            vx[i] = nvx;
            vy[i] = nvy;
            vz[i] = nvz;
            if (drift != 0.0) {
                x = wrap_pos(x + nvx * drift, box);
                y = wrap_pos(y + nvy * drift, box);
                z = wrap_pos(z + nvz * drift, box);
                px[i] = x;
                py[i] = y;
                pz[i] = z;
            }
            if constexpr (MODE == 2) {
                // a particle that drifted more than a box away stays unbinned
                // and is counted, as the next step's range check would
                if (!(in_box(x, box) && in_box(y, box) && in_box(z, box))) {
                    atomicAdd(err, 1ULL);
                    key[i] = 0xFFFFFFFFu;
                    ok = false;
                }
            }
            if (MODE == 2 && ok) {
                const int nt = ng / ts;
                const uint32_t k = (uint32_t)((((int)x / ts) * nt + (int)y / ts) * nt + (int)z / ts);
                key[i] = k;
                uint32_t s = (k * 2654435761u) >> 25;  // 7 bits
                bool done = false;
                for (int probe = 0; probe < HS; ++probe) {
                    const uint32_t old = atomicCAS(&hkey[s], 0xFFFFFFFFu, k);
                    if (old == 0xFFFFFFFFu || old == k) {
                        atomicAdd(&hcnt[s], 1u);
                        done = true;
                        break;
                    }
                    s = (s + 1) & (HS - 1);
                }
                if (!done) atomicAdd(&cnt[k], 1u);
            }
        }
    }
    if constexpr (MODE == 2) {
        __syncthreads();
        for (int s = threadIdx.x; s < HS; s += kPartThreads)
            if (hkey[s] != 0xFFFFFFFFu) atomicAdd(&cnt[hkey[s]], hcnt[s]);
    }
}

int launch_interp(int prec, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                  int64_t np, int ng, const void* phi, double kick, double drift, int mode,
                  unsigned long long* err, hipStream_t st, int ts, uint32_t* cnt, uint32_t* key) {
    if (np <= 0) return PM_OK;
    const unsigned nb = (unsigned)((np + kPartThreads - 1) / kPartThreads);
    const dim3 g(nb), b(kPartThreads);
#define PM_INTERP(RL, M)                                                                                    \
    hipLaunchKernelGGL((k_interp<RL, M>), g, b, 0, st, px, py, pz, vx, vy, vz, np, ng, (const RL*)phi, kick, \
                       drift, ts, cnt, key, err)
    if (prec == PM_FP32) {
        if (mode == 0)
            PM_INTERP(float, 0);
        else if (mode == 1)
            PM_INTERP(float, 1);
        else
            PM_INTERP(float, 2);
    } else {
        if (mode == 0)
            PM_INTERP(double, 0);
        else if (mode == 1)
            PM_INTERP(double, 1);
        else
            PM_INTERP(double, 2);
    }
#undef PM_INTERP
    return check_launch("interp");
}

}  // namespace pm
