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

// Centred-difference force (centered_derivatives, utils.cpp:130-132) of the
// periodic potential, trilinearly interpolated (trilerp1, utils_inline.h:55-70)
// at a particle in cell (ix, iy, iz) with offsets (tx, ty, tz).  The 32
// potential values sit on 12 z-runs: 4 runs of 4 (z-1..z+2) through the
// central x/y cells and 8 runs of 2 (z, z+1) for the x and y arms; away from
// the periodic z seam each run is a single vector load.
template <typename Rl>
__device__ __forceinline__ void cic_force(const Rl* __restrict__ phi, int N, double x, double y, double z,
                                          double (&F)[3]) {
    const int ix = (int)x, iy = (int)y, iz = (int)z;
    const Rl tx = (Rl)(x - ix), ty = (Rl)(y - iy), tz = (Rl)(z - iz);
    const Rl mx = Rl(1) - tx, my = Rl(1) - ty, mz = Rl(1) - tz;
    const uint32_t n = (uint32_t)N;
    // periodic neighbours -1, 0, +1, +2 on x and y, as element offsets
    const uint32_t X[4] = {(ix == 0 ? n - 1 : ix - 1) * n * n, (uint32_t)ix * n * n,
                           (ix + 1 == N ? 0u : ix + 1) * n * n, (ix + 2 >= N ? ix + 2 - N : ix + 2) * n * n};
    const uint32_t Y[4] = {(iy == 0 ? n - 1 : iy - 1) * n, (uint32_t)iy * n, (iy + 1 == N ? 0u : iy + 1) * n,
                           (iy + 2 >= N ? iy + 2 - N : iy + 2) * n};
    Rl cl[2][2][4];  // cl[a][b][c]: x offset a, y offset b (0, +1), z offset c-1
    Rl xr[2][2][2];  // xr[s][b][c]: x offset -1 (s=0) or +2 (s=1), y offset b, z offset c
    Rl yr[2][2][2];  // yr[s][a][c]: y offset -1 (s=0) or +2 (s=1), x offset a, z offset c
    if (iz >= 1 && iz + 2 < N) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                load_run4(mesh_at(phi, X[a + 1] + Y[b + 1] + iz - 1), cl[a][b]);
                load_run2(mesh_at(phi, X[a * 3] + Y[b + 1] + iz), xr[a][b]);
                load_run2(mesh_at(phi, X[b + 1] + Y[a * 3] + iz), yr[a][b]);
            }
    } else {
        const uint32_t Z[4] = {(uint32_t)(iz == 0 ? N - 1 : iz - 1), (uint32_t)iz, (uint32_t)(iz + 1 == N ? 0 : iz + 1),
                               (uint32_t)(iz + 2 >= N ? iz + 2 - N : iz + 2)};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
#pragma unroll
                for (int c = 0; c < 4; ++c) cl[a][b][c] = *mesh_at(phi, X[a + 1] + Y[b + 1] + Z[c]);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    xr[a][b][c] = *mesh_at(phi, X[a * 3] + Y[b + 1] + Z[c + 1]);
                    yr[a][b][c] = *mesh_at(phi, X[b + 1] + Y[a * 3] + Z[c + 1]);
                }
            }
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
                const Rl xm = a == 0 ? xr[0][b][cc] : cl[0][b][cc + 1];
                const Rl xp = a == 0 ? cl[1][b][cc + 1] : xr[1][b][cc];
                gx[a][b][cc] = (xm - xp) * h;
                const Rl ym = b == 0 ? yr[0][a][cc] : cl[a][0][cc + 1];
                const Rl yp = b == 0 ? cl[a][1][cc + 1] : yr[1][a][cc];
                gy[a][b][cc] = (ym - yp) * h;
                gz[a][b][cc] = (cl[a][b][cc] - cl[a][b][cc + 2]) * h;
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
    uint32_t kk = 0xFFFFFFFFu;  // MODE 2: this particle's next tile key
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
                kk = (uint32_t)((((int)x / ts) * nt + (int)y / ts) * nt + (int)z / ts);
                key[i] = kk;
            }
        }
    }
    if constexpr (MODE == 2) {
        // Runs of equal keys in consecutive lanes (particles in a coherent
        // order share tiles) are counted once: the first lane of each run
        // adds the run length.
        const uint32_t lane = __lane_id();
        const uint32_t prev = __shfl_up(kk, 1);
        const bool head = kk != 0xFFFFFFFFu && (lane == 0 || prev != kk);
        const unsigned long long bounds = __ballot(head || kk == 0xFFFFFFFFu);
        if (head) {
            const unsigned long long above = lane == 63 ? 0ull : (bounds >> (lane + 1)) << (lane + 1);
            const uint32_t run = (above ? (uint32_t)__ffsll((long long)above) - 1u : 64u) - lane;
            uint32_t s = (kk * 2654435761u) >> 25;  // 7 bits
            bool done = false;
            for (int probe = 0; probe < HS; ++probe) {
                const uint32_t old = atomicCAS(&hkey[s], 0xFFFFFFFFu, kk);
                if (old == 0xFFFFFFFFu || old == kk) {
                    atomicAdd(&hcnt[s], run);
                    done = true;
                    break;
                }
                s = (s + 1) & (HS - 1);
            }
            if (!done) atomicAdd(&cnt[kk], run);
        }
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
