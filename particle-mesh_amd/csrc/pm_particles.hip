// Particle kernels of the fused one-box step: CIC deposit (atomic variant)
// and interpolation of the centred-difference force fused with kick and
// drift.  Positions and velocities are FP64; the mesh is FP32 or FP64.
#include <type_traits>

#include "pm_gather.h"
#include "pm_internal.h"
#include "pm_tiles.h"

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
    if (!in_box3(x, y, z, (double)ng)) {
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

// Shape of the staging interpolation (MODE 2): 512 threads (FP32 mesh) or
// 256 (FP64 mesh) with PER particles each and an LDS hash of 128 tiles per
// pass.  Interleaved A/B at 512^3 with the band order, 256 / 512 / 1024
// threads: 3.40 / 3.26 / 3.95 ms (FP32 mesh); 128 / 256 / 512: 4.88 / 4.22 /
// 4.32 ms (FP64 mesh).  One particle per thread for both meshes: with two,
// the gathers of both are in flight together but the registers they take
// halve the waves per SIMD (FP64: 5.02 against 4.26 ms; FP32 3.27 against
// 3.40, no better than 512 threads), and capping the FP64 kernel at 96
// VGPRs for a fifth wave spills (5.31 ms).
#ifndef PM_STAGE_PER32
#define PM_STAGE_PER32 1
#endif
#ifndef PM_STAGE_PER64
#define PM_STAGE_PER64 1
#endif
#ifndef PM_STAGE_THREADS32
#define PM_STAGE_THREADS32 512
#endif
#ifndef PM_STAGE_THREADS64
#define PM_STAGE_THREADS64 256
#endif
constexpr int kStageThreads32 = PM_STAGE_THREADS32, kStageThreads64 = PM_STAGE_THREADS64;
constexpr int kStagePer32 = PM_STAGE_PER32, kStagePer64 = PM_STAGE_PER64;
constexpr int kStageHash = 128;

// update_particles (utils.cpp:141-170) with the force of cic_force, then the
// drift that the reference leaves commented out (utils.cpp:145-146), for
// PER particles per thread, NT apart.  The kernel runs in three phases so
// that each phase has all its loads in flight together: (1) the positions
// and velocities of the thread's particles, (2) their 12 potential runs
// each, (3) force, kick, drift and stores.  A lane past the end or with a
// position outside the box (counted in err) gathers around cell 0 and
// stores nothing.
// MODE 0: v += g*kick, x += v*drift.  MODE 1: write g to (vx, vy, vz).
// MODE 2: as MODE 0, then the drifted position is staged for the next
// step's deposit (stage_pass).
#ifndef PM_INTERP_MINW
#define PM_INTERP_MINW 1
#endif
template <typename Rl, int MODE, int NT, int PER>
__global__ __launch_bounds__(NT, PM_INTERP_MINW) void k_interp(double* __restrict__ px, double* __restrict__ py,
                                               double* __restrict__ pz, double* __restrict__ vx,
                                               double* __restrict__ vy, double* __restrict__ vz, int64_t np,
                                               int ng, const Rl* __restrict__ phi, double kick, double drift, int ts,
                                               uint32_t* __restrict__ fill, const uint32_t* __restrict__ coff,
                                               uint32_t* __restrict__ spill, void* __restrict__ stage,
                                               unsigned long long* __restrict__ err, BandMap bm) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    using Rec = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    constexpr int HS = MODE == 2 ? kStageHash : 1;
    static_assert((HS & (HS - 1)) == 0, "hash size must be a power of two");
    __shared__ uint32_t hkey[HS], hcnt[HS], hbase[HS];
    const int64_t i0 = (int64_t)band_remap(blockIdx.x, gridDim.x, bm) * (NT * PER) + threadIdx.x;
    const double box = (double)ng;
    int64_t idx[PER];
    bool live[PER], ok[PER];
    double x[PER], y[PER], z[PER], u[PER][3];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        idx[q] = i0 + (int64_t)q * NT;
        live[q] = idx[q] < np;
        const int64_t j = live[q] ? idx[q] : 0;
        x[q] = stream_load(px + j);
        y[q] = stream_load(py + j);
        z[q] = stream_load(pz + j);
        if constexpr (MODE != 1) {
            u[q][0] = stream_load(vx + j);
            u[q][1] = stream_load(vy + j);
            u[q][2] = stream_load(vz + j);
        }
    }
    CellIdx<Rl> c[PER];
    Rl cl[PER][2][2][4], xr[PER][2][2][2], yr[PER][2][2][2];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        ok[q] = live[q] & in_box3(x[q], y[q], z[q], box);
        c[q] = cell_of<Rl>(ok[q] ? x[q] : 0.0, ok[q] ? y[q] : 0.0, ok[q] ? z[q] : 0.0, ng);
        gather_runs_main(phi, ng, c[q], cl[q], xr[q], yr[q]);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) gather_runs_seam(phi, ng, c[q], cl[q], xr[q], yr[q]);
    uint32_t kk[PER];
    Rec rec[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        kk[q] = kNo;
        Rl f[3];
        force_from_runs(cl[q], xr[q], yr[q], c[q].tx, c[q].ty, c[q].tz, f);
        const int64_t i = idx[q];
        if (!ok[q]) {
            if (live[q]) atomicAdd(err, 1ULL);
            continue;
        }
        if constexpr (MODE == 1) {
            vx[i] = (double)f[0];
            vy[i] = (double)f[1];
            vz[i] = (double)f[2];
        } else {
            const double nvx = u[q][0] + (double)f[0] * kick, nvy = u[q][1] + (double)f[1] * kick,
                         nvz = u[q][2] + (double)f[2] * kick;
            stream_store(vx + i, nvx);
            stream_store(vy + i, nvy);
            stream_store(vz + i, nvz);
            double xn = x[q], yn = y[q], zn = z[q];
            if (drift != 0.0) {
                xn = wrap_pos(xn + nvx * drift, box);
                yn = wrap_pos(yn + nvy * drift, box);
                zn = wrap_pos(zn + nvz * drift, box);
                stream_store(px + i, xn);
                stream_store(py + i, yn);
                stream_store(pz + i, zn);
            }
            if constexpr (MODE == 2) {
                // a particle that drifted more than a box away stays unstaged
                // and is counted, as the next step's range check would
                if (!in_box3(xn, yn, zn, box)) {
                    atomicAdd(err, 1ULL);
                    continue;
                }
                rec[q] = tile_record<Rec>(xn, yn, zn, ts);
                kk[q] = tile_key(xn, yn, zn, ng, ts);
            }
        }
    }
    if constexpr (MODE == 2) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int64_t pos = stage_pass<HS, NT>(kk[q], hkey, hcnt, hbase, fill, coff, spill);
            if (pos >= 0) static_cast<Rec*>(stage)[pos] = rec[q];
        }
    }
}

int launch_interp(int prec, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                  int64_t np, int ng, const void* phi, double kick, double drift, int mode,
                  unsigned long long* err, hipStream_t st, SortedScratch* s) {
    if (np <= 0) return PM_OK;
    if (mode == 2 && (!s || !s->stage)) return fail(PM_ERR_ARG, "interp: staging buffers missing");
    const int nt = mode != 2 ? kPartThreads : prec == PM_FP32 ? kStageThreads32 : kStageThreads64,
              per = mode != 2 ? 1 : prec == PM_FP32 ? kStagePer32 : kStagePer64;
    const unsigned nb = (unsigned)((np + (int64_t)nt * per - 1) / ((int64_t)nt * per));
    const dim3 g(nb), b(nt);
    const BandMap bm = band_map(np, nb, nt * per);
    const int ts = sorted_tile(ng);
    uint32_t* fill = s ? s->fill : nullptr;
    const uint32_t* coff = s ? s->coff : nullptr;
    uint32_t* spill = s ? s->spill : nullptr;
    void* stage = s ? s->stage : nullptr;
#define PM_INTERP(RL, M, NTV, PERV)                                                                          \
    hipLaunchKernelGGL((k_interp<RL, M, NTV, PERV>), g, b, 0, st, px, py, pz, vx, vy, vz, np, ng,          \
                       (const RL*)phi, kick, drift, ts, fill, coff, spill, stage, err, bm)
    if (prec == PM_FP32) {
        if (mode == 0)
            PM_INTERP(float, 0, kPartThreads, 1);
        else if (mode == 1)
            PM_INTERP(float, 1, kPartThreads, 1);
        else
            PM_INTERP(float, 2, kStageThreads32, kStagePer32);
    } else {
        if (mode == 0)
            PM_INTERP(double, 0, kPartThreads, 1);
        else if (mode == 1)
            PM_INTERP(double, 1, kPartThreads, 1);
        else
            PM_INTERP(double, 2, kStageThreads64, kStagePer64);
    }
#undef PM_INTERP
    return check_launch("interp");
}

}  // namespace pm
