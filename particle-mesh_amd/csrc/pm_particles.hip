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

// Shape of the staging interpolation (MODE 2): 256 threads with two
// particles each, whose loads and gathers are in flight together (measured
// against one particle per thread and against 512 / 1024 threads), and an
// LDS hash of 128 tiles per pass.
#ifndef PM_STAGE_PER
#define PM_STAGE_PER 2
#endif
constexpr int kStageThreads = 256;
constexpr int kStagePer = PM_STAGE_PER;
constexpr int kStageHash = 128;

// One particle of k_interp: returns its next tile (kNo for none) and, in
// MODE 2, its staged record.
template <typename Rl, int MODE, typename Rec>
__device__ __forceinline__ uint32_t interp_one(int64_t i, double* __restrict__ px, double* __restrict__ py,
                                               double* __restrict__ pz, double* __restrict__ vx,
                                               double* __restrict__ vy, double* __restrict__ vz, int ng,
                                               const Rl* __restrict__ phi, double kick, double drift, int ts,
                                               unsigned long long* __restrict__ err, Rec& rec) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    const double box = (double)ng;
    double x = px[i], y = py[i], z = pz[i];
    if (!(in_box(x, box) && in_box(y, box) && in_box(z, box))) {
        atomicAdd(err, 1ULL);
        return kNo;
    }
    double F[3];
    cic_force(phi, ng, x, y, z, F);
    if constexpr (MODE == 1) {
        vx[i] = F[0];
        vy[i] = F[1];
        vz[i] = F[2];
        return kNo;
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
        if constexpr (MODE != 2) {
            return kNo;
        } else {
            // a particle that drifted more than a box away stays unstaged and
            // is counted, as the next step's range check would
            if (!(in_box(x, box) && in_box(y, box) && in_box(z, box))) {
                atomicAdd(err, 1ULL);
                return kNo;
            }
            const int lt = __builtin_ctz((unsigned)ts), nt = ng >> lt;  // ts: power of two
            // in-tile offsets: u32 at 2^-28 cells (FP32 mesh) or exact doubles
            const int msk = ~(ts - 1);  // tile origin = cell with the low bits cleared
            const double lx = x - ((int)x & msk), ly = y - ((int)y & msk), lz = z - ((int)z & msk);
            if constexpr (sizeof(Rl) == 4)
                rec = StageQ{(uint32_t)(lx * 268435456.0), (uint32_t)(ly * 268435456.0),
                             (uint32_t)(lz * 268435456.0)};
            else
                rec = StageD{lx, ly, lz};
            return (uint32_t)((((int)x >> lt) * nt + ((int)y >> lt)) * nt + ((int)z >> lt));
        }
    }
}

// update_particles (utils.cpp:141-170) with the force of cic_force, then the
// drift that the reference leaves commented out (utils.cpp:145-146), for
// PER particles per thread (NT apart: their loads are in flight together).
// MODE 0: v += g*kick, x += v*drift.  MODE 1: write g to (vx, vy, vz).
// MODE 2: as MODE 0, then the drifted position is staged for the next
// step's deposit (stage_commit).
template <typename Rl, int MODE, int NT, int PER>
__global__ __launch_bounds__(NT) void k_interp(double* __restrict__ px, double* __restrict__ py,
                                               double* __restrict__ pz, double* __restrict__ vx,
                                               double* __restrict__ vy, double* __restrict__ vz, int64_t np,
                                               int ng, const Rl* __restrict__ phi, double kick, double drift, int ts,
                                               uint32_t* __restrict__ fill, const uint32_t* __restrict__ coff,
                                               uint32_t* __restrict__ spill, void* __restrict__ stage,
                                               unsigned long long* __restrict__ err) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    using Rec = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    constexpr int HS = MODE == 2 ? kStageHash : 1;
    static_assert((HS & (HS - 1)) == 0, "hash size must be a power of two");
    __shared__ uint32_t hkey[HS], hcnt[HS], hbase[HS];
    const int64_t i0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * (NT * PER) + threadIdx.x;
    uint32_t kk[PER];
    Rec rec[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int64_t i = i0 + (int64_t)q * NT;
        kk[q] = i < np ? interp_one<Rl, MODE>(i, px, py, pz, vx, vy, vz, ng, phi, kick, drift, ts, err, rec[q]) : kNo;
    }
    if constexpr (MODE == 2) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            for (int s = threadIdx.x; s < HS; s += NT) {
                hkey[s] = kNo;
                hcnt[s] = 0u;
            }
            __syncthreads();
            stage_commit<HS, NT>(kk[q], rec[q], hkey, hcnt, hbase, fill, coff, spill, static_cast<Rec*>(stage));
            __syncthreads();
        }
    }
}

int launch_interp(int prec, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                  int64_t np, int ng, const void* phi, double kick, double drift, int mode,
                  unsigned long long* err, hipStream_t st, SortedScratch* s) {
    if (np <= 0) return PM_OK;
    if (mode == 2 && (!s || !s->stage)) return fail(PM_ERR_ARG, "interp: staging buffers missing");
    const int nt = mode == 2 ? kStageThreads : kPartThreads, per = mode == 2 ? kStagePer : 1;
    const unsigned nb = (unsigned)((np + (int64_t)nt * per - 1) / ((int64_t)nt * per));
    const dim3 g(nb), b(nt);
    const int ts = sorted_tile(ng);
    uint32_t* fill = s ? s->fill : nullptr;
    const uint32_t* coff = s ? s->coff : nullptr;
    uint32_t* spill = s ? s->spill : nullptr;
    void* stage = s ? s->stage : nullptr;
#define PM_INTERP(RL, M, NTV, PERV)                                                                          \
    hipLaunchKernelGGL((k_interp<RL, M, NTV, PERV>), g, b, 0, st, px, py, pz, vx, vy, vz, np, ng,          \
                       (const RL*)phi, kick, drift, ts, fill, coff, spill, stage, err)
    if (prec == PM_FP32) {
        if (mode == 0)
            PM_INTERP(float, 0, kPartThreads, 1);
        else if (mode == 1)
            PM_INTERP(float, 1, kPartThreads, 1);
        else
            PM_INTERP(float, 2, kStageThreads, kStagePer);
    } else {
        if (mode == 0)
            PM_INTERP(double, 0, kPartThreads, 1);
        else if (mode == 1)
            PM_INTERP(double, 1, kPartThreads, 1);
        else
            PM_INTERP(double, 2, kStageThreads, kStagePer);
    }
#undef PM_INTERP
    return check_launch("interp");
}

}  // namespace pm
