// Split-precision particle state of the fused step with an FP32 mesh:
// positions as an integer cell plus a fixed-point fraction in one u32 per
// axis, velocities in FP32, resident in the plan across steps.
//
// A position x in [0, ng) is held as q = x * 2^S, S = 32 - log2(ng): the
// cell is q >> S and the fraction the low S bits (2^-23 cells at ng = 512),
// so the whole periodic box is the u32 range and a drift is an integer add
// that wraps at the box edge by itself.  The state takes 24 bytes per
// particle instead of the 48 of FP64 positions and velocities, which is what
// the interpolation streams twice per step.  The deposit reads the same
// in-tile records as the FP64 path (StageQ), formed from q exactly.
#include "pm_gather.h"
#include "pm_internal.h"
#include "pm_tiles.h"

namespace pm {

__host__ __device__ inline int split_shift(int ng) { return 32 - ilog2(ng); }

// element a (axis) of particle i in a state array: three planes of np
// (a particle's triple as one 12-byte access measured slower: 2.39 -> 2.43
// ms for the 512^3 interpolation)
__device__ __forceinline__ int64_t sidx(int64_t i, int a, int64_t np) { return (int64_t)a * np + i; }

// tile of q: sh = S + log2(ts) = 32 - log2(tiles per axis), 32 for one tile
__device__ __forceinline__ uint32_t split_key(uint32_t qx, uint32_t qy, uint32_t qz, int sh, uint32_t nt) {
    const auto t = [sh](uint32_t q) { return (uint32_t)((uint64_t)q >> sh); };
    return (t(qx) * nt + t(qy)) * nt + t(qz);
}

// In-tile record of one axis: cell in the tile << 28 | fraction at 2^-28
// cells (truncated for S > 28), the StageQ layout of tile_record.
__device__ __forceinline__ uint32_t split_rec(uint32_t q, int S, uint32_t tmask) {
    const uint32_t cell = (q >> S) & tmask;
    const uint32_t frac = q & ((1u << S) - 1u);
    return (cell << 28) | (S >= 28 ? frac >> (S - 28) : frac << (28 - S));
}

// FP64 position -> u32 fixed point, nearest (x * 2^S is exact).  A position
// outside [0, ng) (or not finite) is counted as a range error and wrapped
// into the box (0 if not finite).
__device__ __forceinline__ uint32_t split_pos(double x, double box, double scale, bool& bad) {
    if (!(x >= 0.0 && x < box)) {
        bad = true;
        x = isfinite(x) ? x - floor(x / box) * box : 0.0;
        if (!(x >= 0.0 && x < box)) x = 0.0;
    }
    return (uint32_t)(unsigned long long)llrint(x * scale);  // 2^32 wraps to 0: the box is periodic
}

__global__ __launch_bounds__(256) void k_split_load(const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ vx,
                                                    const double* __restrict__ vy, const double* __restrict__ vz,
                                                    int64_t np, int ng, uint32_t* __restrict__ q,
                                                    float* __restrict__ u, unsigned long long* __restrict__ err) {
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= np) return;
    const double box = (double)ng, scale = ldexp(1.0, split_shift(ng));
    bool bad = false;
    q[sidx(i, 0, np)] = split_pos(x[i], box, scale, bad);
    q[sidx(i, 1, np)] = split_pos(y[i], box, scale, bad);
    q[sidx(i, 2, np)] = split_pos(z[i], box, scale, bad);
    u[sidx(i, 0, np)] = (float)vx[i];
    u[sidx(i, 1, np)] = (float)vy[i];
    u[sidx(i, 2, np)] = (float)vz[i];
    if (bad) atomicAdd(err, 1ULL);
}

__global__ __launch_bounds__(256) void k_split_store(const uint32_t* __restrict__ q, const float* __restrict__ u,
                                                     int64_t np, int ng, double* __restrict__ x,
                                                     double* __restrict__ y, double* __restrict__ z,
                                                     double* __restrict__ vx, double* __restrict__ vy,
                                                     double* __restrict__ vz) {
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= np) return;
    const double inv = ldexp(1.0, -split_shift(ng));
    x[i] = (double)q[sidx(i, 0, np)] * inv;
    y[i] = (double)q[sidx(i, 1, np)] * inv;
    z[i] = (double)q[sidx(i, 2, np)] * inv;
    vx[i] = (double)u[sidx(i, 0, np)];
    vy[i] = (double)u[sidx(i, 1, np)];
    vz[i] = (double)u[sidx(i, 2, np)];
}

// Restaging (the first step after a load, or a step after a bucket
// overflow; every kernel a no-op unless *gate != 0).  The gate is copied
// from the overflow flag first, because the capacity scan clears that flag.
__global__ void k_split_gate(const uint32_t* __restrict__ spill, uint32_t* __restrict__ gate) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *gate = *spill;
}

// tile counts: runs of equal tiles in consecutive lanes add once
__global__ __launch_bounds__(256) void k_split_count(const uint32_t* __restrict__ q, int64_t np, int ng, int ts,
                                                     uint32_t* __restrict__ cnt, const uint32_t* __restrict__ gate) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    if (*gate == 0u) return;
    const int lt = __builtin_ctz((unsigned)ts), sh = split_shift(ng) + lt;
    const uint32_t nt = (uint32_t)(ng >> lt);
    const uint32_t lane = __lane_id();
    // grid-stride over blocks of 256 (a capped grid: the gated no-op launch
    // of a normal step stays cheap)
    for (int64_t i0 = (int64_t)blockIdx.x * 256; i0 < np; i0 += (int64_t)gridDim.x * 256) {
        const int64_t i = i0 + threadIdx.x;
        const uint32_t kk =
            i < np ? split_key(q[sidx(i, 0, np)], q[sidx(i, 1, np)], q[sidx(i, 2, np)], sh, nt) : kNo;
        const uint32_t prev = __shfl_up(kk, 1);
        const bool head = kk != kNo && (lane == 0 || prev != kk);
        const unsigned long long bounds = __ballot(head || kk == kNo);
        if (head) {
            const unsigned long long above = lane == 63 ? 0ull : (bounds >> (lane + 1)) << (lane + 1);
            const uint32_t run = (above ? (uint32_t)__ffsll((long long)above) - 1u : 64u) - lane;
            atomicAdd(&cnt[kk], run);
        }
    }
}

constexpr int kSplitThreads = 512, kSplitHash = 128;

// records of the current positions into the buckets (after the capacity scan)
__global__ __launch_bounds__(kSplitThreads) void k_split_stage(const uint32_t* __restrict__ q, int64_t np, int ng,
                                                               int ts, uint32_t* __restrict__ fill,
                                                               const uint32_t* __restrict__ coff,
                                                               uint32_t* __restrict__ spill,
                                                               StageQ* __restrict__ stage,
                                                               const uint32_t* __restrict__ gate) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    __shared__ uint32_t hkey[kSplitHash], hcnt[kSplitHash], hbase[kSplitHash];
    if (*gate == 0u) return;
    const int S = split_shift(ng), lt = __builtin_ctz((unsigned)ts);
    const uint32_t nt = (uint32_t)(ng >> lt), tmask = (uint32_t)ts - 1u;
    // grid-stride over blocks (capped grid, as in k_split_count); the loop
    // bound is uniform in the workgroup, as the reservation's barriers need
    for (int64_t i0 = (int64_t)blockIdx.x * kSplitThreads; i0 < np; i0 += (int64_t)gridDim.x * kSplitThreads) {
        const int64_t i = i0 + threadIdx.x;
        uint32_t kk = kNo;
        StageQ rec{};
        if (i < np) {
            const uint32_t a = q[sidx(i, 0, np)], b = q[sidx(i, 1, np)], c = q[sidx(i, 2, np)];
            kk = split_key(a, b, c, S + lt, nt);
            rec = StageQ{split_rec(a, S, tmask), split_rec(b, S, tmask), split_rec(c, S, tmask)};
        }
        const int64_t pos = stage_pass<kSplitHash, kSplitThreads>(kk, hkey, hcnt, hbase, fill, coff, spill);
        if (pos >= 0) stage[pos] = rec;
    }
}

// update_particles (utils.cpp:141-170) on the split state: the force of
// cic_force at the cell and fraction of q (pm_gather.h), the kick in FP32,
// then the drift as the integer add of the displacement rounded to 2^-S
// cells (wrapping at the box edge), and the drifted position's record
// staged for the next deposit, as the FP64 staging interpolation does.  A
// kick that leaves a velocity non-finite is counted as a range error and
// that particle keeps its velocity and position.
template <bool DRIFT>
__global__ __launch_bounds__(kSplitThreads) void k_split_interp(
    uint32_t* __restrict__ q, float* __restrict__ u, int64_t np, int ng, const float* __restrict__ phi, float kick,
    float dscale, int ts, uint32_t* __restrict__ fill, const uint32_t* __restrict__ coff,
    uint32_t* __restrict__ spill, StageQ* __restrict__ stage, unsigned long long* __restrict__ err, BandMap bm) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    __shared__ uint32_t hkey[kSplitHash], hcnt[kSplitHash], hbase[kSplitHash];
    const int64_t i = (int64_t)band_remap(blockIdx.x, gridDim.x, bm) * kSplitThreads + threadIdx.x;
    const bool live = i < np;
    const int64_t j = live ? i : 0;
    const uint32_t a = stream_load(q + sidx(j, 0, np)), b = stream_load(q + sidx(j, 1, np)),
                   c = stream_load(q + sidx(j, 2, np));
    const float u0 = stream_load(u + sidx(j, 0, np)), u1 = stream_load(u + sidx(j, 1, np)),
                u2 = stream_load(u + sidx(j, 2, np));
    const int S = split_shift(ng), lt = __builtin_ctz((unsigned)ts);
    const uint32_t fmask = (1u << S) - 1u;
    const float fs = ldexpf(1.0f, -S);
    const CellIdx<float> cc = cell_at<float>((int)(a >> S), (int)(b >> S), (int)(c >> S), (float)(a & fmask) * fs,
                                             (float)(b & fmask) * fs, (float)(c & fmask) * fs, ng);
    float cl[2][2][4], xr[2][2][2], yr[2][2][2];
    gather_runs_main(phi, ng, cc, cl, xr, yr);
    // the shorter seam reload (the full one: 2.39 -> 2.48 ms at 512^3)
    seam_reload_missing(phi, ng, cc, cl, xr, yr);
    float f[3];
    force_from_runs(cl, xr, yr, cc.tx, cc.ty, cc.tz, f);
    float n0 = fmaf(f[0], kick, u0), n1 = fmaf(f[1], kick, u1), n2 = fmaf(f[2], kick, u2);
    const bool fin = isfinite(n0) && isfinite(n1) && isfinite(n2);
    if (!fin) {
        n0 = u0;
        n1 = u1;
        n2 = u2;
    }
    uint32_t na = a, nb = b, nc = c;
    if (DRIFT && fin) {
        na = a + (uint32_t)__float2int_rn(n0 * dscale);
        nb = b + (uint32_t)__float2int_rn(n1 * dscale);
        nc = c + (uint32_t)__float2int_rn(n2 * dscale);
    }
    uint32_t kk = kNo;
    StageQ rec{};
    if (live) {
        stream_store(u + sidx(i, 0, np), n0);
        stream_store(u + sidx(i, 1, np), n1);
        stream_store(u + sidx(i, 2, np), n2);
        if (DRIFT) {
            stream_store(q + sidx(i, 0, np), na);
            stream_store(q + sidx(i, 1, np), nb);
            stream_store(q + sidx(i, 2, np), nc);
        }
        if (!fin) atomicAdd(err, 1ULL);
        const uint32_t nt = (uint32_t)(ng >> lt), tmask = (uint32_t)ts - 1u;
        kk = split_key(na, nb, nc, S + lt, nt);
        rec = StageQ{split_rec(na, S, tmask), split_rec(nb, S, tmask), split_rec(nc, S, tmask)};
    }
    const int64_t pos = stage_pass<kSplitHash, kSplitThreads>(kk, hkey, hcnt, hbase, fill, coff, spill);
    if (pos >= 0) stage[pos] = rec;
}

// ================================================================ launchers
int launch_split_load(const double* x, const double* y, const double* z, const double* vx, const double* vy,
                      const double* vz, int64_t np, int ng, SplitState& st, unsigned long long* err,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_split_load, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, x, y, z, vx, vy, vz, np,
                       ng, st.q, st.u, err);
    return check_launch("split_load");
}

int launch_split_store(const SplitState& st, int64_t np, int ng, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, hipStream_t s) {
    hipLaunchKernelGGL(k_split_store, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, st.q, st.u, np, ng, x, y,
                       z, vx, vy, vz);
    return check_launch("split_store");
}

int launch_split_restage(const SplitState& st, int64_t np, int ng, SortedScratch& s, hipStream_t hs) {
    const int ts = sorted_tile(ng);
    hipLaunchKernelGGL(k_split_gate, dim3(1), dim3(64), 0, hs, s.spill, st.gate);
    // the tile counts are zero here: cleared at allocation and by every
    // count-consuming scan
    hipLaunchKernelGGL(k_split_count, dim3((unsigned)std::min<int64_t>((np + 255) / 256, 2048)), dim3(256), 0, hs,
                       st.q, np, ng, ts, s.cnt, st.gate);
    int rc = check_launch("split_count");
    if (!rc) rc = launch_cap_scan_from(ng, s, kCapFromCnt, st.gate, hs);
    if (rc) return rc;
    hipLaunchKernelGGL(k_split_stage, dim3((unsigned)std::min<int64_t>((np + kSplitThreads - 1) / kSplitThreads, 2048)),
                       dim3(kSplitThreads), 0, hs, st.q, np, ng, ts, s.fill, s.coff, s.spill,
                       static_cast<StageQ*>(s.stage), st.gate);
    return check_launch("split_stage");
}

int launch_split_interp(SplitState& st, int64_t np, int ng, const float* phi, double kick, double drift,
                        SortedScratch& s, unsigned long long* err, hipStream_t hs) {
    const unsigned nb = (unsigned)((np + kSplitThreads - 1) / kSplitThreads);
    const BandMap bm = band_map(np, nb, kSplitThreads);
    const float dscale = (float)ldexp(drift, split_shift(ng));
    if (drift != 0.0)
        hipLaunchKernelGGL(k_split_interp<true>, dim3(nb), dim3(kSplitThreads), 0, hs, st.q, st.u, np, ng, phi, (float)kick, dscale, sorted_tile(ng), s.fill,
                           s.coff, s.spill, static_cast<StageQ*>(s.stage), err, bm);
    else
        hipLaunchKernelGGL(k_split_interp<false>, dim3(nb), dim3(kSplitThreads), 0, hs, st.q, st.u, np, ng, phi, (float)kick, dscale, sorted_tile(ng), s.fill,
                           s.coff, s.spill, static_cast<StageQ*>(s.stage), err, bm);
    return check_launch("split_interp");
}

}  // namespace pm
