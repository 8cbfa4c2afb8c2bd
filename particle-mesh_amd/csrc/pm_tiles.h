// Tile shells of the sorted deposit, shared by the deposit kernels and the
// FFT pass that can fold them while it reads the density.
#pragma once

#include "pm_common.h"

namespace pm {

template <int TS> __host__ __device__ constexpr int shell_size() {
    return (TS + 1) * (TS + 1) * (TS + 1) - TS * TS * TS;
}

// Shell index of a node with local coordinates (a, b, c), max == TS:
// face x (a == TS): (TS+1)^2 entries, face y (b == TS, a < TS): TS (TS+1),
// face z (c == TS, a, b < TS): TS^2.
template <int TS> __device__ __forceinline__ int shell_index(int a, int b, int c) {
    constexpr int E = TS + 1;
    if (a == TS) return b * E + c;
    if (b == TS) return E * E + a * E + c;
    return E * E + TS * E + a * TS + b;
}

// The periodic ghost fold (reorder_slab.cpp:43-67) at tile granularity for
// one mesh node (x, y, z) on a lower face of its tile: v (the tile's own
// sum) plus the shells of the lower neighbour tiles that hold the node, in
// the fixed order d = 1..7 (bit 0: x-neighbour, 1: y, 2: z).  on: which of
// a == 0, b == 0, c == 0 hold (same bits).
template <typename Rl, int TS>
__device__ __forceinline__ Rl fold_node(Rl v, const Rl* __restrict__ shell, int ng, int x, int y, int z, int on) {
    constexpr int SH = shell_size<TS>();
    const int nt = ng / TS;
    const int tx = x / TS, ty = y / TS, tz = z / TS;
    const int a = x - tx * TS, b = y - ty * TS, c = z - tz * TS;
    for (int d = 1; d < 8; ++d) {
        if ((d & on) != d) continue;
        const int dx = d & 1, dy = (d >> 1) & 1, dz = (d >> 2) & 1;
        const int sx = tx - dx < 0 ? nt - 1 : tx - dx;
        const int sy = ty - dy < 0 ? nt - 1 : ty - dy;
        const int sz = tz - dz < 0 ? nt - 1 : tz - dz;
        const size_t st = ((size_t)sx * nt + sy) * nt + sz;
        v += shell[st * SH + shell_index<TS>(a + TS * dx, b + TS * dy, c + TS * dz)];
    }
    return v;
}

constexpr int kDepBatch = 4;  // particles in flight per thread in the tile kernels

// CIC weights (cic_deposit, utils_inline.h:25-49) of one particle in local
// cell (a, b, c) with fractional offsets (fx, fy, fz), added in 2^-32 fixed
// point to the (TS+1)^3 LDS node tile.
template <int TS>
__device__ __forceinline__ void tile_cic_add(unsigned long long* acc, int a, int b, int c, double fx, double fy,
                                             double fz, double fix) {
    constexpr int E = TS + 1;
    const double sx0 = 1.0 - fx, sy0 = 1.0 - fy;
    const double sxy00 = sx0 * sy0, sxy10 = fx * sy0, sxy01 = sx0 * fy, sxy11 = fx * fy;
    const double m0 = fix * (1.0 - fz), m1 = fix * fz;
    unsigned long long* n0 = acc + (a * E + b) * E + c;
    atomicAdd(n0, (unsigned long long)__double2ll_rn(sxy00 * m0));
    atomicAdd(n0 + E * E, (unsigned long long)__double2ll_rn(sxy10 * m0));
    atomicAdd(n0 + E, (unsigned long long)__double2ll_rn(sxy01 * m0));
    atomicAdd(n0 + E * E + E, (unsigned long long)__double2ll_rn(sxy11 * m0));
    atomicAdd(n0 + 1, (unsigned long long)__double2ll_rn(sxy00 * m1));
    atomicAdd(n0 + E * E + 1, (unsigned long long)__double2ll_rn(sxy10 * m1));
    atomicAdd(n0 + E + 1, (unsigned long long)__double2ll_rn(sxy01 * m1));
    atomicAdd(n0 + E * E + E + 1, (unsigned long long)__double2ll_rn(sxy11 * m1));
}

// The TS^3 owned nodes of tile t go to the mesh, the upper shell to the
// tile's shell buffer (after a barrier on the accumulation).
template <typename Rl, int TS>
__device__ __forceinline__ void tile_write(const unsigned long long* acc, uint32_t t, int ox, int oy, int oz,
                                           int ng, Rl* __restrict__ rho, Rl* __restrict__ shell) {
    constexpr int E = TS + 1, SH = shell_size<TS>();
    const double unfix = 1.0 / 4294967296.0;
    const size_t N = (size_t)ng;
    for (int i = threadIdx.x; i < TS * TS * TS; i += 256) {
        const int a = i / (TS * TS), b = (i / TS) % TS, c = i % TS;
        rho[((size_t)(ox + a) * N + (oy + b)) * N + (oz + c)] = (Rl)((double)acc[(a * E + b) * E + c] * unfix);
    }
    Rl* sh = shell + (size_t)t * SH;
    for (int i = threadIdx.x; i < SH; i += 256) {
        int a, b, c;
        if (i < E * E) {
            a = TS;
            b = i / E;
            c = i % E;
        } else if (i < E * E + TS * E) {
            const int j = i - E * E;
            a = j / E;
            b = TS;
            c = j % E;
        } else {
            const int j = i - E * E - TS * E;
            a = j / TS;
            b = j % TS;
            c = TS;
        }
        sh[i] = (Rl)((double)acc[(a * E + b) * E + c] * unfix);
    }
}

// Slot reservation in the staging buckets and the write of a staged record
// (called by every thread of the workgroup; kk = kNo for no particle).
// Runs of equal tiles in consecutive lanes (particles in a coherent order
// share tiles) reserve their slots together: the first lane of a run takes
// run-length slots, first within the workgroup (LDS hash of HS slots), then
// one global fill atomic per distinct tile; a slot past the bucket's
// capacity raises *spill instead of writing.
template <int HS, int NT, typename Rec>
__device__ __forceinline__ void stage_commit(uint32_t kk, const Rec& rec, uint32_t* hkey, uint32_t* hcnt,
                                             uint32_t* hbase, uint32_t* __restrict__ fill,
                                             const uint32_t* __restrict__ coff, uint32_t* __restrict__ spill,
                                             Rec* __restrict__ stage) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    const uint32_t prev = __shfl_up(kk, 1);
    const bool head = kk != kNo && (lane == 0 || prev != kk);
    const unsigned long long bounds = __ballot(head || kk == kNo);
    uint32_t hs = kNo, lrank = 0u, gb = 0u;
    if (head) {
        const unsigned long long above = lane == 63 ? 0ull : (bounds >> (lane + 1)) << (lane + 1);
        const uint32_t run = (above ? (uint32_t)__ffsll((long long)above) - 1u : 64u) - lane;
        uint32_t s = (kk * 2654435761u) >> (32 - ilog2(HS));
        for (int probe = 0; probe < HS; ++probe) {
            const uint32_t old = atomicCAS(&hkey[s], kNo, kk);
            if (old == kNo || old == kk) {
                hs = s;
                lrank = atomicAdd(&hcnt[s], run);
                break;
            }
            s = (s + 1) & (HS - 1);
        }
        if (hs == kNo) gb = atomicAdd(&fill[kk], run);  // hash full: reserve directly
    }
    __syncthreads();
    for (int s = threadIdx.x; s < HS; s += NT)
        if (hkey[s] != kNo) hbase[s] = atomicAdd(&fill[hkey[s]], hcnt[s]);
    __syncthreads();
    if (hs != kNo) gb = hbase[hs] + lrank;
    // every lane of a run takes its head's base plus its distance
    const unsigned long long upto = bounds & (lane == 63 ? ~0ull : (2ull << lane) - 1ull);
    const int hl = upto ? 63 - __clzll((long long)upto) : 0;
    const uint32_t base = __shfl(gb, hl);
    if (kk != kNo) {
        const uint32_t slot = base + (lane - (uint32_t)hl);
        const uint32_t c0 = coff[kk];
        if (slot < coff[kk + 1] - c0)
            stage[(size_t)c0 + slot] = rec;
        else
            *spill = 1u;
    }
}

}  // namespace pm
