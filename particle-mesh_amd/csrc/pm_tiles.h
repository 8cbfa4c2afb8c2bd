// Tiles of the sorted deposit (column walk, shells, CIC adds in LDS fixed
// point) and the staging commit of the interpolation.
#pragma once

#include "pm_common.h"
#include "pm_internal.h"

namespace pm {

// Column walk of the sorted deposits.  A workgroup deposits a segment of
// consecutive tiles along z of one (x, y) tile column, accumulating each
// tile's (TS+1)^3 nodes in LDS fixed point.  The upper z plane (c == TS) of
// a tile is the lower plane (c == 0) of the next one, so it is carried into
// the next tile's accumulator instead of being written out: within a
// segment the z part of the ghost fold (reorder_slab.cpp:43-67) is exact
// integer addition.  Written out per tile: the TS^3 owned nodes to the mesh
// and, for c < TS, the upper x and y faces to the tile's xy shell; per
// segment: the upper z plane of its last tile to the segment's z seam.
// The fold (k_fold_faces, k_fold_seams) then adds the lower neighbours' xy
// shells to each tile's lower x and y faces, and the seams to the first plane
// of each segment.
#ifndef PM_WALK_SEG
#define PM_WALK_SEG 8
#endif
constexpr int kWalkSeg = PM_WALK_SEG;  // tiles per segment (a power of two)
__host__ __device__ inline int walk_seg(int nt) { return nt < kWalkSeg ? nt : kWalkSeg; }

// xy shell of a tile: x face (a == TS, b = 0..TS, c < TS) at b*TS + c, then
// the y face (b == TS, a < TS, c < TS) at (TS+1)*TS + a*TS + c.
template <int TS> __host__ __device__ constexpr int xy_shell_size() { return (TS + 1) * TS + TS * TS; }
template <int TS> __device__ __forceinline__ int xy_shell_index(int a, int b, int c) {
    return a == TS ? b * TS + c : (TS + 1) * TS + a * TS + c;
}
// z seam of a segment: the nodes (a, b) = 0..TS of its last tile's plane c == TS.
template <int TS> __host__ __device__ constexpr int zseam_size() { return (TS + 1) * (TS + 1); }

// Position of a workgroup's segment: column (tx, ty), segment sg of nseg.
struct WalkSeg {
    int nt, seg, nseg, tx, ty, sg;
    uint32_t w;
};
__device__ __forceinline__ WalkSeg walk_segment(int ng, int TS) {
    WalkSeg s;
    s.nt = ng / TS;
    s.seg = walk_seg(s.nt);
    s.nseg = s.nt / s.seg;
    s.w = xcd_remap(blockIdx.x, gridDim.x);
    s.tx = (int)(s.w / (uint32_t)(s.nt * s.nseg));
    s.ty = (int)((s.w / (uint32_t)s.nseg) % (uint32_t)s.nt);
    s.sg = (int)(s.w % (uint32_t)s.nseg);
    return s;
}

// End of one tile of the walk (after a barrier on its accumulation): D
// accumulators acc[d * E^3 + node] (signed fixed point, value = acc *
// scale[d]) go out as the owned nodes and the xy shell of tile t, then the
// plane c == TS moves to c == 0 and the rest is cleared for the next tile.
template <typename Rl, int TS, int D, int NT>
__device__ __forceinline__ void walk_tile_out(unsigned long long* acc, const double (&scale)[D], uint32_t t, int ox,
                                              int oy, int oz, int ng, Rl* const (&out)[D],
                                              Rl* const (&shell)[D]) {
    constexpr int E = TS + 1, E3 = E * E * E, XYS = xy_shell_size<TS>();
    const size_t N = (size_t)ng;
    for (int i = threadIdx.x; i < TS * TS * TS; i += NT) {
        const int a = i / (TS * TS), b = (i / TS) % TS, c = i % TS;
        const size_t node = ((size_t)(ox + a) * N + (oy + b)) * N + (oz + c);
        const int n = (a * E + b) * E + c;
#pragma unroll
        for (int d = 0; d < D; ++d) out[d][node] = (Rl)((double)(long long)acc[d * E3 + n] * scale[d]);
    }
    for (int i = threadIdx.x; i < XYS; i += NT) {
        int a, b, c;
        if (i < E * TS) {
            a = TS;
            b = i / TS;
            c = i % TS;
        } else {
            const int j = i - E * TS;
            a = j / TS;
            b = TS;
            c = j % TS;
        }
        const int n = (a * E + b) * E + c;
#pragma unroll
        for (int d = 0; d < D; ++d)
            shell[d][(size_t)t * XYS + i] = (Rl)((double)(long long)acc[d * E3 + n] * scale[d]);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < D * E * E; r += NT) {  // one thread per z row
        unsigned long long* row = acc + (r / (E * E)) * E3 + (r % (E * E)) * E;
        row[0] = row[TS];
#pragma unroll
        for (int c = 1; c <= TS; ++c) row[c] = 0ULL;
    }
    __syncthreads();
}

// End of the segment: the carried plane (now at c == 0) is its z seam.
template <typename Rl, int TS, int D, int NT>
__device__ __forceinline__ void walk_seam_out(const unsigned long long* acc, const double (&scale)[D],
                                              const WalkSeg& s, Rl* const (&shell)[D]) {
    constexpr int E = TS + 1, E3 = E * E * E, XYS = xy_shell_size<TS>(), ZS = zseam_size<TS>();
    const size_t seam0 = (size_t)s.nt * s.nt * s.nt * XYS + (size_t)s.w * ZS;
    for (int r = threadIdx.x; r < D * ZS; r += NT) {
        const int d = r / ZS, n = r % ZS;
        shell[d][seam0 + n] = (Rl)((double)(long long)acc[d * E3 + n * E] * scale[d]);
    }
}

// Column walk with H halo planes above the tile and the node box shifted
// one node down: local node n of a tile sits at tile origin + n - 1, so
// local nodes 0 .. TS+H-1 cover a stencil that reaches one node below a
// tile's cells and up to H-1 above its upper corner nodes.  A tile owns the
// local nodes 0 .. TS-1 on each axis (global origin - 1 .. origin + TS - 2);
// the H upper planes go out as shells (x and y) or are carried up the
// column (z), exactly as the walk above does with H = 1.
// xy shell: x face (a = TS..TS+H-1, b = 0..E-1, c < TS) at
// ((a - TS) E + b) TS + c, then the y face (a < TS, b = TS..TS+H-1, c < TS).
template <int TS, int H> __host__ __device__ constexpr int halo_shell_size() {
    return H * (TS + H) * TS + TS * H * TS;
}
template <int TS, int H> __device__ __forceinline__ int halo_shell_index(int a, int b, int c) {
    constexpr int E = TS + H;
    return a >= TS ? ((a - TS) * E + b) * TS + c : H * E * TS + (a * H + (b - TS)) * TS + c;
}
// z seam of a segment: the carried planes c < H, nodes (a, b) of E x E, at (a E + b) H + c.
template <int TS, int H> __host__ __device__ constexpr int halo_seam_size() { return H * (TS + H) * (TS + H); }

template <typename Rl, int TS, int H, int NT>
__device__ __forceinline__ void walk_halo_tile_out(unsigned long long* acc, double scale, uint32_t t, int gx0,
                                                   int gy0, int gz0, int ng, Rl* __restrict__ out,
                                                   Rl* __restrict__ shell) {
    static_assert(H <= TS, "the carried planes must not overlap");
    constexpr int E = TS + H, XYS = halo_shell_size<TS, H>();
    const size_t N = (size_t)ng;
    const int m = ng - 1;  // ng is a power of two
    for (int i = threadIdx.x; i < TS * TS * TS; i += NT) {
        const int a = i / (TS * TS), b = (i / TS) % TS, c = i % TS;
        const size_t node = ((size_t)((gx0 + a) & m) * N + (size_t)((gy0 + b) & m)) * N + (size_t)((gz0 + c) & m);
        out[node] = (Rl)((double)(long long)acc[(a * E + b) * E + c] * scale);
    }
    for (int i = threadIdx.x; i < XYS; i += NT) {
        int a, b, c;
        if (i < H * E * TS) {
            a = TS + i / (E * TS);
            b = (i / TS) % E;
            c = i % TS;
        } else {
            const int j = i - H * E * TS;
            a = j / (H * TS);
            b = TS + (j / TS) % H;
            c = j % TS;
        }
        shell[(size_t)t * XYS + i] = (Rl)((double)(long long)acc[(a * E + b) * E + c] * scale);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < E * E; r += NT) {  // one thread per z row
        unsigned long long* row = acc + r * E;
#pragma unroll
        for (int c = 0; c < H; ++c) row[c] = row[TS + c];
#pragma unroll
        for (int c = H; c < E; ++c) row[c] = 0ULL;
    }
    __syncthreads();
}

template <typename Rl, int TS, int H, int NT>
__device__ __forceinline__ void walk_halo_seam_out(const unsigned long long* acc, double scale, const WalkSeg& s,
                                                   Rl* __restrict__ shell) {
    constexpr int E = TS + H, XYS = halo_shell_size<TS, H>(), ZS = halo_seam_size<TS, H>();
    const size_t seam0 = (size_t)s.nt * s.nt * s.nt * XYS + (size_t)s.w * ZS;
    for (int r = threadIdx.x; r < ZS; r += NT) {
        const int ab = r / H, c = r % H;
        shell[seam0 + r] = (Rl)((double)(long long)acc[ab * E + c] * scale);
    }
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
    atomicAdd(n0, (unsigned long long)fix_rne(sxy00 * m0));
    atomicAdd(n0 + E * E, (unsigned long long)fix_rne(sxy10 * m0));
    atomicAdd(n0 + E, (unsigned long long)fix_rne(sxy01 * m0));
    atomicAdd(n0 + E * E + E, (unsigned long long)fix_rne(sxy11 * m0));
    atomicAdd(n0 + 1, (unsigned long long)fix_rne(sxy00 * m1));
    atomicAdd(n0 + E * E + 1, (unsigned long long)fix_rne(sxy10 * m1));
    atomicAdd(n0 + E + 1, (unsigned long long)fix_rne(sxy01 * m1));
    atomicAdd(n0 + E * E + E + 1, (unsigned long long)fix_rne(sxy11 * m1));
}

// Slot reservation in the staging buckets (called by every thread of the
// workgroup; kk = kNo for no particle); returns the record's position in the
// bucket arrays, or -1 (no particle, or the bucket is full: then *spill is
// raised and the caller writes nothing).  Runs of equal tiles in consecutive
// lanes (particles in a coherent order share tiles) reserve their slots
// together: the first lane of a run takes run-length slots, first within the
// workgroup (LDS hash of HS slots), then one global fill atomic per distinct
// tile.  The hash must be cleared (hkey = kNo, hcnt = 0) and a barrier
// passed before the call.
template <int HS, int NT>
__device__ __forceinline__ int64_t stage_reserve(uint32_t kk, uint32_t* hkey, uint32_t* hcnt, uint32_t* hbase,
                                                 uint32_t* __restrict__ fill, const uint32_t* __restrict__ coff,
                                                 uint32_t* __restrict__ spill) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    // the bucket bounds are loaded first, so that they arrive during the
    // barriers and the fill atomics (interpolation at 512^3 FP32 in an
    // interleaved A/B: 3.24 -> 3.17 ms)
    uint32_t c0 = 0u, c1 = 0u;
    if (kk != kNo) {
        c0 = coff[kk];
        c1 = coff[kk + 1];
    }
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
    if (kk == kNo) return -1;
    const uint32_t slot = base + (lane - (uint32_t)hl);
    if (slot < c1 - c0) return (int64_t)c0 + slot;
    *spill = 1u;
    return -1;
}

// Reservation of HS-hash passes: clears the hash, reserves (stage_reserve).
template <int HS, int NT>
__device__ __forceinline__ int64_t stage_pass(uint32_t kk, uint32_t* hkey, uint32_t* hcnt, uint32_t* hbase,
                                              uint32_t* __restrict__ fill, const uint32_t* __restrict__ coff,
                                              uint32_t* __restrict__ spill) {
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    for (int s = threadIdx.x; s < HS; s += NT) {
        hkey[s] = kNo;
        hcnt[s] = 0u;
    }
    __syncthreads();
    const int64_t pos = stage_reserve<HS, NT>(kk, hkey, hcnt, hbase, fill, coff, spill);
    __syncthreads();
    return pos;
}

// Tile of a position in [0, ng)^3 for power-of-two tiles of ts cells, and
// the offsets inside it: u32 at 2^-28 cells (FP32 mesh) or 42-bit at 2^-38
// cells (FP64 mesh).
__device__ __forceinline__ uint32_t tile_key(double x, double y, double z, int ng, int ts) {
    const int lt = __builtin_ctz((unsigned)ts), nt = ng >> lt;
    return (uint32_t)((((int)x >> lt) * nt + ((int)y >> lt)) * nt + ((int)z >> lt));
}
// The record of in-tile offsets (lx, ly, lz) in [0, ts)^3 (exact: a position
// minus its tile origin): u32 fixed point at 2^-28 cells for the FP32 mesh
// (cell = q >> 28, fraction = the low 28 bits), 42-bit fields at 2^-38 cells
// for the FP64 mesh, both far below the mesh's own rounding.  Truncation, so
// that a position and its record always name the same cell.
template <typename Rec> __device__ __forceinline__ Rec tile_record_of(double lx, double ly, double lz);
template <> __device__ __forceinline__ StageQ tile_record_of<StageQ>(double lx, double ly, double lz) {
    return StageQ{(uint32_t)(lx * 268435456.0), (uint32_t)(ly * 268435456.0), (uint32_t)(lz * 268435456.0)};
}
template <> __device__ __forceinline__ StageD tile_record_of<StageD>(double lx, double ly, double lz) {
    return pack_stage_d((uint64_t)(lx * kQScaleD), (uint64_t)(ly * kQScaleD), (uint64_t)(lz * kQScaleD));
}
template <typename Rec> __device__ __forceinline__ Rec tile_record(double x, double y, double z, int ts) {
    const int msk = ~(ts - 1);  // tile origin = cell with the low bits cleared
    return tile_record_of<Rec>(x - ((int)x & msk), y - ((int)y & msk), z - ((int)z & msk));
}

// Split-precision positions: one u32 per axis, q = x * 2^S with
// S = 32 - log2(ng), so the cell is q >> S and the fraction the low S bits.
__host__ __device__ inline int split_shift(int ng) { return 32 - ilog2(ng); }
// tile of q: sh = S + log2(ts) = 32 - log2(tiles per axis), 32 for one tile
__device__ __forceinline__ uint32_t split_key(uint32_t qx, uint32_t qy, uint32_t qz, int sh, uint32_t nt) {
    const auto t = [sh](uint32_t q) { return (uint32_t)((uint64_t)q >> sh); };
    return (t(qx) * nt + t(qy)) * nt + t(qz);
}
// Element a (axis) of slot s in an array of n slots of the split state: the
// three values of a slot side by side (a tile's bucket is then one
// contiguous stream per array), or three planes of n.
#ifndef PM_STATE_TRIPLES
#define PM_STATE_TRIPLES 1
#endif
__host__ __device__ inline int64_t sel(int64_t n, int64_t s, int a) {
    return PM_STATE_TRIPLES ? 3 * s + a : (int64_t)a * n + s;
}
// In-tile record of one axis: cell in the tile << 28 | fraction at 2^-28
// cells (truncated for S > 28), the StageQ layout of tile_record.
__device__ __forceinline__ uint32_t split_rec(uint32_t q, int S, uint32_t tmask) {
    const uint32_t cell = (q >> S) & tmask;
    const uint32_t frac = q & ((1u << S) - 1u);
    return (cell << 28) | (S >= 28 ? frac >> (S - 28) : frac << (28 - S));
}

// The split state by mesh precision: positions as Q fixed point over the
// box (q = x 2^S, S = bits of Q - log2 ng: 2^-23 cells at 512^3 in u32 for
// the FP32 mesh, 2^-55 cells in u64 for the FP64 mesh, finer than an FP64
// position), velocities as U.
template <typename Rl> struct SplitWords;
template <> struct SplitWords<float> {
    using Q = uint32_t;
    using U = float;
};
template <> struct SplitWords<double> {
    using Q = uint64_t;
    using U = double;
};
template <typename Q> __host__ __device__ inline int qshift(int ng) { return 8 * (int)sizeof(Q) - ilog2(ng); }
// tile of q: sh = S + log2(ts) = bits of Q - log2(tiles per axis), all of Q
// for one tile
template <typename Q> __device__ __forceinline__ uint32_t qkey(Q qx, Q qy, Q qz, int sh, uint32_t nt) {
    if constexpr (sizeof(Q) == 4) return split_key(qx, qy, qz, sh, nt);  // a 64-bit shift: 32 gives 0
    const auto t = [sh](Q q) { return sh >= 64 ? 0u : (uint32_t)(q >> sh); };
    return (t(qx) * nt + t(qy)) * nt + t(qz);
}

// Cell in the tile and offsets in the cell of a record: the deposits take
// every particle through a record (read from the staging, or formed from its
// position), so that staged and binned paths give the same weights.
struct TileOffs {
    int a, b, c;
    double fx, fy, fz;
};
__device__ __forceinline__ TileOffs tile_offs(const StageQ& r) {
    constexpr uint32_t m = 0x0FFFFFFFu;
    constexpr double s = 1.0 / 268435456.0;
    return TileOffs{(int)(r.x >> 28), (int)(r.y >> 28), (int)(r.z >> 28), (r.x & m) * s, (r.y & m) * s, (r.z & m) * s};
}
__device__ __forceinline__ TileOffs tile_offs(const StageD& r) {
    uint64_t qx, qy, qz;
    unpack_stage_d(r, qx, qy, qz);
    return TileOffs{(int)(qx >> 38), (int)(qy >> 38), (int)(qz >> 38), (double)(qx & kQMaskD) / kQScaleD,
                    (double)(qy & kQMaskD) / kQScaleD, (double)(qz & kQMaskD) / kQScaleD};
}

// Cell in the tile and offsets of a split position triple, as the tile
// deposits form them: the FP32 mesh through the StageQ record (2^-28 cells),
// the FP64 mesh with the fraction truncated to 2^-38 cells (the StageD
// resolution).
template <typename Q>
__device__ __forceinline__ TileOffs split_offs(Q qa, Q qb, Q qc, int S, uint32_t tmask) {
    if constexpr (sizeof(Q) == 4) {
        return tile_offs(StageQ{split_rec(qa, S, tmask), split_rec(qb, S, tmask), split_rec(qc, S, tmask)});
    } else {
        const auto cell = [&](Q q) { return (int)((q >> S) & tmask); };
        const auto frac = [&](Q q) { return (double)((q & ((1ull << S) - 1ull)) >> (S - 38)) / kQScaleD; };
        return TileOffs{cell(qa), cell(qb), cell(qc), frac(qa), frac(qb), frac(qc)};
    }
}

}  // namespace pm
