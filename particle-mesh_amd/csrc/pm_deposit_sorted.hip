// Deterministic CIC deposit: particles binned by mesh tile, each tile's
// mass accumulated exactly in 64-bit fixed point in LDS.
//
// deposit_particles (utils.cpp:70-102) gives every thread a private grid
// and sums the grids (utils.cpp:104-113); the GPU equivalent of a private
// grid is a workgroup's LDS tile:
//   1. bin_count   : tile of each particle's cell, slot within the tile
//                    (LDS-hash aggregation: one global atomic per distinct
//                    tile per chunk of 2048 particles)
//   2. scan        : exclusive prefix sum of the tile counts
//   3. bin_scatter : particle indices in tile order
//   4. tiles       : one workgroup per segment of consecutive tiles of TS^3
//                    cells along z (the column walk of pm_tiles.h); the CIC
//                    weights of a tile's particles are added in 2^-32 fixed
//                    point to a (TS+1)^3 LDS node tile (integer adds: exact,
//                    order-free), the TS^3 owned nodes are written to the
//                    mesh, the upper x / y faces to the tile's xy shell, and
//                    the upper z plane is carried in LDS into the next tile
//                    (the segment's last one goes to its z seam)
//   5. fold        : the lower x / y face nodes of each tile get the xy
//                    shells of its lower neighbours, a segment's first plane
//                    the seams, in a fixed order -- the periodic ghost fold
//                    of reorder_slab.cpp:43-67 at tile granularity (by mesh
//                    lines: face lines whole, then the seam nodes).
// Every mesh node is written by exactly one workgroup: no global float
// atomics, and the result does not depend on the particle order.
#include <type_traits>

#include "pm_internal.h"
#include "pm_tiles.h"

namespace pm {

constexpr uint32_t kInvalidKey = 0xFFFFFFFFu;

// 1. tile id and slot of every particle.  A workgroup takes a chunk of
// kBinChunk particles, counts them per tile in an LDS hash table (open
// addressing), and issues one global atomic per distinct tile of the chunk,
// all in parallel; a particle's slot is that atomic's result plus its rank
// within the chunk.  Spatially coherent chunks touch few tiles.
constexpr int kBinPer = 8;                    // particles per thread
constexpr int kBinChunk = 256 * kBinPer;      // particles per workgroup
constexpr int kBinHash = 2 * kBinChunk;       // hash slots (load factor <= 1/2)

__device__ __forceinline__ uint32_t bin_hash(uint32_t k) { return (k * 2654435761u) >> 20; }

__global__ __launch_bounds__(256) void k_bin_count(const double* __restrict__ px, const double* __restrict__ py,
                                                   const double* __restrict__ pz, int64_t np, int ng, int ts,
                                                   uint32_t* __restrict__ cnt, uint32_t* __restrict__ key,
                                                   uint32_t* __restrict__ slot,
                                                   unsigned long long* __restrict__ err,
                                                   const uint32_t* __restrict__ gate) {
    static_assert((kBinHash & (kBinHash - 1)) == 0, "hash size must be a power of two");
    if (gate && *gate == 0u) return;
    __shared__ uint32_t hkey[kBinHash];
    __shared__ uint32_t hcnt[kBinHash];
    const int64_t nchunk = (np + kBinChunk - 1) / kBinChunk;
    // grid-stride over the chunks (the grid is capped, so that a gated no-op
    // launch stays cheap)
    for (int64_t ch = xcd_remap(blockIdx.x, gridDim.x); ch < nchunk; ch += gridDim.x) {
    for (int h = threadIdx.x; h < kBinHash; h += 256) {
        hkey[h] = kInvalidKey;
        hcnt[h] = 0u;
    }
    __syncthreads();
    // each thread takes kBinPer consecutive particles: neighbouring particles
    // usually share a tile, so runs of equal keys cost one hash insert
    const int64_t base = ch * kBinChunk + (int64_t)threadIdx.x * kBinPer;
    const int lt = __builtin_ctz((unsigned)ts), nt = ng >> lt;  // ts: power of two
    const double box = (double)ng;
    uint32_t k[kBinPer], h[kBinPer], rank[kBinPer];
    double x[kBinPer], y[kBinPer], z[kBinPer];
    const bool aligned = (((uintptr_t)px | (uintptr_t)py | (uintptr_t)pz) & 15) == 0;
    if (aligned && base + kBinPer <= np) {
        const double2* vx = reinterpret_cast<const double2*>(px + base);
        const double2* vy = reinterpret_cast<const double2*>(py + base);
        const double2* vz = reinterpret_cast<const double2*>(pz + base);
#pragma unroll
        for (int j = 0; j < kBinPer / 2; ++j) {
            const double2 a = vx[j], b = vy[j], c = vz[j];
            x[2 * j] = a.x;
            x[2 * j + 1] = a.y;
            y[2 * j] = b.x;
            y[2 * j + 1] = b.y;
            z[2 * j] = c.x;
            z[2 * j + 1] = c.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kBinPer; ++j) {
            const bool in = base + j < np;
            x[j] = in ? px[base + j] : -1.0;
            y[j] = in ? py[base + j] : -1.0;
            z[j] = in ? pz[base + j] : -1.0;
        }
    }
#pragma unroll
    for (int j = 0; j < kBinPer; ++j) {
        h[j] = 0u;
        rank[j] = 0u;
        if (in_box3(x[j], y[j], z[j], box)) {
            k[j] = (uint32_t)((((int)x[j] >> lt) * nt + ((int)y[j] >> lt)) * nt + ((int)z[j] >> lt));
        } else {
            k[j] = kInvalidKey;
            if (base + j < np) atomicAdd(err, 1ULL);
        }
    }
    int start = 0;
#pragma unroll
    for (int j = 1; j <= kBinPer; ++j) {
        if (j < kBinPer && k[j] == k[start]) continue;
        if (k[start] != kInvalidKey) {
            uint32_t s = bin_hash(k[start]) & (kBinHash - 1);
            while (true) {  // at most kBinChunk distinct keys: a free slot always exists
                const uint32_t old = atomicCAS(&hkey[s], kInvalidKey, k[start]);
                if (old == kInvalidKey || old == k[start]) break;
                s = (s + 1) & (kBinHash - 1);
            }
            const uint32_t r = atomicAdd(&hcnt[s], (uint32_t)(j - start));
#pragma unroll
            for (int m = 0; m < kBinPer; ++m)
                if (m >= start && m < j) {
                    h[m] = s;
                    rank[m] = r + (uint32_t)(m - start);
                }
        }
        start = j;
    }
    __syncthreads();
    // one global atomic per distinct tile; hcnt becomes the chunk's base slot
    for (int s = threadIdx.x; s < kBinHash; s += 256) {
        const uint32_t kk = hkey[s];
        if (kk != kInvalidKey) hcnt[s] = atomicAdd(&cnt[kk], hcnt[s]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBinPer; ++j) {
        const int64_t i = base + j;
        if (i < np) {
            key[i] = k[j];
            slot[i] = k[j] == kInvalidKey ? 0u : hcnt[h[j]] + rank[j];
        }
    }
    __syncthreads();  // the hash is reset for the next chunk
    }
}

// Exclusive scan by one workgroup of 1024 threads of the n values load(i)
// into out[0..n] (out[n] = total).  The values pass through LDS in chunks of
// 16384 with coalesced loads and stores, each thread scanning 16 consecutive
// ones; then a wave scan by shuffles and one over the 16 wave totals.
constexpr int kScanThreads = 1024;

template <typename Load>
__device__ __forceinline__ void scan_workgroup(int n, Load load, uint32_t* __restrict__ out) {
    constexpr int T = kScanThreads, PER = 16, CH = T * PER;
    __shared__ uint32_t buf[CH + CH / PER];  // entry k at k + k / 16: the thread rows have stride 17
    __shared__ uint32_t wsum[T / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t carry = 0u;
    for (int base = 0; base < n; base += CH) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int k = j * T + t, i = base + k;
            buf[k + k / PER] = i < n ? load(i) : 0u;
        }
        __syncthreads();
        uint32_t v[PER], sum = 0u;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            v[j] = buf[t * (PER + 1) + j];
            sum += v[j];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint32_t before = 0u, total = 0u;
#pragma unroll
        for (int w = 0; w < T / 64; ++w) {
            const uint32_t x = wsum[w];
            before += w < wv ? x : 0u;
            total += x;
        }
        uint32_t run = carry + before + incl - sum;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            buf[t * (PER + 1) + j] = run;
            run += v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int k = j * T + t, i = base + k;
            if (i < n) out[i] = buf[k + k / PER];
        }
        carry += total;
        __syncthreads();  // buf and wsum are refilled by the next chunk
    }
    if (t == 0) out[n] = carry;
}

// 2. exclusive scan of the n tile counts into off[0..n], the counts zeroed
// behind it for the next binning.
__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t* __restrict__ cnt, uint32_t* __restrict__ off, int n,
                                                       const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    scan_workgroup(
        n,
        [&](int i) {
            const uint32_t c = cnt[i];
            cnt[i] = 0u;
            return c;
        },
        off);
}

// 3. particle indices in tile order.
__global__ __launch_bounds__(256) void k_bin_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ slot,
                                                     const uint32_t* __restrict__ off, int64_t np,
                                                     uint32_t* __restrict__ idx, const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x; i < np; i += stride) {
        const uint32_t k = key[i];
        if (k != kInvalidKey) idx[off[k] + slot[i]] = (uint32_t)i;
    }
}


// 4. one workgroup per segment of the column walk, particles by index
// (binning of positions).
// Fixed-point unit of the tile deposits: 2^-q of a density unit, q = 32
// unless a particle's mass m is so large (m >= 2^19, fewer than ng^3 / 2^19
// particles) that m 2^32 would reach 2^51, the range of fix_rne; a tile's
// sum stays below ng^3 2^q <= 2^62 either way.
__device__ __forceinline__ int fix_bits(double mass) {
    int q = 32;
    while (q > 0 && ldexp(mass, q) >= 2251799813685248.0) --q;  // 2^51
    return q;
}

template <typename Rl, int TS>
__global__ __launch_bounds__(256) void k_deposit_tiles(const double* __restrict__ px, const double* __restrict__ py,
                                                       const double* __restrict__ pz,
                                                       const uint32_t* __restrict__ idx,
                                                       const uint32_t* __restrict__ off, int ng, double mass,
                                                       Rl* __restrict__ rho, Rl* __restrict__ shell,
                                                       const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    using St = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    constexpr int E = TS + 1, E3 = E * E * E;
    __shared__ unsigned long long acc[E3];
    const WalkSeg sw = walk_segment(ng, TS);
    const int qb = fix_bits(mass);
    const double scale[1] = {ldexp(1.0, -qb)};
    Rl* const out[1] = {rho};
    Rl* const sh[1] = {shell};
    for (int i = threadIdx.x; i < E3; i += 256) acc[i] = 0ULL;
    __syncthreads();
    const double fix = ldexp(mass, qb);  // 2^q * m
    const int ox = sw.tx * TS, oy = sw.ty * TS;
    for (int k = 0; k < sw.seg; ++k) {
        const int tz = sw.sg * sw.seg + k, oz = tz * TS;
        const uint32_t t = ((uint32_t)sw.tx * sw.nt + sw.ty) * sw.nt + tz;
        const uint32_t beg = off[t], end = off[t + 1];
        // kDepBatch particles per thread per pass: their index and position
        // loads are all in flight before the first LDS atomic
        for (uint32_t q0 = beg + threadIdx.x; q0 < end; q0 += 256 * kDepBatch) {
            uint32_t p[kDepBatch];
            double x[kDepBatch], y[kDepBatch], z[kDepBatch];
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                const uint32_t q = q0 + 256 * j;
                p[j] = q < end ? idx[q] : kInvalidKey;
            }
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                if (p[j] != kInvalidKey) {
                    x[j] = px[p[j]];
                    y[j] = py[p[j]];
                    z[j] = pz[p[j]];
                }
            }
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                if (p[j] == kInvalidKey) continue;
                // the staged records' fixed-point offsets, formed the same way
                const TileOffs o = tile_offs(tile_record_of<St>(x[j] - ox, y[j] - oy, z[j] - oz));
                tile_cic_add<TS>(acc, o.a, o.b, o.c, o.fx, o.fy, o.fz, fix);
            }
        }
        __syncthreads();
        walk_tile_out<Rl, TS, 1, 256>(acc, scale, t, ox, oy, oz, ng, out, sh);
    }
    walk_seam_out<Rl, TS, 1, 256>(acc, scale, sw, sh);
}

// 4'. as 4, particles from the staging arrays that the interpolation of the
// previous step filled in tile order: tile t holds fill[t] records from
// coff[t], in-tile offsets in fixed point (2^-28 cells, FP32 mesh; 2^-38
// cells, FP64 mesh), which the binned path forms in the same way, so both
// give the same weights bit for bit.  Skipped when a tile overflowed its capacity (the binned path runs
// instead).
template <typename Rl, int TS>
__global__ __launch_bounds__(256) void k_deposit_staged(const void* __restrict__ stage,
                                                        const uint32_t* __restrict__ coff,
                                                        const uint32_t* __restrict__ fill,
                                                        const uint32_t* __restrict__ spill, int ng, double mass,
                                                        Rl* __restrict__ rho, Rl* __restrict__ shell) {
    if (*spill != 0u) return;
    using St = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    const St* __restrict__ sp = static_cast<const St*>(stage);
    constexpr int E = TS + 1, E3 = E * E * E;
    __shared__ unsigned long long acc[E3];
    const WalkSeg sw = walk_segment(ng, TS);
    const int qb = fix_bits(mass);
    const double scale[1] = {ldexp(1.0, -qb)};
    Rl* const out[1] = {rho};
    Rl* const sh[1] = {shell};
    for (int i = threadIdx.x; i < E3; i += 256) acc[i] = 0ULL;
    __syncthreads();
    const double fix = ldexp(mass, qb);
    for (int k = 0; k < sw.seg; ++k) {
        const int tz = sw.sg * sw.seg + k;
        const uint32_t t = ((uint32_t)sw.tx * sw.nt + sw.ty) * sw.nt + tz;
        const uint32_t beg = coff[t], end = beg + fill[t];
        for (uint32_t q0 = beg + threadIdx.x; q0 < end; q0 += 256 * kDepBatch) {
            St v[kDepBatch];
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                const uint32_t q = q0 + 256 * j;
                if (q < end) v[j] = sp[q];
            }
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                if (q0 + 256 * j >= end) continue;
                const TileOffs o = tile_offs(v[j]);
                tile_cic_add<TS>(acc, o.a, o.b, o.c, o.fx, o.fy, o.fz, fix);
            }
        }
        __syncthreads();
        walk_tile_out<Rl, TS, 1, 256>(acc, scale, t, sw.tx * TS, sw.ty * TS, tz * TS, ng, out, sh);
    }
    walk_seam_out<Rl, TS, 1, 256>(acc, scale, sw, sh);
}

// 4''. as 4, particles from the tile buckets of the split-precision state
// (pm_split.hip: positions as fixed-point cell + fraction, u32 for an FP32
// mesh, u64 for an FP64 mesh, in triples of cap slots; tile t holds bfill[t]
// slots from boff[t], holes among them, whose tile does not match): the
// in-tile offsets are formed from the positions exactly (split_offs: at the
// staged records' resolution, 2^-28 / 2^-38 cells), so the weights are those
// of the staged deposit of the same positions.  Also leaves each tile's used
// slot count in bnow, the slots the step's interpolation takes.
template <typename Rl, int TS>
__device__ __forceinline__ void deposit_state(const typename SplitWords<Rl>::Q* __restrict__ q0,
                                              const typename SplitWords<Rl>::Q* __restrict__ q1,
                                              const uint32_t* __restrict__ ctl, int64_t cap,
                                              const uint32_t* __restrict__ boff, const uint32_t* __restrict__ bfill,
                                              uint32_t* __restrict__ bnow, int ng, double mass, Rl* __restrict__ rho,
                                              Rl* __restrict__ shell) {
    using Q = typename SplitWords<Rl>::Q;
    constexpr int E = TS + 1, E3 = E * E * E;
    __shared__ unsigned long long acc[E3];
    const Q* __restrict__ q = ctl[3] ? q1 : q0;  // the current set of slots
    const WalkSeg sw = walk_segment(ng, TS);
    const int qb = fix_bits(mass);
    const double scale[1] = {ldexp(1.0, -qb)};
    Rl* const out[1] = {rho};
    Rl* const sh[1] = {shell};
    const int S = qshift<Q>(ng), ksh = S + ilog2(TS);
    const uint32_t nt = (uint32_t)sw.nt, tmask = TS - 1;
    for (int i = threadIdx.x; i < E3; i += 256) acc[i] = 0ULL;
    __syncthreads();
    const double fix = ldexp(mass, qb);
    for (int k = 0; k < sw.seg; ++k) {
        const int tz = sw.sg * sw.seg + k;
        const uint32_t t = ((uint32_t)sw.tx * sw.nt + sw.ty) * sw.nt + tz;
        const uint32_t beg = boff[t], n = min(bfill[t], boff[t + 1] - beg), end = beg + n;
        if (threadIdx.x == 0) bnow[t] = n;
        for (uint32_t s0 = beg + threadIdx.x; s0 < end; s0 += 256 * kDepBatch) {
            Q a[kDepBatch], b[kDepBatch], c[kDepBatch];
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                const uint32_t s = s0 + 256 * j;
                const int64_t r = s < end ? s : beg;
                a[j] = q[sel(cap, r, 0)];
                b[j] = q[sel(cap, r, 1)];
                c[j] = q[sel(cap, r, 2)];
            }
#pragma unroll
            for (int j = 0; j < kDepBatch; ++j) {
                if (s0 + 256 * j >= end || qkey<Q>(a[j], b[j], c[j], ksh, nt) != t) continue;  // a hole
                const TileOffs o = split_offs<Q>(a[j], b[j], c[j], S, tmask);
                tile_cic_add<TS>(acc, o.a, o.b, o.c, o.fx, o.fy, o.fz, fix);
            }
        }
        __syncthreads();
        walk_tile_out<Rl, TS, 1, 256>(acc, scale, t, sw.tx * TS, sw.ty * TS, tz * TS, ng, out, sh);
    }
    walk_seam_out<Rl, TS, 1, 256>(acc, scale, sw, sh);
}
template <int TS>
__global__ __launch_bounds__(256) void k_deposit_state(const uint32_t* __restrict__ q0, const uint32_t* __restrict__ q1,
                                                       const uint32_t* __restrict__ ctl, int64_t cap,
                                                       const uint32_t* __restrict__ boff,
                                                       const uint32_t* __restrict__ bfill, uint32_t* __restrict__ bnow,
                                                       int ng, double mass, float* __restrict__ rho,
                                                       float* __restrict__ shell) {
    deposit_state<float, TS>(q0, q1, ctl, cap, boff, bfill, bnow, ng, mass, rho, shell);
}
template <int TS>
__global__ __launch_bounds__(256) void k_deposit_state64(const uint64_t* __restrict__ q0,
                                                         const uint64_t* __restrict__ q1,
                                                         const uint32_t* __restrict__ ctl, int64_t cap,
                                                         const uint32_t* __restrict__ boff,
                                                         const uint32_t* __restrict__ bfill,
                                                         uint32_t* __restrict__ bnow, int ng, double mass,
                                                         double* __restrict__ rho, double* __restrict__ shell) {
    deposit_state<double, TS>(q0, q1, ctl, cap, boff, bfill, bnow, ng, mass, rho, shell);
}

// Capacities of the staging buckets for the next step: count + count/8 + 32
// per tile (counts from the binning offsets on a first step, else from the
// fill counters of the staged step), or of the split state's buckets from
// its tile counts (count + count/2 + 64 rounded up to 32; the counts are
// cleared), exclusive
// scan into coff[0..n]; the fill counters and the overflow flag are cleared
// for the next staging.  A non-null gate makes the kernel a no-op unless
// *gate != 0.
__global__ __launch_bounds__(kScanThreads) void k_cap_scan(uint32_t* __restrict__ fill,
                                                           const uint32_t* __restrict__ off,
                                                           uint32_t* __restrict__ cnt,
                                                           uint32_t* __restrict__ coff, uint32_t* __restrict__ spill,
                                                           int n, int from, const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    scan_workgroup(
        n,
        [&](int i) {
            uint32_t c;
            if (from == kCapFromOff) {
                c = off[i + 1] - off[i];
            } else if (from == kCapFromCnt) {
                c = cnt[i];
                cnt[i] = 0u;
                fill[i] = 0u;
                return (c + (c >> 1) + 64u + 31u) & ~31u;  // split_capacity: arrivals, whole lines
            } else {
                c = fill[i];
            }
            fill[i] = 0u;
            return c + (c >> 3) + 32u;
        },
        coff);
    if (threadIdx.x == 0) *spill = 0u;
}

// 5. the fold: each tile's lower-face nodes get the shells of its lower x,
// y and xy neighbours, and a segment's first tile the seams of the previous
// segment (own, x, y and xy columns), in the fixed order d = 1..7 (bit 0: x
// neighbour, 1: y, 2: z).  The column walks leave H shell planes per tile
// (H = 1 for the CIC deposit, node box unshifted; H = 3 for the divergence
// deposit, node box shifted one node down, SH = 1).  In the walk's frame
// (g = node + SH) a node folds when its tile-local index is below H on x or
// y (every node of such an (x, y) line) or on z in the first tile of a
// segment (H nodes per segment on every other line).  k_fold_faces takes the
// face lines whole, one workgroup per line, so that the mesh is read and
// written in whole lines; k_fold_seams the segments' first planes on the
// other lines.  (One workgroup per tile, with the faces in runs of TS nodes,
// took 0.128 against 0.114 ms FP32 and 0.16 against 0.13 ms FP64 at 512^3.)
template <typename Rl, int TS, int H, int SH>
__device__ __forceinline__ Rl fold_value(const Rl* __restrict__ shell, int ng, int gx, int gy, int gz, Rl v) {
    constexpr int E = TS + H, XYS = halo_shell_size<TS, H>(), ZS = halo_seam_size<TS, H>();
    const int nt = ng / TS, seg = walk_seg(nt), nseg = nt / seg;
    const int a = gx & (TS - 1), b = gy & (TS - 1), c = gz & (TS - 1);
    const int tx = gx / TS, ty = gy / TS, tz = gz / TS, sg = tz / seg;
    const bool first = tz == sg * seg;
    const int on = (a < H) | ((b < H) << 1) | ((first && c < H) << 2);
    if (on == 0) return v;
    const Rl* __restrict__ seam = shell + (size_t)nt * nt * nt * XYS;
    Rl sv[8];
#pragma unroll
    for (int d = 1; d < 8; ++d) {
        const int dx = d & 1, dy = (d >> 1) & 1;
        const int sx = tx - dx < 0 ? nt - 1 : tx - dx;
        const int sy = ty - dy < 0 ? nt - 1 : ty - dy;
        const int na = a + TS * dx, nb = b + TS * dy;
        sv[d] = Rl(0);
        if ((d & on) != d) continue;
        if (d < 4) {
            sv[d] = shell[(((size_t)sx * nt + sy) * nt + tz) * XYS + halo_shell_index<TS, H>(na, nb, c)];
        } else {
            const int ss = sg == 0 ? nseg - 1 : sg - 1;
            sv[d] = seam[(((size_t)sx * nt + sy) * nseg + ss) * ZS + ((size_t)na * E + nb) * H + c];
        }
    }
#pragma unroll
    for (int d = 1; d < 8; ++d)
        if ((d & on) == d) v += sv[d];
    return v;
}

// face lines: H nt ng lines with a < H, then (TS - H) nt x H nt with b < H <= a
template <typename Rl, int TS, int H, int SH>
__global__ __launch_bounds__(256) void k_fold_faces(const Rl* __restrict__ shell, int ng, Rl* __restrict__ mesh) {
    const int nt = ng / TS, m = ng - 1;
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x), nxf = (uint32_t)(H * nt) * ng;
    int gx, gy;
    if (w < nxf) {
        const int ax = (int)(w / (uint32_t)ng);
        gx = (ax / H) * TS + ax % H;
        gy = (int)(w % (uint32_t)ng);
    } else {
        const uint32_t r = w - nxf, hy = (uint32_t)(H * nt);
        const int bx = (int)(r / hy), by = (int)(r % hy);
        gx = (bx / (TS - H)) * TS + H + bx % (TS - H);
        gy = (by / H) * TS + by % H;
    }
    Rl* __restrict__ line = mesh + ((size_t)((gx - SH) & m) * ng + (size_t)((gy - SH) & m)) * ng;
    for (int z = threadIdx.x; z < ng; z += 256) {
        const int gz = (z + SH) & m;
        line[z] = fold_value<Rl, TS, H, SH>(shell, ng, gx, gy, gz, line[z]);
    }
}

// the segments' first H planes on the lines that are on no face
template <typename Rl, int TS, int H, int SH>
__global__ __launch_bounds__(256) void k_fold_seams(const Rl* __restrict__ shell, int ng, Rl* __restrict__ mesh) {
    const int nt = ng / TS, seg = walk_seg(nt), nseg = nt / seg, m = ng - 1;
    const int side = (TS - H) * nt;  // lines per axis that are on no face
    const int64_t n = (int64_t)side * side * nseg * H;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int c = (int)(i % H), sgi = (int)((i / H) % nseg);
        const int64_t l = i / ((int64_t)H * nseg);
        const int bx = (int)(l / side), by = (int)(l % side);
        const int gx = (bx / (TS - H)) * TS + H + bx % (TS - H);
        const int gy = (by / (TS - H)) * TS + H + by % (TS - H);
        const int gz = sgi * seg * TS + c;
        Rl* __restrict__ p =
            mesh + ((size_t)((gx - SH) & m) * ng + (size_t)((gy - SH) & m)) * ng + (size_t)((gz - SH) & m);
        *p = fold_value<Rl, TS, H, SH>(shell, ng, gx, gy, gz, *p);
    }
}

template <typename Rl, int TS, int H, int SH>
static void launch_fold_lines(int ng, const void* shell, void* mesh, hipStream_t st) {
    const int nt = ng / TS, side = (TS - H) * nt, nseg = nt / walk_seg(nt);
    const unsigned faces = (unsigned)((size_t)ng * ng - (size_t)side * side);
    hipLaunchKernelGGL((k_fold_faces<Rl, TS, H, SH>), dim3(faces), dim3(256), 0, st, (const Rl*)shell, ng, (Rl*)mesh);
    const int64_t n = (int64_t)side * side * nseg * H;
    if (n > 0) {
        const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL((k_fold_seams<Rl, TS, H, SH>), dim3(nb), dim3(256), 0, st, (const Rl*)shell, ng,
                           (Rl*)mesh);
    }
}

int sorted_tile(int ng) { return ng >= 16 ? 16 : 8; }

size_t sorted_shell_elems(int ng) {
    const int ts = sorted_tile(ng);
    const size_t nt = (size_t)(ng / ts);
    const size_t nseg = nt * nt * (nt / (size_t)walk_seg((int)nt));
    return ts == 16 ? nt * nt * nt * xy_shell_size<16>() + nseg * zseam_size<16>()
                    : nt * nt * nt * xy_shell_size<8>() + nseg * zseam_size<8>();
}

size_t halo_shell_elems(int ng) {
    const int ts = sorted_tile(ng);
    const size_t nt = (size_t)(ng / ts);
    const size_t nseg = nt * nt * (nt / (size_t)walk_seg((int)nt));
    return ts == 16 ? nt * nt * nt * halo_shell_size<16, kDivHalo>() + nseg * halo_seam_size<16, kDivHalo>()
                    : nt * nt * nt * halo_shell_size<8, kDivHalo>() + nseg * halo_seam_size<8, kDivHalo>();
}

int launch_fold_halo(int prec, int ts, int ng, const void* shell, void* mesh, hipStream_t st) {
#define PM_FH(RL, TSV) launch_fold_lines<RL, TSV, kDivHalo, 1>(ng, shell, mesh, st)
    if (prec == PM_FP32) {
        if (ts == 16)
            PM_FH(float, 16);
        else
            PM_FH(float, 8);
    } else {
        if (ts == 16)
            PM_FH(double, 16);
        else
            PM_FH(double, 8);
    }
#undef PM_FH
    return check_launch("fold_halo");
}

int launch_bin(const double* px, const double* py, const double* pz, int64_t np, int ng, SortedScratch& s,
               unsigned long long* err, hipStream_t st, const uint32_t* gate) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts;
    const uint32_t ntiles = (uint32_t)nt * nt * nt;
    if (np > 0xFFFFFFF0LL) return fail(PM_ERR_ARG, "sorted deposit: more than 2^32 particles");
    // the counts are zero here: cleared at allocation, then by every k_scan
    if (np > 0) {
        const unsigned nchunk = (unsigned)((np + kBinChunk - 1) / kBinChunk);
        hipLaunchKernelGGL(k_bin_count, dim3(std::min(nchunk, 4096u)), dim3(256), 0, st, px, py, pz, np, ng, ts,
                           s.cnt, s.key, s.slot, err, gate);
    }
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, s.cnt, s.off, (int)ntiles, gate);
    if (np > 0) {
        const unsigned nb = (unsigned)((np + 255) / 256);
        // grid-stride over at most 16384 workgroups (a gated no-op stays cheap)
        hipLaunchKernelGGL(k_bin_scatter, dim3(std::min(nb, 16384u)), dim3(256), 0, st, s.key, s.slot, s.off, np,
                           s.idx, gate);
    }
    return check_launch("bin");
}

int launch_fold_shells(int prec, int ts, int ng, const void* shell, void* rho, hipStream_t st) {
    if (prec == PM_FP32) {
        if (ts == 16)
            launch_fold_lines<float, 16, 1, 0>(ng, shell, rho, st);
        else
            launch_fold_lines<float, 8, 1, 0>(ng, shell, rho, st);
    } else {
        if (ts == 16)
            launch_fold_lines<double, 16, 1, 0>(ng, shell, rho, st);
        else
            launch_fold_lines<double, 8, 1, 0>(ng, shell, rho, st);
    }
    return check_launch("fold_shells");
}

int launch_deposit_binned(int prec, const double* px, const double* py, const double* pz, int ng, double mass,
                          void* rho, SortedScratch& s, hipStream_t st, const uint32_t* gate, bool fold) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts, seg = walk_seg(nt);
    const uint32_t ntiles = (uint32_t)nt * nt * nt;
#define PM_DT(RL, TSV)                                                                                                 \
    hipLaunchKernelGGL((k_deposit_tiles<RL, TSV>), dim3(ntiles / seg), dim3(256), 0, st, px, py, pz, s.idx, s.off, ng,  \
                       mass, (RL*)rho, (RL*)s.shell, gate)
    if (prec == PM_FP32) {
        if (ts == 16)
            PM_DT(float, 16);
        else
            PM_DT(float, 8);
    } else {
        if (ts == 16)
            PM_DT(double, 16);
        else
            PM_DT(double, 8);
    }
#undef PM_DT
    int rc = check_launch("deposit_tiles");
    if (rc || !fold) return rc;
    return launch_fold_shells(prec, ts, ng, s.shell, rho, st);
}

int launch_cap_scan(int ng, SortedScratch& s, bool from_off, hipStream_t st) {
    return launch_cap_scan_from(ng, s, from_off ? kCapFromOff : kCapFromFill, nullptr, st);
}

int launch_cap_scan_from(int ng, SortedScratch& s, int from, const uint32_t* gate, hipStream_t st) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts;
    hipLaunchKernelGGL(k_cap_scan, dim3(1), dim3(kScanThreads), 0, st, s.fill, s.off, s.cnt, s.coff, s.spill,
                       nt * nt * nt, from, gate);
    return check_launch("cap_scan");
}

int launch_deposit_staged(int prec, const double* px, const double* py, const double* pz, int64_t np, int ng,
                          double mass, void* rho, SortedScratch& s, unsigned long long* err, hipStream_t st,
                          bool rescan, bool bin) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts, seg = walk_seg(nt);
    const uint32_t ntiles = (uint32_t)nt * nt * nt;
    // fallback, run only if some tile overflowed its staging capacity (none
    // for the split state, px == nullptr: its restaging ran before)
    int rc = PM_OK;
    if (px) {
        rc = bin ? launch_bin(px, py, pz, np, ng, s, err, st, s.spill) : PM_OK;
        if (!rc) rc = launch_deposit_binned(prec, px, py, pz, ng, mass, rho, s, st, s.spill, false);
        if (rc) return rc;
    }
#define PM_DS(RL, TSV)                                                                                               \
    hipLaunchKernelGGL((k_deposit_staged<RL, TSV>), dim3(ntiles / seg), dim3(256), 0, st, s.stage, s.coff, s.fill,    \
                       s.spill, ng, mass, (RL*)rho, (RL*)s.shell)
    if (prec == PM_FP32) {
        if (ts == 16)
            PM_DS(float, 16);
        else
            PM_DS(float, 8);
    } else {
        if (ts == 16)
            PM_DS(double, 16);
        else
            PM_DS(double, 8);
    }
#undef PM_DS
    rc = check_launch("deposit_staged");
    if (!rc) rc = launch_fold_shells(prec, ts, ng, s.shell, rho, st);
    if (!rc && rescan) rc = launch_cap_scan(ng, s, false, st);
    return rc;
}

int launch_cap_scan_raw(int n, uint32_t* fill, uint32_t* cnt, uint32_t* coff, uint32_t* flag, const uint32_t* gate,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_cap_scan, dim3(1), dim3(kScanThreads), 0, st, fill, nullptr, cnt, coff, flag, n, kCapFromCnt,
                       gate);
    return check_launch("cap_scan");
}

int launch_deposit_state(int ng, double mass, const SplitState& st, void* rho, SortedScratch& s, hipStream_t hs) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts, seg = walk_seg(nt);
    const unsigned nsegs = (unsigned)(nt * nt * (nt / seg));
#define PM_DS(K, RL, TSV)                                                                                            \
    hipLaunchKernelGGL((K<TSV>), dim3(nsegs), dim3(256), 0, hs, static_cast<const typename SplitWords<RL>::Q*>(st.q[0]), \
                       static_cast<const typename SplitWords<RL>::Q*>(st.q[1]), st.ctl, st.cap, st.boff, st.bfill,   \
                       st.bnow, ng, mass, (RL*)rho, (RL*)s.shell)
    if (st.prec == PM_FP64) {
        if (ts == 16)
            PM_DS(k_deposit_state64, double, 16);
        else
            PM_DS(k_deposit_state64, double, 8);
    } else {
        if (ts == 16)
            PM_DS(k_deposit_state, float, 16);
        else
            PM_DS(k_deposit_state, float, 8);
    }
#undef PM_DS
    const int rc = check_launch("deposit_state");
    return rc ? rc : launch_fold_shells(st.prec, ts, ng, s.shell, rho, hs);
}

int launch_deposit_sorted(int prec, const double* px, const double* py, const double* pz, int64_t np, int ng,
                          double mass, void* rho, SortedScratch& s, unsigned long long* err, hipStream_t st) {
    int rc = launch_bin(px, py, pz, np, ng, s, err, st);
    if (rc) return rc;
    return launch_deposit_binned(prec, px, py, pz, ng, mass, rho, s, st);
}

}  // namespace pm
