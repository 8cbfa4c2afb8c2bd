// Split-precision particle state of the fused step, resident in the plan in
// tile order.
//
// Positions: an integer cell plus a fixed-point fraction in one word per
// axis, q = x * 2^S: with an FP32 mesh a u32 (S = 32 - log2(ng), 2^-23 cells
// at ng = 512) and FP32 velocities, 24 bytes per particle against the 48 of
// FP64 positions and velocities; with an FP64 mesh a u64 (S = 64 - log2(ng),
// 2^-55 cells at 512, finer than an FP64 position) and FP64 velocities.  The
// whole periodic box is the word's range, so a drift is an integer add that
// wraps at the box edge by itself.
//
// The particles live in the buckets of the deposit's 16^3-cell tiles: tile t
// owns slots [boff[t], boff[t+1]) of the position and velocity arrays (a
// slot's three values side by side, sel()) and of the particle indices (ids,
// for pm_split_store), of which bfill[t] are used.  Every kernel then works tile by tile on contiguous
// slots:
//   deposit  the column walk of the sorted deposit (pm_tiles.h), reading a
//            tile's positions straight from its bucket;
//   interp   one workgroup per tile stages the potential around the tile
//            (19^3 nodes) in LDS and interpolates, kicks and drifts the
//            tile's particles from it: no gathers from HBM, no records.  A
//            particle that stays in its tile (all but 0.25-0.4% per step at
//            512^3) is written back in place; one that leaves goes to the
//            tile's outbox with its index and its slot becomes a hole (x in
//            the other half of the box, so that its tile no longer matches
//            and the tile kernels skip it).  Every slot of the tile's range
//            is written, holes too, so that the stores cover whole lines
//            (buckets start on 32-slot boundaries).  The tile's holes, old
//            and new, are listed for the arrivals;
//   arrive   moves the outboxes' particles to their new tiles: into a
//            listed hole of the new tile while there is one, else appended
//            to its bucket (global atomic on bfill).  The buckets' slack
//            takes the arrivals that find no hole;
//   compact  when a bucket runs out of slack (the mover goes to the pending
//            list and raises the flag), and after a load (every particle
//            pending), the next step first rebuilds all buckets from the
//            occupied slots and the pending list into the other of two sets
//            of slots (count, capacity scan, placement), which becomes the
//            current one (ctl[3]).  Every compaction kernel is a no-op unless
//            the flag is set.  At 512^3, with the holes refilled, a
//            bucket's slack (count/2 + 64 slots) lasts past 25 steps of the
//            clustered benchmark ICs; in the strongly clustered regime (150
//            steps) a rebuild still comes every few steps.
// A tile processes only the slots it held when the step's deposit read it
// (bnow), so an arrival is not moved twice in one step.  Results do not
// depend on the slot order: the deposit sums exactly, the rest is per
// particle.
#include "pm_gather.h"
#include "pm_internal.h"
#include "pm_tiles.h"

namespace pm {

constexpr uint32_t kNoId = 0xFFFFFFFFu;

// both sets of slots, passed by value; a kernel takes set ctl[3]
template <typename Rl> struct StateSets {
    typename SplitWords<Rl>::Q* q[2];
    typename SplitWords<Rl>::U* u[2];
    uint32_t* id[2];
};
template <typename Rl> static StateSets<Rl> sets(const SplitState& st) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    return StateSets<Rl>{{static_cast<Q*>(st.q[0]), static_cast<Q*>(st.q[1])},
                         {static_cast<U*>(st.u[0]), static_cast<U*>(st.u[1])},
                         {st.id[0], st.id[1]}};
}

// FP64 position -> fixed point, nearest (x * 2^S is exact).  A position
// outside [0, ng) (or not finite) is counted as a range error and wrapped
// into the box (0 if not finite).
template <typename Q> __device__ __forceinline__ Q split_pos(double x, double box, double scale, bool& bad) {
    if (!(x >= 0.0 && x < box)) {
        bad = true;
        x = isfinite(x) ? x - floor(x / box) * box : 0.0;
        if (!(x >= 0.0 && x < box)) x = 0.0;
    }
    if constexpr (sizeof(Q) == 4)
        return (uint32_t)(unsigned long long)llrint(x * scale);  // 2^32 wraps to 0: the box is periodic
    else
        return (uint64_t)__double2ull_rn(x * scale);  // below 2^64 (x * 2^S exact, an integer above 2^52)
}

// load: every particle into the pending list (the first step compacts it
// into the buckets)
template <typename Rl>
__global__ __launch_bounds__(256) void k_state_load(const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ vx,
                                                    const double* __restrict__ vy, const double* __restrict__ vz,
                                                    int64_t np, int ng, typename SplitWords<Rl>::Q* __restrict__ pq,
                                                    typename SplitWords<Rl>::U* __restrict__ pu,
                                                    uint32_t* __restrict__ pid, unsigned long long* __restrict__ err) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= np) return;
    const double box = (double)ng, scale = ldexp(1.0, qshift<Q>(ng));
    bool bad = false;
    pq[sel(np, i, 0)] = split_pos<Q>(x[i], box, scale, bad);
    pq[sel(np, i, 1)] = split_pos<Q>(y[i], box, scale, bad);
    pq[sel(np, i, 2)] = split_pos<Q>(z[i], box, scale, bad);
    pu[sel(np, i, 0)] = (U)vx[i];
    pu[sel(np, i, 1)] = (U)vy[i];
    pu[sel(np, i, 2)] = (U)vz[i];
    pid[i] = (uint32_t)i;
    if (bad) atomicAdd(err, 1ULL);
}

// store: occupied bucket slots, then the pending list (non-empty only after
// an overflow that no step has compacted yet)
template <typename Rl>
__global__ __launch_bounds__(256) void k_state_store(StateSets<Rl> ss, int64_t cap,
                                                     const typename SplitWords<Rl>::Q* __restrict__ pq,
                                                     const typename SplitWords<Rl>::U* __restrict__ pu,
                                                     const uint32_t* __restrict__ pid, int64_t np,
                                                     const uint32_t* __restrict__ ctl, int ng,
                                                     double* __restrict__ x, double* __restrict__ y,
                                                     double* __restrict__ z, double* __restrict__ vx,
                                                     double* __restrict__ vy, double* __restrict__ vz) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    const int cs = (int)ctl[3];
    const Q* __restrict__ q = ss.q[cs];
    const U* __restrict__ u = ss.u[cs];
    const uint32_t* __restrict__ id = ss.id[cs];
    const double inv = ldexp(1.0, -qshift<Q>(ng));
    const int64_t npend = ctl[0];
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < cap + npend; k += (int64_t)gridDim.x * 256) {
        const bool st = k < cap;
        const int64_t s = st ? k : k - cap, stride = st ? cap : np;
        const Q* qq = st ? q : pq;
        const U* uu = st ? u : pu;
        const uint32_t i = st ? id[s] : pid[s];
        if (i == kNoId) continue;
        x[i] = (double)qq[sel(stride, s, 0)] * inv;
        y[i] = (double)qq[sel(stride, s, 1)] * inv;
        z[i] = (double)qq[sel(stride, s, 2)] * inv;
        vx[i] = (double)uu[sel(stride, s, 0)];
        vy[i] = (double)uu[sel(stride, s, 1)];
        vz[i] = (double)uu[sel(stride, s, 2)];
    }
}

// ------------------------------------------------------------ compaction
// ctl[0]: pending count, ctl[1]: compaction flag, ctl[2]: the flag that the
// capacity scan clears (unused), ctl[3]: the current set of slots
template <typename F>
__device__ __forceinline__ void for_each_entry(int64_t cap, const uint32_t* __restrict__ ctl, F f) {
    const int64_t n = cap + (int64_t)ctl[0];
    for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x; k0 < n; k0 += (int64_t)gridDim.x * blockDim.x)
        f(k0 + threadIdx.x, n);
}

// tile counts of the occupied slots and the pending entries
template <typename Rl>
__global__ __launch_bounds__(256) void k_state_count(StateSets<Rl> ss, int64_t cap,
                                                     const typename SplitWords<Rl>::Q* __restrict__ pq, int64_t np,
                                                     const uint32_t* __restrict__ ctl, int ng, int ts,
                                                     uint32_t* __restrict__ cnt) {
    using Q = typename SplitWords<Rl>::Q;
    if (ctl[1] == 0u) return;
    const Q* __restrict__ q = ss.q[ctl[3]];
    const uint32_t* __restrict__ id = ss.id[ctl[3]];
    const int lt = __builtin_ctz((unsigned)ts), sh = qshift<Q>(ng) + lt;
    const uint32_t nt = (uint32_t)(ng >> lt);
    const uint32_t lane = __lane_id();
    for_each_entry(cap, ctl, [&](int64_t k, int64_t n) {
        uint32_t kk = kNoId;
        if (k < cap) {
            if (id[k] != kNoId) kk = qkey<Q>(q[sel(cap, k, 0)], q[sel(cap, k, 1)], q[sel(cap, k, 2)], sh, nt);
        } else if (k < n) {
            const int64_t p = k - cap;
            kk = qkey<Q>(pq[sel(np, p, 0)], pq[sel(np, p, 1)], pq[sel(np, p, 2)], sh, nt);
        }
        // runs of equal tiles in consecutive lanes add once
        const uint32_t prev = __shfl_up(kk, 1);
        const bool head = kk != kNoId && (lane == 0 || prev != kk);
        const unsigned long long bounds = __ballot(head || kk == kNoId);
        if (head) {
            const unsigned long long above = lane == 63 ? 0ull : (bounds >> (lane + 1)) << (lane + 1);
            const uint32_t run = (above ? (uint32_t)__ffsll((long long)above) - 1u : 64u) - lane;
            atomicAdd(&cnt[kk], run);
        }
    });
}

// every entry into its tile's bucket of the other set of slots
template <typename Rl>
__global__ __launch_bounds__(256) void k_state_place(StateSets<Rl> ss, int64_t cap,
                                                     const typename SplitWords<Rl>::Q* __restrict__ pq,
                                                     const typename SplitWords<Rl>::U* __restrict__ pu,
                                                     const uint32_t* __restrict__ pid, int64_t np,
                                                     const uint32_t* __restrict__ ctl, int ng, int ts,
                                                     const uint32_t* __restrict__ boff, uint32_t* __restrict__ bfill) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    if (ctl[1] == 0u) return;
    const int cs = (int)ctl[3];
    const Q* __restrict__ q = ss.q[cs];
    const U* __restrict__ u = ss.u[cs];
    const uint32_t* __restrict__ id = ss.id[cs];
    Q* __restrict__ q2 = ss.q[cs ^ 1];
    U* __restrict__ u2 = ss.u[cs ^ 1];
    uint32_t* __restrict__ id2 = ss.id[cs ^ 1];
    const int lt = __builtin_ctz((unsigned)ts), sh = qshift<Q>(ng) + lt;
    const uint32_t nt = (uint32_t)(ng >> lt);
    const uint32_t lane = __lane_id();
    for_each_entry(cap, ctl, [&](int64_t k, int64_t n) {
        uint32_t kk = kNoId, i = kNoId;
        Q a = 0u, b = 0u, c = 0u;
        U v0 = 0, v1 = 0, v2 = 0;
        if (k < cap) {
            i = id[k];
            if (i != kNoId) {
                a = q[sel(cap, k, 0)];
                b = q[sel(cap, k, 1)];
                c = q[sel(cap, k, 2)];
                v0 = u[sel(cap, k, 0)];
                v1 = u[sel(cap, k, 1)];
                v2 = u[sel(cap, k, 2)];
            }
        } else if (k < n) {
            const int64_t p = k - cap;
            i = pid[p];
            a = pq[sel(np, p, 0)];
            b = pq[sel(np, p, 1)];
            c = pq[sel(np, p, 2)];
            v0 = pu[sel(np, p, 0)];
            v1 = pu[sel(np, p, 1)];
            v2 = pu[sel(np, p, 2)];
        }
        if (i != kNoId) kk = qkey<Q>(a, b, c, sh, nt);
        const uint32_t prev = __shfl_up(kk, 1);
        const bool head = kk != kNoId && (lane == 0 || prev != kk);
        const unsigned long long bounds = __ballot(head || kk == kNoId);
        uint32_t base = 0u;
        if (head) {
            const unsigned long long above = lane == 63 ? 0ull : (bounds >> (lane + 1)) << (lane + 1);
            const uint32_t run = (above ? (uint32_t)__ffsll((long long)above) - 1u : 64u) - lane;
            base = atomicAdd(&bfill[kk], run);
        }
        const unsigned long long upto = bounds & (lane == 63 ? ~0ull : (2ull << lane) - 1ull);
        const int hl = upto ? 63 - __clzll((long long)upto) : 0;
        const uint32_t slot = __shfl(base, hl) + (lane - (uint32_t)hl);
        if (kk != kNoId) {  // capacities are the counts + slack: every entry fits
            const int64_t s = (int64_t)boff[kk] + slot;
            q2[sel(cap, s, 0)] = a;
            q2[sel(cap, s, 1)] = b;
            q2[sel(cap, s, 2)] = c;
            u2[sel(cap, s, 0)] = v0;
            u2[sel(cap, s, 1)] = v1;
            u2[sel(cap, s, 2)] = v2;
            id2[s] = i;
        }
    });
}

// ids of every slot of a set to kNoId: the current set (a load), or behind
// the compaction flag the other set (before the placement)
template <typename Rl>
__global__ __launch_bounds__(256) void k_state_clear_ids(StateSets<Rl> ss, const uint32_t* __restrict__ ctl, int other,
                                                         int64_t n) {
    if (other && ctl[1] == 0u) return;
    uint32_t* __restrict__ id = ss.id[ctl[3] ^ (uint32_t)other];
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) id[k] = kNoId;
}

__global__ void k_state_done(uint32_t* __restrict__ ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && ctl[1] != 0u) {
        ctl[0] = 0u;
        ctl[1] = 0u;
        ctl[3] ^= 1u;  // the rebuilt set is the current one
    }
}

// ---------------------------------------------------------------- interp
// update_particles (utils.cpp:141-170) on a tile's particles: the force of
// cic_force (centered_derivatives + trilerp1, the arithmetic of pm_gather.h
// on the same 32 potential values, here read from the tile's LDS copy), the
// kick in the mesh precision, then the drift as the integer add of the
// displacement rounded to 2^-S cells.  A kick that leaves a velocity
// non-finite is counted as a range error and leaves that particle as it was.
// FP32: 1024 threads a tile, at most 64 VGPRs (eight waves per SIMD, two
// workgroups per CU, 27 KB of LDS each); FP64: 512 threads a tile, at most
// 128 VGPRs (four waves per SIMD, two workgroups per CU, 55 KB of LDS each).
#ifndef PM_TILE_THREADS
#define PM_TILE_THREADS 1024
#endif
#ifndef PM_TILE_THREADS64
#define PM_TILE_THREADS64 512
#endif
#ifndef PM_TILE_BATCH
#define PM_TILE_BATCH 1
#endif
constexpr int kTileThreads = PM_TILE_THREADS, kTileBatch = PM_TILE_BATCH;
template <typename Rl> constexpr int tile_threads() { return sizeof(Rl) == 4 ? kTileThreads : PM_TILE_THREADS64; }

// per tile, where its leavers go and the kernels that read them: position,
// velocity and particle index of each, and their count
template <typename Rl> struct Outbox {
    typename SplitWords<Rl>::Q* q;
    typename SplitWords<Rl>::U* u;
    uint32_t* i;
    uint32_t* cnt;
};
template <typename Rl> static Outbox<Rl> outbox(const SplitState& st) {
    return Outbox<Rl>{static_cast<typename SplitWords<Rl>::Q*>(st.oq), static_cast<typename SplitWords<Rl>::U*>(st.ou),
                      st.oid, st.ocnt};
}
// per tile, the holes of its range that the last interpolation saw (slots
// whose particle left, in this step or before), their count, and how many
// of them k_state_arrive has given to arrivals
struct Holes {
    uint32_t* s;
    uint32_t* cnt;
    uint32_t* used;
};
static Holes holes(const SplitState& st) { return Holes{st.hole, st.hcnt, st.huse}; }

// where a particle that leaves its tile is appended: the new tile's bucket,
// or, if that is full, the pending list (which makes the next step compact)
template <typename Rl> struct Dest {
    const uint32_t* boff;
    uint32_t* bfill;
    typename SplitWords<Rl>::Q* pq;
    typename SplitWords<Rl>::U* pu;
    uint32_t* pid;
    int64_t np;
    uint32_t* ctl;
};
template <typename Rl> static Dest<Rl> dest(const SplitState& st, int64_t np) {
    return Dest<Rl>{st.boff, st.bfill, static_cast<typename SplitWords<Rl>::Q*>(st.pq),
                    static_cast<typename SplitWords<Rl>::U*>(st.pu), st.pid, np, st.ctl};
}
template <typename Rl, typename Q = typename SplitWords<Rl>::Q, typename U = typename SplitWords<Rl>::U>
__device__ __forceinline__ void state_append(const Dest<Rl>& d, Q* __restrict__ q, U* __restrict__ u,
                                             uint32_t* __restrict__ id, int64_t cap, uint32_t t2, Q na, Q nb, Q nc,
                                             U n0, U n1, U n2, uint32_t i) {
    const uint32_t slot = atomicAdd(&d.bfill[t2], 1u);
    const int64_t c2 = (int64_t)d.boff[t2 + 1] - d.boff[t2];
    Q* dq = q;
    uint32_t* di = id;
    U* du = u;
    int64_t ds, stride = cap;
    if ((int64_t)slot < c2) {
        ds = (int64_t)d.boff[t2] + slot;
    } else {
        ds = atomicAdd(&d.ctl[0], 1u);  // pending entries never exceed np
        dq = d.pq;
        du = d.pu;
        di = d.pid;
        stride = d.np;
        d.ctl[1] = 1u;
    }
    dq[sel(stride, ds, 0)] = na;
    dq[sel(stride, ds, 1)] = nb;
    dq[sel(stride, ds, 2)] = nc;
    du[sel(stride, ds, 0)] = n0;
    du[sel(stride, ds, 1)] = n1;
    du[sel(stride, ds, 2)] = n2;
    di[ds] = i;
}

// (the sets by reference: indexed by the current set in the kernel's
// arguments, their pointers stay global; a by-value copy indexed at run time
// would leave flat loads and stores)
template <typename Rl, int TS, bool DRIFT>
__device__ __forceinline__ void state_interp(const StateSets<Rl>& ss, int64_t cap, const uint32_t* __restrict__ boff,
                                             const uint32_t* __restrict__ bnow, const Dest<Rl>& dst,
                                             const Outbox<Rl>& ob, const Holes& ho,
                                             const uint32_t* __restrict__ ctl, int ng, const Rl* __restrict__ phi,
                                             Rl kick, Rl dscale, unsigned long long* __restrict__ err) {
    using Q = typename SplitWords<Rl>::Q;
    using U = Rl;
    constexpr int NT = tile_threads<Rl>();
    constexpr int W = TS + 3;  // nodes -1 .. TS+1 around the tile's cells
    __shared__ U ph[W * W * W];
    __shared__ uint32_t nmov, nhole;
    const int cs = (int)ctl[3];
    Q* __restrict__ q = ss.q[cs];
    U* __restrict__ u = ss.u[cs];
    uint32_t* __restrict__ id = ss.id[cs];
    const int lt = ilog2(TS), S = qshift<Q>(ng), sh = S + lt;
    const uint32_t nt = (uint32_t)(ng >> lt);
    const uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = (int)(t / (nt * nt)), ty = (int)((t / nt) % nt), tz = (int)(t % nt);
    const int m = ng - 1;
    const int64_t beg = boff[t], end = beg + bnow[t];
    if (threadIdx.x == 0) {
        nmov = 0u;
        nhole = 0u;
    }
    constexpr int64_t kStep = (int64_t)NT * kTileBatch;
    Q qa[kTileBatch], qb[kTileBatch], qc[kTileBatch];
    U ua[kTileBatch], ub[kTileBatch], uc[kTileBatch];
    // a batch's slots past the end read the tile's first slot (every bucket
    // has at least 64 slots) and are skipped
    const auto load_batch = [&](int64_t s0) {
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
            const int64_t s = s0 + (int64_t)NT * j;
            const int64_t r = s < end ? s : beg;
            qa[j] = stream_load(q + sel(cap, r, 0));
            qb[j] = stream_load(q + sel(cap, r, 1));
            qc[j] = stream_load(q + sel(cap, r, 2));
            ua[j] = stream_load(u + sel(cap, r, 0));
            ub[j] = stream_load(u + sel(cap, r, 1));
            uc[j] = stream_load(u + sel(cap, r, 2));
        }
    };
    // all of a thread's staging loads in flight at once, then the first
    // batch of particles, then the LDS writes: the two latencies overlap
    constexpr int NE = W * W * W, NIT = (NE + NT - 1) / NT;
    U pv[NIT];
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
        const int e = threadIdx.x + j * NT;
        if (e < NE) {
            const int a = e / (W * W), r = e - a * W * W, b = r / W, c = r - b * W;
            const int gx = (tx * TS - 1 + a) & m, gy = (ty * TS - 1 + b) & m, gz = (tz * TS - 1 + c) & m;
            pv[j] = phi[((size_t)gx * ng + gy) * ng + gz];
        }
    }
    int64_t s0 = beg + threadIdx.x;
    load_batch(s0);
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
        const int e = threadIdx.x + j * NT;
        if (e < NE) ph[e] = pv[j];
    }
    __syncthreads();
    const Q fmask = ((Q)1 << S) - 1u;
    U fs;
    if constexpr (sizeof(U) == 4)
        fs = ldexpf(1.0f, -S);
    else
        fs = ldexp(1.0, -S);
    constexpr Q kHole = (Q)1 << (8 * sizeof(Q) - 1);  // x in the other half of the box
    unsigned nerr = 0u;
    while (s0 < end) {
#pragma unroll
        for (int j = 0; j < kTileBatch; ++j) {
            const int64_t s = s0 + (int64_t)NT * j;
            if (s >= end) continue;
            const Q a = qa[j], b = qb[j], c = qc[j];
            // what the slot gets back: a hole (another tile's key) keeps its
            // values.  Every slot of the range is written, so that the
            // stores cover whole 128-byte lines (a partial line costs the
            // memory a read-modify-write).
            Q wa = a, wb = b, wc = c;
            U w0 = ua[j], w1 = ub[j], w2 = uc[j];
            bool hole = DRIFT;  // a hole already, unless its particle is the tile's
            if (qkey<Q>(a, b, c, sh, nt) == t) {
                hole = false;
                const int lx = (int)(a >> S) - tx * TS, ly = (int)(b >> S) - ty * TS, lz = (int)(c >> S) - tz * TS;
                const U fx = (U)(a & fmask) * fs, fy = (U)(b & fmask) * fs, fz = (U)(c & fmask) * fs;
                // the 12 z-runs of gather_runs_main from LDS: node (ix+x,
                // iy+y, iz+z) of the cell at local (lx, ly, lz) is
                // L(lx+1+x, ly+1+y, lz+1+z)
                U cl[2][2][4], xr[2][2][2], yr[2][2][2];
                const auto L = [&](int x, int y, int z) { return ph[(x * W + y) * W + z]; };
#pragma unroll
                for (int aa = 0; aa < 2; ++aa)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) cl[aa][bb][k] = L(lx + 1 + aa, ly + 1 + bb, lz + k);
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            xr[aa][bb][k] = L(lx + 3 * aa, ly + 1 + bb, lz + 1 + k);
                            yr[aa][bb][k] = L(lx + 1 + bb, ly + 3 * aa, lz + 1 + k);
                        }
                    }
                U f[3];
                force_from_runs(cl, xr, yr, fx, fy, fz, f);
                U n0 = fma(f[0], kick, ua[j]), n1 = fma(f[1], kick, ub[j]), n2 = fma(f[2], kick, uc[j]);
                const bool fin = isfinite(n0) && isfinite(n1) && isfinite(n2);
                if (!fin) {
                    n0 = ua[j];
                    n1 = ub[j];
                    n2 = uc[j];
                    ++nerr;
                }
                Q na = a, nb = b, nc = c;
                if (DRIFT && fin) {
                    if constexpr (sizeof(Q) == 4) {
                        na = a + (uint32_t)__float2int_rn(n0 * dscale);
                        nb = b + (uint32_t)__float2int_rn(n1 * dscale);
                        nc = c + (uint32_t)__float2int_rn(n2 * dscale);
                    } else {
                        na = a + (uint64_t)__double2ll_rn(n0 * dscale);
                        nb = b + (uint64_t)__double2ll_rn(n1 * dscale);
                        nc = c + (uint64_t)__double2ll_rn(n2 * dscale);
                    }
                }
                w0 = n0;
                w1 = n1;
                w2 = n2;
                const uint32_t t2 = DRIFT ? qkey<Q>(na, nb, nc, sh, nt) : t;
                if (t2 == t) {
                    wa = na;
                    wb = nb;
                    wc = nc;
                } else {
                    // a mover: its slot becomes a hole (x in the other half of
                    // the box: another tile); the particle, with its index,
                    // goes into the tile's outbox (an LDS counter, no global
                    // round trip in the loop), which k_state_arrive moves to
                    // its new tile.  A full outbox: appended here.
                    wa = a ^ kHole;
                    hole = true;
                    const uint32_t pid = id[s];
                    id[s] = kNoId;
                    const uint32_t k = atomicAdd(&nmov, 1u);
                    if (k < (uint32_t)kOutbox) {
                        const int64_t e = (int64_t)t * kOutbox + k, no = (int64_t)gridDim.x * kOutbox;
                        ob.q[sel(no, e, 0)] = na;
                        ob.q[sel(no, e, 1)] = nb;
                        ob.q[sel(no, e, 2)] = nc;
                        ob.u[sel(no, e, 0)] = n0;
                        ob.u[sel(no, e, 1)] = n1;
                        ob.u[sel(no, e, 2)] = n2;
                        ob.i[e] = pid;
                    } else {
                        state_append<Rl>(dst, q, u, id, cap, t2, na, nb, nc, n0, n1, n2, pid);
                    }
                }
            }
            if (hole) {  // listed for the arrivals of the next k_state_arrive
                const uint32_t k = atomicAdd(&nhole, 1u);
                if (k < (uint32_t)kOutbox) ho.s[(int64_t)t * kOutbox + k] = (uint32_t)s;
            }
            stream_store(u + sel(cap, s, 0), w0);
            stream_store(u + sel(cap, s, 1), w1);
            stream_store(u + sel(cap, s, 2), w2);
            if (DRIFT) {
                stream_store(q + sel(cap, s, 0), wa);
                stream_store(q + sel(cap, s, 1), wb);
                stream_store(q + sel(cap, s, 2), wc);
            }
        }
        s0 += kStep;
        if (s0 < end) load_batch(s0);
    }
    if (nerr) atomicAdd(err, (unsigned long long)nerr);
    if (DRIFT) {
        __syncthreads();
        if (threadIdx.x == 0) {
            ob.cnt[t] = min(nmov, (uint32_t)kOutbox);
            ho.cnt[t] = min(nhole, (uint32_t)kOutbox);
            ho.used[t] = 0u;
        }
    }
}
template <int TS, bool DRIFT>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_state_interp(
    StateSets<float> ss, int64_t cap, const uint32_t* __restrict__ boff, const uint32_t* __restrict__ bnow,
    Dest<float> dst, Outbox<float> ob, Holes ho, const uint32_t* __restrict__ ctl, int ng,
    const float* __restrict__ phi, float kick, float dscale, unsigned long long* __restrict__ err) {
    state_interp<float, TS, DRIFT>(ss, cap, boff, bnow, dst, ob, ho, ctl, ng, phi, kick, dscale, err);
}
template <int TS, bool DRIFT>
__global__ __launch_bounds__(PM_TILE_THREADS64) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_state_interp64(
    StateSets<double> ss, int64_t cap, const uint32_t* __restrict__ boff, const uint32_t* __restrict__ bnow,
    Dest<double> dst, Outbox<double> ob, Holes ho, const uint32_t* __restrict__ ctl, int ng,
    const double* __restrict__ phi, double kick, double dscale, unsigned long long* __restrict__ err) {
    state_interp<double, TS, DRIFT>(ss, cap, boff, bnow, dst, ob, ho, ctl, ng, phi, kick, dscale, err);
}

// The leavers of the last interpolation: one wave per tile moves its outbox
// entries to their new tiles, into a hole of the new tile's range if the
// last interpolation listed one that no arrival has taken yet (the buckets
// then keep their length), else appended to the bucket; then empties the
// outbox, so that a second call is a no-op.  The interpolation's own appends
// (full outbox) never take holes: a hole in a range that the same
// interpolation walks could be moved twice in one step.
template <typename Rl, int TS>
__global__ __launch_bounds__(256) void k_state_arrive(StateSets<Rl> ss, int64_t cap, Dest<Rl> dst, Outbox<Rl> ob,
                                                      Holes ho, int ng, uint32_t ntiles) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    const uint32_t t = blockIdx.x * 4 + threadIdx.x / 64;
    if (t >= ntiles) return;
    const int cs = (int)dst.ctl[3];
    Q* __restrict__ q = ss.q[cs];
    U* __restrict__ u = ss.u[cs];
    uint32_t* __restrict__ id = ss.id[cs];
    const int lt = ilog2(TS), sh = qshift<Q>(ng) + lt;
    const uint32_t nt = (uint32_t)(ng >> lt);
    const uint32_t n = ob.cnt[t];
    const int64_t no = (int64_t)ntiles * kOutbox;
    for (uint32_t k = threadIdx.x & 63u; k < n; k += 64u) {
        const int64_t e = (int64_t)t * kOutbox + k;
        const Q na = ob.q[sel(no, e, 0)], nb = ob.q[sel(no, e, 1)], nc = ob.q[sel(no, e, 2)];
        const U n0 = ob.u[sel(no, e, 0)], n1 = ob.u[sel(no, e, 1)], n2 = ob.u[sel(no, e, 2)];
        const uint32_t pid = ob.i[e], t2 = qkey<Q>(na, nb, nc, sh, nt);
        const uint32_t h = atomicAdd(&ho.used[t2], 1u);
        if (h < ho.cnt[t2]) {
            const int64_t s = ho.s[(int64_t)t2 * kOutbox + h];
            q[sel(cap, s, 0)] = na;
            q[sel(cap, s, 1)] = nb;
            q[sel(cap, s, 2)] = nc;
            u[sel(cap, s, 0)] = n0;
            u[sel(cap, s, 1)] = n1;
            u[sel(cap, s, 2)] = n2;
            id[s] = pid;
        } else {
            state_append<Rl>(dst, q, u, id, cap, t2, na, nb, nc, n0, n1, n2, pid);
        }
    }
    if ((threadIdx.x & 63u) == 0u && n) ob.cnt[t] = 0u;
}

// ================================================================ launchers
static unsigned grid_for(int64_t n, unsigned cap_blocks) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap_blocks));
}

template <typename Rl>
static int state_load(const double* x, const double* y, const double* z, const double* vx, const double* vy,
                      const double* vz, int64_t np, int ng, SplitState& st, unsigned long long* err, hipStream_t s) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    hipLaunchKernelGGL(k_state_clear_ids<Rl>, dim3(grid_for(st.cap, 16384)), dim3(256), 0, s, sets<Rl>(st), st.ctl, 0,
                       st.cap);
    hipLaunchKernelGGL(k_state_load<Rl>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, x, y, z, vx, vy, vz, np,
                       ng, static_cast<Q*>(st.pq), static_cast<U*>(st.pu), st.pid, err);
    // every particle pending, no outbox entries; the next step compacts
    const int ts = sorted_tile(ng);
    const size_t ntiles = (size_t)(ng / ts) * (ng / ts) * (ng / ts);
    // outbox counts, hole counts and their use (contiguous): none
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3((unsigned)std::min<size_t>((3 * ntiles + 255) / 256, 1024)), dim3(256), 0,
                       s, st.ocnt, 3 * ntiles, 0u);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, s, st.ctl, (size_t)1, (uint32_t)np);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, s, st.ctl + 1, (size_t)1, 1u);
    return check_launch("split_load");
}
int launch_state_load(const double* x, const double* y, const double* z, const double* vx, const double* vy,
                      const double* vz, int64_t np, int ng, SplitState& st, unsigned long long* err, hipStream_t s) {
    return st.prec == PM_FP64 ? state_load<double>(x, y, z, vx, vy, vz, np, ng, st, err, s)
                              : state_load<float>(x, y, z, vx, vy, vz, np, ng, st, err, s);
}

template <typename Rl>
static int state_store(const SplitState& st, int64_t np, int ng, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, hipStream_t s) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    hipLaunchKernelGGL(k_state_store<Rl>, dim3(grid_for(st.cap + np, 16384)), dim3(256), 0, s, sets<Rl>(st), st.cap,
                       static_cast<const Q*>(st.pq), static_cast<const U*>(st.pu), st.pid, np, st.ctl, ng, x, y, z, vx,
                       vy, vz);
    return check_launch("split_store");
}
int launch_state_store(const SplitState& st, int64_t np, int ng, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, hipStream_t s) {
    return st.prec == PM_FP64 ? state_store<double>(st, np, ng, x, y, z, vx, vy, vz, s)
                              : state_store<float>(st, np, ng, x, y, z, vx, vy, vz, s);
}

template <typename Rl> static int state_compact(SplitState& st, int64_t np, int ng, SortedScratch& s, hipStream_t hs) {
    using Q = typename SplitWords<Rl>::Q;
    using U = typename SplitWords<Rl>::U;
    const int ts = sorted_tile(ng);
    const int nt = ng / ts;
    const int64_t n = st.cap + np;
    // the tile counts (s.cnt) are zero here: cleared at allocation and by
    // every count-consuming scan
    // capped grids (grid-stride loops): in a step without compaction every
    // kernel here is a no-op, and a workgroup that only exits still costs
    // its launch
    hipLaunchKernelGGL(k_state_count<Rl>, dim3(grid_for(n, 1024)), dim3(256), 0, hs, sets<Rl>(st), st.cap,
                       static_cast<const Q*>(st.pq), np, st.ctl, ng, ts, s.cnt);
    int rc = check_launch("split_count");
    if (rc) return rc;
    // capacities (count + count/2 + 64, rounded up to 32) into boff, bfill cleared
    rc = launch_cap_scan_raw(nt * nt * nt, st.bfill, s.cnt, st.boff, st.ctl + 2, st.ctl + 1, hs);
    if (rc) return rc;
    hipLaunchKernelGGL(k_state_clear_ids<Rl>, dim3(grid_for(st.cap, 1024)), dim3(256), 0, hs, sets<Rl>(st), st.ctl, 1,
                       st.cap);
    hipLaunchKernelGGL(k_state_place<Rl>, dim3(grid_for(n, 1024)), dim3(256), 0, hs, sets<Rl>(st), st.cap,
                       static_cast<const Q*>(st.pq), static_cast<const U*>(st.pu), st.pid, np, st.ctl, ng, ts, st.boff,
                       st.bfill);
    hipLaunchKernelGGL(k_state_done, dim3(1), dim3(64), 0, hs, st.ctl);
    return check_launch("split_compact");
}
int launch_state_compact(SplitState& st, int64_t np, int ng, SortedScratch& s, hipStream_t hs) {
    return st.prec == PM_FP64 ? state_compact<double>(st, np, ng, s, hs) : state_compact<float>(st, np, ng, s, hs);
}

int launch_state_interp(SplitState& st, int64_t np, int ng, const void* phi, double kick, double drift,
                        unsigned long long* err, hipStream_t hs) {
    const int ts = sorted_tile(ng);
    const unsigned ntiles = (unsigned)((ng / ts) * (ng / ts) * (ng / ts));
#define PM_SI(K, RL, TSV, DR, NT)                                                                                    \
    hipLaunchKernelGGL((K<TSV, DR>), dim3(ntiles), dim3(NT), 0, hs, sets<RL>(st), st.cap, st.boff, st.bnow,          \
                       dest<RL>(st, np), outbox<RL>(st), holes(st), st.ctl, ng, static_cast<const RL*>(phi), (RL)kick, \
                       (RL)ldexp(drift, qshift<typename SplitWords<RL>::Q>(ng)), err)
#define PM_SI_ALL(K, RL, NT)            \
    if (ts == 16) {                     \
        if (drift != 0.0)               \
            PM_SI(K, RL, 16, true, NT);  \
        else                            \
            PM_SI(K, RL, 16, false, NT); \
    } else {                            \
        if (drift != 0.0)               \
            PM_SI(K, RL, 8, true, NT);   \
        else                            \
            PM_SI(K, RL, 8, false, NT);  \
    }
    if (st.prec == PM_FP64) {
        PM_SI_ALL(k_state_interp64, double, PM_TILE_THREADS64)
    } else {
        PM_SI_ALL(k_state_interp, float, kTileThreads)
    }
#undef PM_SI_ALL
#undef PM_SI
    return check_launch("split_interp");
}

template <typename Rl> static int state_arrive(SplitState& st, int64_t np, int ng, hipStream_t hs) {
    const int ts = sorted_tile(ng);
    const unsigned ntiles = (unsigned)((ng / ts) * (ng / ts) * (ng / ts));
    const dim3 grid((ntiles + 3) / 4);
    if (ts == 16)
        hipLaunchKernelGGL((k_state_arrive<Rl, 16>), grid, dim3(256), 0, hs, sets<Rl>(st), st.cap, dest<Rl>(st, np),
                           outbox<Rl>(st), holes(st), ng, ntiles);
    else
        hipLaunchKernelGGL((k_state_arrive<Rl, 8>), grid, dim3(256), 0, hs, sets<Rl>(st), st.cap, dest<Rl>(st, np),
                           outbox<Rl>(st), holes(st), ng, ntiles);
    return check_launch("split_arrive");
}
int launch_state_arrive(SplitState& st, int64_t np, int ng, hipStream_t hs) {
    return st.prec == PM_FP64 ? state_arrive<double>(st, np, ng, hs) : state_arrive<float>(st, np, ng, hs);
}

}  // namespace pm
