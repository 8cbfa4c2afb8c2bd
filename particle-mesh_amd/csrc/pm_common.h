// Shared device helpers for the PM kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace pm {

constexpr int kXcds = 8;  // MI355X: 256 CUs in 8 XCDs, one L2 each

// Workgroups are dispatched round-robin over the XCDs (block b -> XCD b % 8).
// Remap so that each XCD walks one contiguous range of logical blocks, which
// keeps neighbouring particles / mesh lines in the same L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
    const uint32_t q = nb / kXcds, r = nb % kXcds;
    const uint32_t xcd = bid % kXcds, idx = bid / kXcds;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Block order of the particle kernels.  Particles are usually stored in
// Lagrangian order (the lattice of the initial conditions, np = L^3), so a
// block of consecutive particles gathers from a small patch of the mesh.
// Swept plane by plane, one mesh plane serves the particles of four lattice
// planes, a whole plane apart, and drops out of the XCD's L2 in between.
// With a band map each XCD sweeps its range of whole lattice planes in bands
// of rows instead: band by band, each band through all the XCD's planes,
// so the mesh rows around a band stay in L2 while the next planes use them.
// A permutation of the blocks: any particle order gives the same results.
struct BandMap {
    uint32_t bpp, bb;  // blocks per lattice plane, blocks per band; bb == 0: plain order
};
__device__ __forceinline__ uint32_t band_remap(uint32_t bid, uint32_t nb, BandMap m) {
    const uint32_t l = xcd_remap(bid, nb);
    if (m.bb == 0) return l;
    // the host checked nb % kXcds == 0, len % bpp == 0 and bpp % bb == 0
    const uint32_t len = nb / kXcds;
    if (((len & (len - 1)) | (m.bpp & (m.bpp - 1)) | (m.bb & (m.bb - 1))) == 0) {
        // powers of two (lattices of 2^k): shifts instead of the four
        // divisions below, which each wave would otherwise run before its
        // first load (interpolation at 512^3 FP32, A/B: 3.17 -> 3.16 ms)
        const uint32_t lb = __builtin_ctz(m.bb), lp = __builtin_ctz(len) - __builtin_ctz(m.bpp);
        const uint32_t r = l & (len - 1), s = l - r;
        const uint32_t o = r & (m.bb - 1), q = r >> lb;
        return s + (q & ((1u << lp) - 1u)) * m.bpp + (q >> lp) * m.bb + o;
    }
    const uint32_t r = l % len, s = l - r;  // r: dispatch order in the XCD
    const uint32_t planes = len / m.bpp, o = r % m.bb, q = r / m.bb;
    return s + (q % planes) * m.bpp + (q / planes) * m.bb + o;
}

// Position inside the periodic box [0, box)^3 (false for NaN).  The six
// compares are combined with bitwise &: with a short-circuit && the compiler
// sinks the y and z loads behind the x test, and the three loads of a
// position then wait on each other.
__device__ __forceinline__ bool in_box3(double x, double y, double z, double box) {
    return (x >= 0.0) & (x < box) & (y >= 0.0) & (y < box) & (z >= 0.0) & (z < box);
}

// Loads and stores of particle streams that a kernel reads once and writes
// once: non-temporal, so that they pass through L2 without evicting the
// mesh planes that the gathers of neighbouring particle rows re-read (the
// interpolation streams 13 GB per launch at 512^3 particles).  Interpolation
// in an interleaved A/B at 512^3: 3.86 -> 3.40 ms (FP32 mesh), 5.41 -> 4.95
// ms (FP64); stores alone 3.81, loads alone 3.61.  Bit 0: loads, 1: stores.
#ifndef PM_STREAM_NT
#define PM_STREAM_NT 3
#endif
template <typename T> __device__ __forceinline__ T stream_load(const T* p) {
    if constexpr (PM_STREAM_NT & 1) return __builtin_nontemporal_load(p);
    return *p;
}
template <typename T> __device__ __forceinline__ void stream_store(T* p, T v) {
    if constexpr (PM_STREAM_NT & 2)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Nearest integer (ties to even) of a double with |v| < 2^51, as a signed
// 64-bit integer: v + 1.5 * 2^52 lands in [2^52, 2^53), where the doubles are
// the integers, so that sum's bit pattern minus that of 1.5 * 2^52 is the
// rounded value -- the same result as __double2ll_rn, in two instructions
// instead of its conversion sequence (the fixed-point deposits convert 8 to
// 32 weights per particle).  The add is kept apart from the product that
// usually feeds it, so v is rounded to a double first, as for __double2ll_rn.
__device__ __forceinline__ long long fix_rne(double v) {
#pragma clang fp contract(off)
    const double t = v + 6755399441055744.0;  // 1.5 * 2^52
    return __double_as_longlong(t) - 0x4338000000000000LL;
}

template <typename R> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename R> using cplx_t = typename vec2<R>::type;

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
// multiply by -i (DIR=-1) or +i (DIR=+1)
template <int DIR, typename C> __device__ __forceinline__ C mul_i(C a) {
    return DIR < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

}  // namespace pm
