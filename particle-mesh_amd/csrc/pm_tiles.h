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

}  // namespace pm
