// Adjoint (reverse-mode) kernels of the PM step.  The reference has no
// gradients; these differentiate the forward step of this library:
//
//   rho = m * sum_p W(x_p, .)            CIC deposit (utils.cpp:70-102)
//   phi = S (rho - 1)                     FFT solve, S symmetric (G real, even)
//   g_a = D_a phi                         centred difference (utils.cpp:130-132)
//   F_p = sum_c W(x_p, c) g[c]            trilinear interpolation (utils_inline.h:55-70)
//
// For a per-particle cotangent u, the vector-Jacobian product is
//   (J_F^T u)_q = sum_a u_q^a grad_x [sum_c W(x_q, c) g_a[c]]        (term A)
//               + m grad_x [sum_c W(x_q, c) lambda[c]]               (term B)
//   lambda = S mu,  mu = sum_a D_a^T U_a,  U_a = sum_p W(x_p, .) u_p^a,
//   (D_a^T U)[c] = (U[c + e_a] - U[c - e_a]) / 2.
// The CIC weights are piecewise linear: grad_x W(x, corner) takes the +-1
// slope along the axis and the linear weights along the other two.
#include <type_traits>

#include "pm_gather.h"
#include "pm_internal.h"
#include "pm_tiles.h"

namespace pm {

// ---------------------------------------------------------------- helpers
template <typename Rl> struct Corner {
    int x[2], y[2], z[2];  // periodic node indices
    Rl tx, ty, tz;         // fractional offsets
};

template <typename Rl>
__device__ __forceinline__ void corner_setup(double x, double y, double z, int N, int (&xa)[4], int (&ya)[4],
                                             int (&za)[4], Rl& tx, Rl& ty, Rl& tz) {
    const int ix = (int)x, iy = (int)y, iz = (int)z;
    tx = (Rl)(x - ix);
    ty = (Rl)(y - iy);
    tz = (Rl)(z - iz);
    xa[0] = ix == 0 ? N - 1 : ix - 1;
    xa[1] = ix;
    xa[2] = ix + 1 == N ? 0 : ix + 1;
    xa[3] = ix + 2 >= N ? ix + 2 - N : ix + 2;
    ya[0] = iy == 0 ? N - 1 : iy - 1;
    ya[1] = iy;
    ya[2] = iy + 1 == N ? 0 : iy + 1;
    ya[3] = iy + 2 >= N ? iy + 2 - N : iy + 2;
    za[0] = iz == 0 ? N - 1 : iz - 1;
    za[1] = iz;
    za[2] = iz + 1 == N ? 0 : iz + 1;
    za[3] = iz + 2 >= N ? iz + 2 - N : iz + 2;
}

// value and gradient (d/dx, d/dy, d/dz) of the trilinear interpolant
template <typename Rl>
__device__ __forceinline__ void tri_val_grad(const Rl (&f)[2][2][2], Rl tx, Rl ty, Rl tz, Rl& val, Rl (&gr)[3]) {
    const Rl mx = Rl(1) - tx, my = Rl(1) - ty, mz = Rl(1) - tz;
    // trilerp1 order (utils_inline.h:55-70) for the value
    const Rl f00 = f[0][0][0] * mx + f[1][0][0] * tx;
    const Rl f10 = f[0][1][0] * mx + f[1][1][0] * tx;
    const Rl fz0 = f00 * my + f10 * ty;
    const Rl g00 = f[0][0][1] * mx + f[1][0][1] * tx;
    const Rl g10 = f[0][1][1] * mx + f[1][1][1] * tx;
    const Rl fz1 = g00 * my + g10 * ty;
    val = fz0 * mz + fz1 * tz;
    // d/dx: slope +-1 along x
    const Rl dx00 = f[1][0][0] - f[0][0][0], dx10 = f[1][1][0] - f[0][1][0];
    const Rl dx01 = f[1][0][1] - f[0][0][1], dx11 = f[1][1][1] - f[0][1][1];
    gr[0] = (dx00 * my + dx10 * ty) * mz + (dx01 * my + dx11 * ty) * tz;
    // d/dy
    gr[1] = (f10 - f00) * mz + (g10 - g00) * tz;
    // d/dz
    gr[2] = fz1 - fz0;
}

// centred differences of phi at the 8 corners (as in k_interp)
template <typename Rl>
__device__ __forceinline__ void corner_gradients(const Rl* __restrict__ phi, int N, const int (&xa)[4],
                                                 const int (&ya)[4], const int (&za)[4], Rl (&gx)[2][2][2],
                                                 Rl (&gy)[2][2][2], Rl (&gz)[2][2][2]) {
    auto P = [&](int a, int b, int c) -> Rl { return phi[((size_t)xa[a] * N + ya[b]) * N + za[c]]; };
    const Rl h = Rl(0.5);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                gx[a][b][c] = (P(a, b + 1, c + 1) - P(a + 2, b + 1, c + 1)) * h;
                gy[a][b][c] = (P(a + 1, b, c + 1) - P(a + 1, b + 2, c + 1)) * h;
                gz[a][b][c] = (P(a + 1, b + 1, c) - P(a + 1, b + 1, c + 2)) * h;
            }
}

// ------------------------------------------------- adjoint of the deposit
// xbar_q += scale * grad_x sum_c W(x_q, c) mbar[c]   (transposed CIC)
template <typename Rl>
__global__ __launch_bounds__(256) void k_adj_deposit(const double* __restrict__ px, const double* __restrict__ py,
                                                     const double* __restrict__ pz, int64_t np, int N,
                                                     const Rl* __restrict__ mbar, double scale,
                                                     double* __restrict__ bx, double* __restrict__ by,
                                                     double* __restrict__ bz, unsigned long long* __restrict__ err) {
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= np) return;
    const double x = px[i], y = py[i], z = pz[i];
    const double box = (double)N;
    if (!in_box3(x, y, z, box)) {
        atomicAdd(err, 1ULL);
        return;
    }
    int xa[4], ya[4], za[4];
    Rl tx, ty, tz;
    corner_setup<Rl>(x, y, z, N, xa, ya, za, tx, ty, tz);
    Rl f[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c) f[a][b][c] = mbar[((size_t)xa[a + 1] * N + ya[b + 1]) * N + za[c + 1]];
    Rl v, gr[3];
    tri_val_grad(f, tx, ty, tz, v, gr);
    bx[i] += scale * (double)gr[0];
    by[i] += scale * (double)gr[1];
    bz[i] += scale * (double)gr[2];
}

// ------------------------------------------- interpolation transpose
// U_a[c] += W(x_p, c) u_p^a for a = 0..2, one workgroup per segment of the
// column walk over the tiles of the sorted binning.  The cotangents have arbitrary sign and scale, so each
// component is accumulated in signed 64-bit fixed point with a power-of-two
// scale 2^40 / 2^ceil(log2 max|u_a|) taken from a global max (k_absmax3):
// integer LDS adds are exact and order-free (LDS float atomics serialise on
// this part).  Owned nodes, xy shells and z seams are written as in the
// forward deposit, and folded by the same kernel.
constexpr int kVec3Threads = 1024;

__device__ __forceinline__ double vec3_scale(unsigned long long maxbits) {
    const double m = __longlong_as_double((long long)maxbits);
    if (!(m > 0.0) || !isfinite(m)) return 1.0;
    int e = 0;
    frexp(m, &e);  // m < 2^e
    return ldexp(1.0, 40 - e);
}

// Particles in flight per thread.  Measured while the reverse sweep still
// used this deposit (it now deposits the divergence directly), full sweep of
// 20 steps at 512^3 in an interleaved A/B: 260.5 ms (2), 265.7 (4), 279.5
// (8).
#ifndef PM_VEC3_BATCH
#define PM_VEC3_BATCH 2
#endif
template <typename Rl, int TS>
__global__ __launch_bounds__(kVec3Threads) void k_deposit_vec3_tiles(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ ux, const double* __restrict__ uy, const double* __restrict__ uz,
    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off, int ng,
    const unsigned long long* __restrict__ umax, Rl* __restrict__ U0, Rl* __restrict__ U1, Rl* __restrict__ U2,
    Rl* __restrict__ shell, size_t shell_stride) {
    // tile t's particles: idx[off[t] .. off[t+1]) of the binning, their
    // in-tile offsets formed in the staging records' fixed point (the same
    // weights as the other tile deposits)
    using St = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    constexpr int E = TS + 1, E3 = E * E * E;
    constexpr int NT = kVec3Threads;
    __shared__ unsigned long long acc[3][E3];
    const WalkSeg sw = walk_segment(ng, TS);
    const double sc[3] = {vec3_scale(umax[0]), vec3_scale(umax[1]), vec3_scale(umax[2])};
    const double inv[3] = {1.0 / sc[0], 1.0 / sc[1], 1.0 / sc[2]};
    Rl* const out[3] = {U0, U1, U2};
    Rl* const sh[3] = {shell, shell + shell_stride, shell + 2 * shell_stride};
    for (int i = threadIdx.x; i < 3 * E3; i += NT) (&acc[0][0])[i] = 0ULL;
    __syncthreads();
    const int ox = sw.tx * TS, oy = sw.ty * TS;
    for (int k = 0; k < sw.seg; ++k) {
        const int tz = sw.sg * sw.seg + k, oz = tz * TS;
        const uint32_t t = ((uint32_t)sw.tx * sw.nt + sw.ty) * sw.nt + tz;
        const uint32_t beg = off[t], end = off[t + 1];
        constexpr int B = PM_VEC3_BATCH;  // particles in flight per thread
        for (uint32_t q0 = beg + threadIdx.x; q0 < end; q0 += NT * B) {
            uint32_t p[B];
            St r[B];
            double u[B][3];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const uint32_t q = q0 + NT * j;
                p[j] = q < end ? idx[q] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (p[j] == 0xFFFFFFFFu) continue;
                r[j] = tile_record_of<St>(px[p[j]] - ox, py[p[j]] - oy, pz[p[j]] - oz);
                u[j][0] = ux[p[j]] * sc[0];
                u[j][1] = uy[p[j]] * sc[1];
                u[j][2] = uz[p[j]] * sc[2];
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (p[j] == 0xFFFFFFFFu) continue;
                const TileOffs o = tile_offs(r[j]);
                const double fx = o.fx, fy = o.fy, fz = o.fz;
                const double w[8] = {(1 - fx) * (1 - fy) * (1 - fz), (1 - fx) * (1 - fy) * fz,
                                     (1 - fx) * fy * (1 - fz),       (1 - fx) * fy * fz,
                                     fx * (1 - fy) * (1 - fz),       fx * (1 - fy) * fz,
                                     fx * fy * (1 - fz),             fx * fy * fz};
                const int n0 = (o.a * E + o.b) * E + o.c;
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int n = n0 + ((m >> 2) & 1) * E * E + ((m >> 1) & 1) * E + (m & 1);
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        atomicAdd(&acc[a][n], (unsigned long long)fix_rne(w[m] * u[j][a]));
                }
            }
        }
        __syncthreads();
        walk_tile_out<Rl, TS, 3, NT>(&acc[0][0], inv, t, ox, oy, oz, ng, out, sh);
    }
    walk_seam_out<Rl, TS, 3, NT>(&acc[0][0], inv, sw, sh);
}

// ------------------------------------ divergence of the transpose, direct
// mu = sum_a D_a^T U_a with U_a = interp^T(u_a) (the mesh that k_div3 forms
// from the three k_deposit_vec3_tiles meshes), deposited in one pass: with
// (D_a f)[i] = (f[i - e_a] - f[i + e_a]) / 2, a particle's CIC weight w_c
// at corner c adds +u_a w_c / 2 to node c - e_a and -u_a w_c / 2 to node
// c + e_a.  That is 32 nodes, cell offsets -1..2 on one axis and 0..1 on
// the others; the 8 corners take one term per axis.  One signed 64-bit
// fixed-point accumulator per node, with the scale of the largest |u_a|;
// the column walk keeps kDivHalo = 3 halo planes above a tile, node box
// shifted one node down (walk_halo_tile_out), in 19^3 x 8 B = 55 KB of LDS,
// so two workgroups share a CU.  Against the vec3 deposit, divergence and
// three folds: one mesh written instead of three, one fold, no k_div3.
// Workgroup size and particles in flight per thread.  Reverse sweep of 20
// steps at 512^3 in an interleaved A/B: 237.2 ms (512 threads, 2 in
// flight), 237.6 (512, 1), 239.3 (512, 4), 244.6 (256, 2).
#ifndef PM_DIV_THREADS
#define PM_DIV_THREADS 512
#endif
#ifndef PM_DIV_BATCH
#define PM_DIV_BATCH 2
#endif
constexpr int kDivThreads = PM_DIV_THREADS;

template <typename Rl, int TS>
__global__ __launch_bounds__(kDivThreads) void k_deposit_div_tiles(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const void* __restrict__ stage, const void* __restrict__ ust, const double* __restrict__ ux,
    const double* __restrict__ uy, const double* __restrict__ uz, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ cnt, int ng, const unsigned long long* __restrict__ umax, Rl* __restrict__ mu,
    Rl* __restrict__ shell, const uint32_t* __restrict__ gate, int run_if) {
    // particles of tile t: buckets with records and staged cotangents
    // (stage, ust), or a binning (idx) with the records formed from the
    // positions and u read by index, rounded to the mesh precision as the
    // staging rounds it
    if (gate && ((*gate != 0u) != (run_if != 0))) return;
    using St = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    const St* __restrict__ sp = static_cast<const St*>(stage);
    const Real3<Rl>* __restrict__ su = static_cast<const Real3<Rl>*>(ust);
    constexpr int H = kDivHalo, E = TS + H, E3 = E * E * E, NT = kDivThreads;
    __shared__ unsigned long long acc[E3];
    const WalkSeg sw = walk_segment(ng, TS);
    const double sc = vec3_scale(max(umax[0], max(umax[1], umax[2])));
    const double hs = 0.5 * sc, inv = 1.0 / sc;
    for (int i = threadIdx.x; i < E3; i += NT) acc[i] = 0ULL;
    __syncthreads();
    const int ox = sw.tx * TS, oy = sw.ty * TS;
    for (int k = 0; k < sw.seg; ++k) {
        const int tz = sw.sg * sw.seg + k, oz = tz * TS;
        const uint32_t t = ((uint32_t)sw.tx * sw.nt + sw.ty) * sw.nt + tz;
        const uint32_t beg = off[t], end = cnt ? beg + cnt[t] : off[t + 1];
        constexpr int B = PM_DIV_BATCH;  // particles in flight per thread
        for (uint32_t q0 = beg + threadIdx.x; q0 < end; q0 += NT * B) {
            bool in[B];
            St r[B];
            double u[B][3];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const uint32_t q = q0 + NT * j;
                in[j] = q < end;
                if (!in[j]) continue;
                if (sp) {
                    r[j] = sp[q];
                    const Real3<Rl> w = su[q];
                    u[j][0] = (double)w.x;
                    u[j][1] = (double)w.y;
                    u[j][2] = (double)w.z;
                } else {
                    const uint32_t p = idx[q];
                    r[j] = tile_record_of<St>(px[p] - ox, py[p] - oy, pz[p] - oz);
                    u[j][0] = (double)(Rl)ux[p];
                    u[j][1] = (double)(Rl)uy[p];
                    u[j][2] = (double)(Rl)uz[p];
                }
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (!in[j]) continue;
                u[j][0] *= hs;
                u[j][1] *= hs;
                u[j][2] *= hs;
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (!in[j]) continue;
                const TileOffs o = tile_offs(r[j]);
                const double wx[2] = {1.0 - o.fx, o.fx}, wy[2] = {1.0 - o.fy, o.fy}, wz[2] = {1.0 - o.fz, o.fz};
                // node (dx, dy, dz) relative to the cell's lower corner
                unsigned long long* const n0 = acc + ((o.a + 1) * E + (o.b + 1)) * E + (o.c + 1);
                auto add = [&](int dx, int dy, int dz, double v) {
                    atomicAdd(n0 + (dx * E + dy) * E + dz, (unsigned long long)fix_rne(v));
                };
#pragma unroll
                for (int e = 0; e < 2; ++e)
#pragma unroll
                    for (int f = 0; f < 2; ++f) {
                        add(-1, e, f, u[j][0] * wx[0] * wy[e] * wz[f]);
                        add(2, e, f, -u[j][0] * wx[1] * wy[e] * wz[f]);
                        add(e, -1, f, u[j][1] * wx[e] * wy[0] * wz[f]);
                        add(e, 2, f, -u[j][1] * wx[e] * wy[1] * wz[f]);
                        add(e, f, -1, u[j][2] * wx[e] * wy[f] * wz[0]);
                        add(e, f, 2, -u[j][2] * wx[e] * wy[f] * wz[1]);
                    }
#pragma unroll
                for (int cx = 0; cx < 2; ++cx)
#pragma unroll
                    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                        for (int cz = 0; cz < 2; ++cz) {
                            const double vx = cx == 0 ? u[j][0] * wx[1] * wy[cy] * wz[cz]
                                                      : -u[j][0] * wx[0] * wy[cy] * wz[cz];
                            const double vy = cy == 0 ? u[j][1] * wx[cx] * wy[1] * wz[cz]
                                                      : -u[j][1] * wx[cx] * wy[0] * wz[cz];
                            const double vz = cz == 0 ? u[j][2] * wx[cx] * wy[cy] * wz[1]
                                                      : -u[j][2] * wx[cx] * wy[cy] * wz[0];
                            add(cx, cy, cz, vx + vy + vz);
                        }
            }
        }
        __syncthreads();
        walk_halo_tile_out<Rl, TS, H, NT>(acc, inv, t, ox - 1, oy - 1, oz - 1, ng, mu, shell);
    }
    walk_halo_seam_out<Rl, TS, H, NT>(acc, inv, sw, shell);
}

// Max of two magnitudes (sign bit clear) by their bit patterns, which order
// like the values for finite ones and put +inf above them and NaN above +inf:
// unlike fmax, a NaN is kept, so a non-finite cotangent reaches the max.
__device__ __forceinline__ double mag_max(double a, double b) {
    return __longlong_as_double(max(__double_as_longlong(a), __double_as_longlong(b)));
}
constexpr unsigned long long kInfBits = 0x7FF0000000000000ULL;

// Workgroup (NT threads) max of m[0..3) into one of kMaxSlots slots per
// component (umax[3 + d * kMaxSlots + slot]): bit patterns of non-negative
// doubles order like the values, so an integer atomic max works, and the
// slots spread the atomics; k_max_slots reduces them into umax[0..3).
template <int NT>
__device__ __forceinline__ void wg_max_to_slots(double (&m)[3], unsigned long long* __restrict__ umax) {
    __shared__ double wm[3][NT / 64];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int o = 32; o > 0; o >>= 1) m[d] = mag_max(m[d], __shfl_xor(m[d], o));
        if (__lane_id() == 0) wm[d][threadIdx.x >> 6] = m[d];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = wm[threadIdx.x][0];
        for (int w = 1; w < NT / 64; ++w) v = mag_max(v, wm[threadIdx.x][w]);
        atomicMax(&umax[3 + threadIdx.x * kMaxSlots + (blockIdx.x & (kMaxSlots - 1))],
                  (unsigned long long)__double_as_longlong(v));
    }
}

// umax[d] = max over the slots of component d; the slots are cleared.  A
// non-finite max (an inf or NaN cotangent, which no fixed-point scale can
// hold) is counted in the range-error counter.
__global__ __launch_bounds__(kMaxSlots) void k_max_slots(unsigned long long* __restrict__ umax,
                                                         unsigned long long* __restrict__ err) {
    __shared__ unsigned long long r[3][kMaxSlots];
    for (int d = 0; d < 3; ++d) {
        r[d][threadIdx.x] = umax[3 + d * kMaxSlots + threadIdx.x];
        umax[3 + d * kMaxSlots + threadIdx.x] = 0ULL;
    }
    __syncthreads();
    for (int h = kMaxSlots / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int d = 0; d < 3; ++d) r[d][threadIdx.x] = max(r[d][threadIdx.x], r[d][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        umax[threadIdx.x] = r[threadIdx.x][0];
        if (r[threadIdx.x][0] >= kInfBits) atomicAdd(err, 1ULL);
    }
}

// max |u_a| over all particles into the slots
__global__ __launch_bounds__(256) void k_absmax3(const double* __restrict__ ux, const double* __restrict__ uy,
                                                 const double* __restrict__ uz, int64_t np,
                                                 unsigned long long* __restrict__ umax) {
    double m[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < np; i += (int64_t)gridDim.x * 256) {
        m[0] = mag_max(m[0], fabs(ux[i]));
        m[1] = mag_max(m[1], fabs(uy[i]));
        m[2] = mag_max(m[2], fabs(uz[i]));
    }
    wg_max_to_slots<256>(m, umax);
}

// ------------------------------------------------- mu = sum_a D_a^T U_a
template <typename Rl>
__global__ __launch_bounds__(256) void k_div3(const Rl* __restrict__ U0, const Rl* __restrict__ U1,
                                              const Rl* __restrict__ U2, int N, Rl* __restrict__ mu) {
    const size_t total = (size_t)N * N * N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % N);
        const int j = (int)((e / N) % N);
        const int i = (int)(e / ((size_t)N * N));
        const size_t sx = (size_t)N * N, sy = (size_t)N;
        const size_t ip = i + 1 == N ? e - (size_t)i * sx : e + sx;
        const size_t im = i == 0 ? e + (size_t)(N - 1) * sx : e - sx;
        const size_t jp = j + 1 == N ? e - (size_t)j * sy : e + sy;
        const size_t jm = j == 0 ? e + (size_t)(N - 1) * sy : e - sy;
        const size_t kp = k + 1 == N ? e - (size_t)k : e + 1;
        const size_t km = k == 0 ? e + (size_t)(N - 1) : e - 1;
        mu[e] = ((U0[ip] - U0[im]) + (U1[jp] - U1[jm]) + (U2[kp] - U2[km])) * Rl(0.5);
    }
}

// -------------------------------------------------- reverse-step gather
// MODE 1: xbar += vjp_scale * (term A + term B) (pure VJP).
// MODE 0: also v -= kick * F(x) (undoes the step's kick).
// MODE 2: as MODE 0, then the undrift that opens the next reverse step,
//         x -= next_drift * v and u = vbar + next_drift * xbar (in place on
//         u), with max |u_a| into umax for that step's divergence deposit.
// Workgroup size of the adjoint gather.  Full reverse sweep of 20 steps at
// 512^3 (FP32 mesh) in an interleaved A/B, 128 / 256 / 512 / 1024 threads:
// 265.1 / 258.6 / 261.7 / 278.7 ms.
#ifndef PM_ADJ_THREADS32
#define PM_ADJ_THREADS32 256
#endif
#ifndef PM_ADJ_THREADS64
#define PM_ADJ_THREADS64 256
#endif
constexpr int kAdjThreads32 = PM_ADJ_THREADS32, kAdjThreads64 = PM_ADJ_THREADS64;
template <typename Rl, int MODE, bool STAGE, int NT>
__global__ __launch_bounds__(NT) void k_adj_gather(double* __restrict__ px, double* __restrict__ py,
                                                    double* __restrict__ pz, double* __restrict__ vx,
                                                    double* __restrict__ vy, double* __restrict__ vz,
                                                    double* __restrict__ ux, double* __restrict__ uy,
                                                    double* __restrict__ uz, double* __restrict__ bx,
                                                    double* __restrict__ by, double* __restrict__ bz, int64_t np,
                                                    int N, const Rl* __restrict__ phi, const Rl* __restrict__ lam,
                                                    double mass, double kick, double vjp_scale, double next_drift,
                                                    unsigned long long* __restrict__ umax,
                                                    unsigned long long* __restrict__ err, AdjStage stg,
                                                    BandMap bm) {
    static_assert(!STAGE || MODE == 2, "only the undrifting gather stages");
    using Rec = typename std::conditional<sizeof(Rl) == 4, StageQ, StageD>::type;
    constexpr uint32_t kNo = 0xFFFFFFFFu;
    constexpr int HS = STAGE ? 128 : 1;
    __shared__ uint32_t hkey[HS], hcnt[HS], hbase[HS];
    uint32_t kk = kNo;  // STAGE: tile of the undrifted position
    Rec rec{};
    double nu_st[3] = {0.0, 0.0, 0.0};  // STAGE: its cotangent u
    const int64_t i = (int64_t)band_remap(blockIdx.x, gridDim.x, bm) * NT + threadIdx.x;
    const double box = (double)N;
    // three load phases, as in k_interp: the particle's streams, then the
    // gathers around its cell (cell 0 for a lane that stores nothing), then
    // the arithmetic and the stores
    const bool live = i < np;
    const int64_t j = live ? i : 0;
    const double x = stream_load(px + j), y = stream_load(py + j), z = stream_load(pz + j);
    const double u[3] = {stream_load(ux + j), stream_load(uy + j), stream_load(uz + j)};
    const double b0[3] = {stream_load(bx + j), stream_load(by + j), stream_load(bz + j)};
    double v0[3] = {0.0, 0.0, 0.0};
    if constexpr (MODE != 1) {
        v0[0] = stream_load(vx + j);
        v0[1] = stream_load(vy + j);
        v0[2] = stream_load(vz + j);
    }
    const bool ok = live & in_box3(x, y, z, box);
    if (live && !ok) atomicAdd(err, 1ULL);
    const CellIdx<Rl> c = cell_of<Rl>(ok ? x : 0.0, ok ? y : 0.0, ok ? z : 0.0, N);
    Rl cl[2][2][4], xr[2][2][2], yr[2][2][2], lf[2][2][2];
    gather_runs_main(phi, N, c, cl, xr, yr);
    gather_corners_main(lam, N, c, lf);
    gather_runs_seam(phi, N, c, cl, xr, yr);
    gather_corners_seam(lam, N, c, lf);
    double m[3] = {0.0, 0.0, 0.0};  // MODE 2: |u| of the next step
    Rl gx[2][2][2], gy[2][2][2], gz[2][2][2];
    corner_differences(cl, xr, yr, gx, gy, gz);
    Rl F[3], dG[3][3], dL[3], lv;
    tri_val_grad(gx, c.tx, c.ty, c.tz, F[0], dG[0]);
    tri_val_grad(gy, c.tx, c.ty, c.tz, F[1], dG[1]);
    tri_val_grad(gz, c.tx, c.ty, c.tz, F[2], dG[2]);
    tri_val_grad(lf, c.tx, c.ty, c.tz, lv, dL);
    if (ok) {
        double nb[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            nb[d] = b0[d] + vjp_scale * (u[0] * (double)dG[0][d] + u[1] * (double)dG[1][d] +
                                         u[2] * (double)dG[2][d] + mass * (double)dL[d]);
        stream_store(bx + i, nb[0]);
        stream_store(by + i, nb[1]);
        stream_store(bz + i, nb[2]);
        if constexpr (MODE != 1) {
            const double nv[3] = {v0[0] - kick * (double)F[0], v0[1] - kick * (double)F[1],
                                  v0[2] - kick * (double)F[2]};
            stream_store(vx + i, nv[0]);
            stream_store(vy + i, nv[1]);
            stream_store(vz + i, nv[2]);
            if constexpr (MODE == 2) {
                auto wrap = [&](double v) {
                    if (v < 0.0) v += box;
                    if (v >= box) v -= box;
                    return v;
                };
                const double xn = wrap(x - next_drift * nv[0]), yn = wrap(y - next_drift * nv[1]),
                             zn = wrap(z - next_drift * nv[2]);
                stream_store(px + i, xn);
                stream_store(py + i, yn);
                stream_store(pz + i, zn);
                if constexpr (STAGE) {
                    if (in_box3(xn, yn, zn, box)) {
                        kk = tile_key(xn, yn, zn, N, stg.ts);
                        rec = tile_record<Rec>(xn, yn, zn, stg.ts);
                    } else {
                        atomicAdd(err, 1ULL);  // as the next deposit's range check would
                    }
                }
                const double nu[3] = {u[0] + next_drift * nb[0], u[1] + next_drift * nb[1],
                                      u[2] + next_drift * nb[2]};
                stream_store(ux + i, nu[0]);
                stream_store(uy + i, nu[1]);
                stream_store(uz + i, nu[2]);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    m[d] = fabs(nu[d]);
                    nu_st[d] = nu[d];
                }
            }
        }
    }
    if constexpr (STAGE) {
        const int64_t pos = stage_pass<HS, NT>(kk, hkey, hcnt, hbase, stg.fill, stg.coff, stg.spill);
        if (pos >= 0) {
            static_cast<Rec*>(stg.rec)[pos] = rec;
            static_cast<Real3<Rl>*>(stg.u)[pos] = Real3<Rl>{(Rl)nu_st[0], (Rl)nu_st[1], (Rl)nu_st[2]};
        }
    }
    if constexpr (MODE == 2) wg_max_to_slots<NT>(m, umax);
}

// x -= drift * v (wrapped); u = vbar + drift * xbar (in place on vbar)
__global__ __launch_bounds__(256) void k_undrift(double* __restrict__ px, double* __restrict__ py,
                                                 double* __restrict__ pz, const double* __restrict__ vx,
                                                 const double* __restrict__ vy, const double* __restrict__ vz,
                                                 const double* __restrict__ bx, const double* __restrict__ by,
                                                 const double* __restrict__ bz, double* __restrict__ ubx,
                                                 double* __restrict__ uby, double* __restrict__ ubz, int64_t np,
                                                 double box, double drift) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    auto wrap = [&](double v) {
        if (v < 0.0) v += box;
        if (v >= box) v -= box;
        return v;
    };
    px[i] = wrap(px[i] - drift * vx[i]);
    py[i] = wrap(py[i] - drift * vy[i]);
    pz[i] = wrap(pz[i] - drift * vz[i]);
    ubx[i] += drift * bx[i];
    uby[i] += drift * by[i];
    ubz[i] += drift * bz[i];
}

// ================================================================ launchers
int launch_adj_deposit(int prec, const double* px, const double* py, const double* pz, int64_t np, int ng,
                       const void* mbar, double scale, double* bx, double* by, double* bz, unsigned long long* err,
                       hipStream_t st) {
    if (np <= 0) return PM_OK;
    const unsigned nb = (unsigned)((np + 255) / 256);
    if (prec == PM_FP32)
        hipLaunchKernelGGL(k_adj_deposit<float>, dim3(nb), dim3(256), 0, st, px, py, pz, np, ng, (const float*)mbar,
                           scale, bx, by, bz, err);
    else
        hipLaunchKernelGGL(k_adj_deposit<double>, dim3(nb), dim3(256), 0, st, px, py, pz, np, ng,
                           (const double*)mbar, scale, bx, by, bz, err);
    return check_launch("adjoint_deposit");
}

int launch_deposit_vec3(int prec, const double* px, const double* py, const double* pz, const double* ux,
                        const double* uy, const double* uz, int64_t np, int ng, const SortedScratch& s,
                        void* const U[3], void* shell3, unsigned long long* umax, unsigned long long* err,
                        hipStream_t st) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts;
    const uint32_t ntiles = (uint32_t)nt * nt * nt;
    const size_t sh = sorted_shell_elems(ng);
    const size_t rs = prec == PM_FP32 ? sizeof(float) : sizeof(double);
    if (np > 0) {
        const unsigned nb = (unsigned)std::min<int64_t>((np + 255) / 256, 4096);
        hipLaunchKernelGGL(k_absmax3, dim3(nb), dim3(256), 0, st, ux, uy, uz, np, umax);
    }
    hipLaunchKernelGGL(k_max_slots, dim3(1), dim3(kMaxSlots), 0, st, umax, err);
#define PM_VEC3(RL, TSV)                                                                                          \
    hipLaunchKernelGGL((k_deposit_vec3_tiles<RL, TSV>), dim3(ntiles / walk_seg(nt)), dim3(kVec3Threads), 0, st, px, \
                       py, pz, ux, uy, uz, s.idx, s.off, ng, umax, (RL*)U[0], (RL*)U[1], (RL*)U[2], (RL*)shell3, sh)
    if (prec == PM_FP32) {
        if (ts == 16)
            PM_VEC3(float, 16);
        else
            PM_VEC3(float, 8);
    } else {
        if (ts == 16)
            PM_VEC3(double, 16);
        else
            PM_VEC3(double, 8);
    }
#undef PM_VEC3
    for (int d = 0; d < 3; ++d) launch_fold_shells(prec, ts, ng, (const char*)shell3 + d * sh * rs, U[d], st);
    return check_launch("deposit_vec3");
}

int launch_deposit_div(int prec, const double* px, const double* py, const double* pz, const double* ux,
                       const double* uy, const double* uz, int64_t np, int ng, const SortedScratch& s, void* mu,
                       void* shell, unsigned long long* umax, unsigned long long* err, hipStream_t st,
                       bool umax_ready, const void* bucket_u) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts;
    const uint32_t ntiles = (uint32_t)nt * nt * nt;
    if (!umax_ready) {  // else the previous reverse-step gather left max |u_a|
        if (np > 0) {
            const unsigned nb = (unsigned)std::min<int64_t>((np + 255) / 256, 4096);
            hipLaunchKernelGGL(k_absmax3, dim3(nb), dim3(256), 0, st, ux, uy, uz, np, umax);
        }
        hipLaunchKernelGGL(k_max_slots, dim3(1), dim3(kMaxSlots), 0, st, umax, err);
    }
    // with buckets: the bucket pass, or (overflow) the binning pass
#define PM_DIV_PASS(RL, TSV, REC, UST, IDX, OFF, CNT, GATE, RUN)                                                 \
    hipLaunchKernelGGL((k_deposit_div_tiles<RL, TSV>), dim3(ntiles / walk_seg(nt)), dim3(kDivThreads), 0, st, px, \
                       py, pz, REC, UST, ux, uy, uz, IDX, OFF, CNT, ng, umax, (RL*)mu, (RL*)shell, GATE, RUN)
#define PM_DIV(RL, TSV)                                                                                     \
    if (bucket_u) {                                                                                         \
        PM_DIV_PASS(RL, TSV, s.stage, bucket_u, nullptr, s.coff, s.fill, s.spill, 0);                       \
        PM_DIV_PASS(RL, TSV, nullptr, nullptr, s.idx, s.off, nullptr, s.spill, 1);                          \
    } else {                                                                                                \
        PM_DIV_PASS(RL, TSV, nullptr, nullptr, s.idx, s.off, nullptr, nullptr, 0);                          \
    }
    if (prec == PM_FP32) {
        if (ts == 16) {
            PM_DIV(float, 16)
        } else {
            PM_DIV(float, 8)
        }
    } else {
        if (ts == 16) {
            PM_DIV(double, 16)
        } else {
            PM_DIV(double, 8)
        }
    }
#undef PM_DIV
#undef PM_DIV_PASS
    const int rc = check_launch("deposit_div");
    if (rc) return rc;
    return launch_fold_halo(prec, ts, ng, shell, mu, st);
}

int launch_div3(int prec, void* const U[3], int ng, void* mu, hipStream_t st) {
    const size_t n = (size_t)ng * ng * ng;
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    if (prec == PM_FP32)
        hipLaunchKernelGGL(k_div3<float>, dim3(nb), dim3(256), 0, st, (const float*)U[0], (const float*)U[1],
                           (const float*)U[2], ng, (float*)mu);
    else
        hipLaunchKernelGGL(k_div3<double>, dim3(nb), dim3(256), 0, st, (const double*)U[0], (const double*)U[1],
                           (const double*)U[2], ng, (double*)mu);
    return check_launch("div3");
}

int launch_adj_gather(int prec, int mode, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                      double* ux, double* uy, double* uz, double* bx, double* by, double* bz, int64_t np, int ng,
                      const void* phi, const void* lam, double mass, double kick, double vjp_scale,
                      unsigned long long* err, hipStream_t st, double next_drift, unsigned long long* umax,
                      const AdjStage* stage) {
    if (np <= 0) return PM_OK;
    if (mode == 2 && !umax) return fail(PM_ERR_ARG, "adjoint gather: no max buffer");
    if (stage && (mode != 2 || !stage->rec || !stage->u)) return fail(PM_ERR_ARG, "adjoint gather: bad staging");
    const int nt = prec == PM_FP32 ? kAdjThreads32 : kAdjThreads64;
    const unsigned nb = (unsigned)((np + nt - 1) / nt);
    const AdjStage stg = stage ? *stage : AdjStage();
    const BandMap bm = band_map(np, nb, nt);
#define PM_GATHER(RL, M, S, NTV)                                                                                         \
    hipLaunchKernelGGL((k_adj_gather<RL, M, S, NTV>), dim3(nb), dim3(NTV), 0, st, px, py, pz, vx, vy, vz, ux, uy, uz, bx, \
                       by, bz, np, ng, (const RL*)phi, (const RL*)lam, mass, kick, vjp_scale, next_drift, umax, err,  \
                       stg, bm)
    if (prec == PM_FP32) {
        if (mode == 0)
            PM_GATHER(float, 0, false, kAdjThreads32);
        else if (mode == 1)
            PM_GATHER(float, 1, false, kAdjThreads32);
        else if (!stage)
            PM_GATHER(float, 2, false, kAdjThreads32);
        else
            PM_GATHER(float, 2, true, kAdjThreads32);
    } else {
        if (mode == 0)
            PM_GATHER(double, 0, false, kAdjThreads64);
        else if (mode == 1)
            PM_GATHER(double, 1, false, kAdjThreads64);
        else if (!stage)
            PM_GATHER(double, 2, false, kAdjThreads64);
        else
            PM_GATHER(double, 2, true, kAdjThreads64);
    }
#undef PM_GATHER
    if (mode == 2) hipLaunchKernelGGL(k_max_slots, dim3(1), dim3(kMaxSlots), 0, st, umax, err);
    return check_launch("adjoint_gather");
}

int launch_undrift(double* px, double* py, double* pz, const double* vx, const double* vy, const double* vz,
                   const double* bx, const double* by, const double* bz, double* ubx, double* uby, double* ubz,
                   int64_t np, int ng, double drift, hipStream_t st) {
    if (np <= 0) return PM_OK;
    const unsigned nb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_undrift, dim3(nb), dim3(256), 0, st, px, py, pz, vx, vy, vz, bx, by, bz, ubx, uby, ubz, np,
                       (double)ng, drift);
    return check_launch("undrift");
}

}  // namespace pm
