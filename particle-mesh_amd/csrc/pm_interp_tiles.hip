// Interpolation + kick + drift over the tile binning of the sorted deposit.
//
// update_particles (utils.cpp:141-170) with the centred difference of
// centered_derivatives (utils.cpp:130-132) taken on the fly, as in k_interp,
// but one workgroup per 16^3-cell tile: the (16+3)^3 potential nodes its
// particles can touch (corners 0..16 and the +-1 difference stencil) are
// staged once in LDS, so every potential value is fetched from HBM once per
// tile instead of once per particle stencil.  Particles are visited through
// the binning's index list (runs of consecutive indices for spatially
// coherent particle orders).
//
// With KEYS, the kernel also bins the drifted positions for the next step:
// each particle's new tile key goes to key[p] and the tile counts are
// aggregated in an LDS hash (one global atomic per distinct tile per
// workgroup; a particle whose key finds the hash full adds itself directly).
#include "pm_internal.h"

namespace pm {

constexpr int kItHash = 128;  // distinct next-step tiles tracked in LDS per workgroup
constexpr uint32_t kNoKey = 0xFFFFFFFFu;

template <typename Rl, int TS, bool KEYS>
__global__ __launch_bounds__(256) void k_interp_tiles(double* __restrict__ px, double* __restrict__ py,
                                                      double* __restrict__ pz, double* __restrict__ vx,
                                                      double* __restrict__ vy, double* __restrict__ vz,
                                                      const uint32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ off, int ng,
                                                      const Rl* __restrict__ phi, double kick, double drift,
                                                      uint32_t* __restrict__ cnt, uint32_t* __restrict__ key,
                                                      unsigned long long* __restrict__ err) {
    constexpr int H = TS + 3, H3 = H * H * H;
    __shared__ Rl f[H3];
    __shared__ uint32_t hkey[KEYS ? kItHash : 1];
    __shared__ uint32_t hcnt[KEYS ? kItHash : 1];
    const int N = ng;
    const int nt = ng / TS;
    const uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = (int)(t / (uint32_t)(nt * nt)), ty = (int)((t / (uint32_t)nt) % (uint32_t)nt),
              tz = (int)(t % (uint32_t)nt);
    const int ox = tx * TS, oy = ty * TS, oz = tz * TS;
    // stage nodes [o-1, o+TS+1] of every axis (periodic)
    for (int i = threadIdx.x; i < H3; i += 256) {
        const int a = i / (H * H), b = (i / H) % H, c = i % H;
        int gx = ox - 1 + a, gy = oy - 1 + b, gz = oz - 1 + c;
        gx += gx < 0 ? N : (gx >= N ? -N : 0);
        gy += gy < 0 ? N : (gy >= N ? -N : 0);
        gz += gz < 0 ? N : (gz >= N ? -N : 0);
        f[i] = phi[((size_t)gx * N + gy) * N + gz];
    }
    if constexpr (KEYS) {
        for (int i = threadIdx.x; i < kItHash; i += 256) {
            hkey[i] = kNoKey;
            hcnt[i] = 0u;
        }
    }
    __syncthreads();
    const double box = (double)ng;
    const uint32_t beg = off[t], end = off[t + 1];
    constexpr int B = 2;  // particles in flight per thread
    for (uint32_t q0 = beg + threadIdx.x; q0 < end; q0 += 256 * B) {
        uint32_t p[B];
        double x[B], y[B], z[B], u[B], v[B], w[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const uint32_t q = q0 + 256 * j;
            p[j] = q < end ? idx[q] : kNoKey;
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (p[j] == kNoKey) continue;
            x[j] = px[p[j]];
            y[j] = py[p[j]];
            z[j] = pz[p[j]];
            u[j] = vx[p[j]];
            v[j] = vy[p[j]];
            w[j] = vz[p[j]];
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (p[j] == kNoKey) continue;
            const int ix = (int)x[j], iy = (int)y[j], iz = (int)z[j];
            const Rl fx = (Rl)(x[j] - ix), fy = (Rl)(y[j] - iy), fz = (Rl)(z[j] - iz);
            const Rl mx = Rl(1) - fx, my = Rl(1) - fy, mz = Rl(1) - fz;
            // LDS index of the particle's cell corner (offset +1 for the halo)
            const int c0 = ((ix - ox + 1) * H + (iy - oy + 1)) * H + (iz - oz + 1);
            auto P = [&](int da, int db, int dc) -> Rl { return f[c0 + (da * H + db) * H + dc]; };
            Rl gx[2][2][2], gy[2][2][2], gz[2][2][2];
            const Rl hf = Rl(0.5);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        gx[a][b][c] = (P(a - 1, b, c) - P(a + 1, b, c)) * hf;
                        gy[a][b][c] = (P(a, b - 1, c) - P(a, b + 1, c)) * hf;
                        gz[a][b][c] = (P(a, b, c - 1) - P(a, b, c + 1)) * hf;
                    }
            auto tri = [&](const Rl (&g)[2][2][2]) -> Rl {  // trilerp1 order
                const Rl f00 = g[0][0][0] * mx + g[1][0][0] * fx;
                const Rl f10 = g[0][1][0] * mx + g[1][1][0] * fx;
                const Rl fz0 = f00 * my + f10 * fy;
                const Rl g00 = g[0][0][1] * mx + g[1][0][1] * fx;
                const Rl g10 = g[0][1][1] * mx + g[1][1][1] * fx;
                const Rl fz1 = g00 * my + g10 * fy;
                return fz0 * mz + fz1 * fz;
            };
            const double nu = u[j] + (double)tri(gx) * kick;
            const double nv = v[j] + (double)tri(gy) * kick;
            const double nw = w[j] + (double)tri(gz) * kick;
            vx[p[j]] = nu;
            vy[p[j]] = nv;
            vz[p[j]] = nw;
            double nx = x[j], ny = y[j], nz = z[j];
            if (drift != 0.0) {
                nx += nu * drift;
                ny += nv * drift;
                nz += nw * drift;
                nx += nx < 0.0 ? box : 0.0;
                nx -= nx >= box ? box : 0.0;
                ny += ny < 0.0 ? box : 0.0;
                ny -= ny >= box ? box : 0.0;
                nz += nz < 0.0 ? box : 0.0;
                nz -= nz >= box ? box : 0.0;
                px[p[j]] = nx;
                py[p[j]] = ny;
                pz[p[j]] = nz;
            }
            if constexpr (KEYS) {
                uint32_t k = kNoKey;
                if (nx >= 0.0 && nx < box && ny >= 0.0 && ny < box && nz >= 0.0 && nz < box)
                    k = (uint32_t)((((int)nx / TS) * nt + (int)ny / TS) * nt + (int)nz / TS);
                else
                    atomicAdd(err, 1ULL);
                key[p[j]] = k;
                if (k != kNoKey) {
                    uint32_t s = (k * 2654435761u) >> 25;  // 7 bits
                    bool done = false;
                    for (int probe = 0; probe < kItHash; ++probe) {
                        const uint32_t old = atomicCAS(&hkey[s], kNoKey, k);
                        if (old == kNoKey || old == k) {
                            atomicAdd(&hcnt[s], 1u);
                            done = true;
                            break;
                        }
                        s = (s + 1) & (kItHash - 1);
                    }
                    if (!done) atomicAdd(&cnt[k], 1u);
                }
            }
        }
    }
    if constexpr (KEYS) {
        __syncthreads();
        for (int s = threadIdx.x; s < kItHash; s += 256)
            if (hkey[s] != kNoKey) atomicAdd(&cnt[hkey[s]], hcnt[s]);
    }
}

int launch_interp_tiles(int prec, double* px, double* py, double* pz, double* vx, double* vy, double* vz, int ng,
                        const void* phi, double kick, double drift, SortedScratch& s, int keys,
                        unsigned long long* err, hipStream_t st) {
    const int ts = sorted_tile(ng);
    const int nt = ng / ts;
    const uint32_t ntiles = (uint32_t)nt * nt * nt;
#define PM_IT(RL, TSV, K)                                                                                          \
    hipLaunchKernelGGL((k_interp_tiles<RL, TSV, K>), dim3(ntiles), dim3(256), 0, st, px, py, pz, vx, vy, vz, s.idx, \
                       s.off, ng, (const RL*)phi, kick, drift, s.cnt, s.key, err)
    if (prec == PM_FP32) {
        if (ts == 16) {
            if (keys)
                PM_IT(float, 16, true);
            else
                PM_IT(float, 16, false);
        } else {
            if (keys)
                PM_IT(float, 8, true);
            else
                PM_IT(float, 8, false);
        }
    } else {
        if (ts == 16) {
            if (keys)
                PM_IT(double, 16, true);
            else
                PM_IT(double, 16, false);
        } else {
            if (keys)
                PM_IT(double, 8, true);
            else
                PM_IT(double, 8, false);
        }
    }
#undef PM_IT
    return check_launch("interp_tiles");
}

}  // namespace pm
