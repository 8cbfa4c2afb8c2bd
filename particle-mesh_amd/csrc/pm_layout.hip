// Reference-layout operators (FP64): the drop-in kernels behind the
// reference's src/utils and src/decomp/slab routines for one box.
#include "pm_internal.h"

namespace pm {

// init_buf (utils.cpp:38-68) for one grid.
__global__ __launch_bounds__(256) void k_init_buf(double* __restrict__ buf, int nx, int ny, int nz) {
    const size_t total = (size_t)nx * ny * nz;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % nz);
        const int j = (int)((e / nz) % ny);
        const int i = (int)(e / ((size_t)ny * nz));
        buf[e] = (i < nx - 1 && j < ny - 1 && k < nz - 1) ? -1.0 : 0.0;
    }
}

// deposit_particles + cic_deposit (utils.cpp:70-102, utils_inline.h:25-49),
// with FP64 atomics in place of the per-thread grids and their reduction
// (utils.cpp:74-77, 104-113).
__global__ __launch_bounds__(256) void k_deposit_cube(const double* __restrict__ px, const double* __restrict__ py,
                                                      const double* __restrict__ pz, int64_t np,
                                                      double* __restrict__ buf, int nx, int ny, int nz, int x0,
                                                      int y0, int z0, double mass,
                                                      unsigned long long* __restrict__ err) {
    const int64_t p = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (p >= np) return;
    const double xp = px[p], yp = py[p], zp = pz[p];
    if (!(xp >= (double)x0 && xp < (double)(x0 + nx - 1) && yp >= (double)y0 && yp < (double)(y0 + ny - 1) &&
          zp >= (double)z0 && zp < (double)(z0 + nz - 1))) {
        atomicAdd(err, 1ULL);
        return;
    }
    const int ix = (int)xp, iy = (int)yp, iz = (int)zp;
    const size_t ys = (size_t)nz, xs = (size_t)ny * nz;
    const size_t base = ((size_t)(ix - x0) * ny + (iy - y0)) * nz + (iz - z0);
    const double tx = xp - (double)ix, ty = yp - (double)iy, tz = zp - (double)iz;
    const double sx0 = 1.0 - tx, sy0 = 1.0 - ty, sz0 = 1.0 - tz;
    const double sxy00 = sx0 * sy0, sxy10 = tx * sy0, sxy01 = sx0 * ty, sxy11 = tx * ty;
    const double m0 = mass * sz0, m1 = mass * tz;
    double* a = buf + base;
    unsafeAtomicAdd(a, sxy00 * m0);
    unsafeAtomicAdd(a + xs, sxy10 * m0);
    unsafeAtomicAdd(a + ys, sxy01 * m0);
    unsafeAtomicAdd(a + xs + ys, sxy11 * m0);
    unsafeAtomicAdd(a + 1, sxy00 * m1);
    unsafeAtomicAdd(a + xs + 1, sxy10 * m1);
    unsafeAtomicAdd(a + ys + 1, sxy01 * m1);
    unsafeAtomicAdd(a + xs + ys + 1, sxy11 * m1);
}

// TransposeSlab + ReorderSlab at one box: every slab node gathers the cube
// node and its + ghost images (reorder_slab.cpp:15-22 keys, :43-67 sums).
__global__ __launch_bounds__(256) void k_reorder_slab(const double* __restrict__ cube, double* __restrict__ slab,
                                                      int ng, int stride) {
    const size_t total = (size_t)ng * ng * ng;
    const size_t n = (size_t)ng + 1;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % ng);
        const int j = (int)((e / ng) % ng);
        const int i = (int)(e / ((size_t)ng * ng));
        double s = 0.0;
        for (int a = 0; a < (i == 0 ? 2 : 1); ++a)
            for (int b = 0; b < (j == 0 ? 2 : 1); ++b)
                for (int c = 0; c < (k == 0 ? 2 : 1); ++c) {
                    const size_t ci = (size_t)(a ? ng : i), cj = (size_t)(b ? ng : j), ck = (size_t)(c ? ng : k);
                    s += cube[(ci * n + cj) * n + ck];
                }
        slab[((size_t)i * ng + j) * stride + k] = s;
    }
}

// ReorderSlabBack at one box (reorder_slab_back.cpp:11-21): slab ->
// ng x (ng+3) x (ng+3) with the -1/+2 periodic halo on y and z.
__global__ __launch_bounds__(256) void k_reorder_slab_back(const double* __restrict__ slab,
                                                           double* __restrict__ out, int ng, int stride) {
    const int m = ng + 3;
    const size_t total = (size_t)ng * m * m;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % m);
        const int j = (int)((e / m) % m);
        const int i = (int)(e / ((size_t)m * m));
        const int sj = (j - 1 + ng) % ng, sk = (k - 1 + ng) % ng;
        out[e] = slab[((size_t)i * ng + sj) * stride + sk];
    }
}

// TransposeSlabBack at one box (transpose_slab_back.cpp:13-28): the x halo,
// plane -1 <- ng-1 and planes ng, ng+1 <- 0, 1, around the ng planes.
__global__ __launch_bounds__(256) void k_transpose_slab_back(const double* __restrict__ in,
                                                             double* __restrict__ halo, int ng) {
    const int m = ng + 3;
    const size_t plane = (size_t)m * m;
    const size_t total = plane * m;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / plane);
        const int si = (i - 1 + ng) % ng;
        halo[e] = in[(size_t)si * plane + e % plane];
    }
}

// centered_derivatives (utils.cpp:115-137).
__global__ __launch_bounds__(256) void k_centered(const double* __restrict__ in, double* __restrict__ dx,
                                                  double* __restrict__ dy, double* __restrict__ dz, int nx1,
                                                  int ny1, int nz1) {
    const int nx = nx1 - 2, ny = ny1 - 2, nz = nz1 - 2;
    const size_t total = (size_t)nx * ny * nz;
    const size_t yz1 = (size_t)ny1 * nz1;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int kk = (int)(e % nz);
        const int jj = (int)((e / nz) % ny);
        const int ii = (int)(e / ((size_t)ny * nz));
        const size_t c = (size_t)(ii + 1) * yz1 + (size_t)(jj + 1) * nz1 + (kk + 1);
        dx[e] = (in[c - yz1] - in[c + yz1]) * 0.5;
        dy[e] = (in[c - nz1] - in[c + nz1]) * 0.5;
        dz[e] = (in[c - 1] - in[c + 1]) * 0.5;
    }
}

__device__ __forceinline__ double trilerp1(const double* __restrict__ f, size_t b, size_t xs, size_t ys,
                                           double tx, double ty, double tz) {
    // utils_inline.h:55-70
    const double mx = 1.0 - tx, my = 1.0 - ty, mz = 1.0 - tz;
    const double f00 = f[b] * mx + f[b + xs] * tx;
    const double f10 = f[b + ys] * mx + f[b + xs + ys] * tx;
    const double fz0 = f00 * my + f10 * ty;
    const double g00 = f[b + 1] * mx + f[b + xs + 1] * tx;
    const double g10 = f[b + ys + 1] * mx + f[b + xs + ys + 1] * tx;
    const double fz1 = g00 * my + g10 * ty;
    return fz0 * mz + fz1 * tz;
}

// update_particles (utils.cpp:141-170).
__global__ __launch_bounds__(256) void k_update(const double* __restrict__ px, const double* __restrict__ py,
                                                const double* __restrict__ pz, const double* __restrict__ gx,
                                                const double* __restrict__ gy, const double* __restrict__ gz,
                                                double* __restrict__ vx, double* __restrict__ vy,
                                                double* __restrict__ vz, double fa_da, int64_t np, int nx, int ny,
                                                int nz, int x0, int y0, int z0,
                                                unsigned long long* __restrict__ err) {
    const int64_t p = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (p >= np) return;
    const double xp = px[p], yp = py[p], zp = pz[p];
    if (!(xp >= (double)x0 && xp < (double)(x0 + nx - 1) && yp >= (double)y0 && yp < (double)(y0 + ny - 1) &&
          zp >= (double)z0 && zp < (double)(z0 + nz - 1))) {
        atomicAdd(err, 1ULL);
        return;
    }
    const int ix = (int)xp, iy = (int)yp, iz = (int)zp;
    const double tx = xp - ix, ty = yp - iy, tz = zp - iz;
    const size_t b = ((size_t)(ix - x0) * ny + (iy - y0)) * nz + (iz - z0);
    const size_t xs = (size_t)ny * nz, ys = (size_t)nz;
    vx[p] += trilerp1(gx, b, xs, ys, tx, ty, tz) * fa_da;
    vy[p] += trilerp1(gy, b, xs, ys, tx, ty, tz) * fa_da;
    vz[p] += trilerp1(gz, b, xs, ys, tx, ty, tz) * fa_da;
}

static unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 16384); }

}  // namespace pm

using namespace pm;

extern "C" int pm_init_buf(double* buf, int nx, int ny, int nz, void* stream) {
    if (!buf || nx < 2 || ny < 2 || nz < 2) return fail(PM_ERR_ARG, "pm_init_buf: bad arguments");
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_init_buf, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, buf, nx, ny, nz);
    return check_launch("pm_init_buf");
}

extern "C" int pm_deposit_particles(const double* posx, const double* posy, const double* posz, int64_t np,
                                    double* buf, int nx, int ny, int nz, int x0, int y0, int z0, double mass,
                                    void* stream) {
    if (np < 0 || !buf || nx < 2 || ny < 2 || nz < 2 || (np > 0 && (!posx || !posy || !posz)))
        return fail(PM_ERR_ARG, "pm_deposit_particles: bad arguments");
    if (np == 0) return PM_OK;
    unsigned long long* err = range_counter();
    if (!err) return fail(PM_ERR_HIP, "pm_deposit_particles: range-error counter not found");
    const unsigned nb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_deposit_cube, dim3(nb), dim3(256), 0, (hipStream_t)stream, posx, posy, posz, np, buf,
                       nx, ny, nz, x0, y0, z0, mass, err);
    return check_launch("pm_deposit_particles");
}

extern "C" int pm_reorder_slab(const double* cube, double* slab, int ng, int stride, void* stream) {
    if (!cube || !slab || ng < 2 || stride < ng) return fail(PM_ERR_ARG, "pm_reorder_slab: bad arguments");
    const size_t n = (size_t)ng * ng * ng;
    hipLaunchKernelGGL(k_reorder_slab, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, cube, slab, ng,
                       stride);
    return check_launch("pm_reorder_slab");
}

extern "C" int pm_transpose_slab(const double* cube, double* recv, int ng, void* stream) {
    if (!cube || !recv || ng < 2) return fail(PM_ERR_ARG, "pm_transpose_slab: bad arguments");
    const size_t bytes = sizeof(double) * (size_t)(ng + 1) * (ng + 1) * (ng + 1);
    const hipError_t e = hipMemcpyAsync(recv, cube, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    if (e != hipSuccess) return fail(PM_ERR_HIP, std::string("pm_transpose_slab: ") + hipGetErrorString(e));
    return PM_OK;
}

extern "C" int pm_reorder_slab_back(const double* slab, double* out, int ng, int stride, void* stream) {
    if (!slab || !out || ng < 2 || stride < ng) return fail(PM_ERR_ARG, "pm_reorder_slab_back: bad arguments");
    const size_t n = (size_t)ng * (ng + 3) * (ng + 3);
    hipLaunchKernelGGL(k_reorder_slab_back, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, slab, out, ng,
                       stride);
    return check_launch("pm_reorder_slab_back");
}

extern "C" int pm_transpose_slab_back(const double* in, double* halo, int ng, void* stream) {
    if (!in || !halo || ng < 2) return fail(PM_ERR_ARG, "pm_transpose_slab_back: bad arguments");
    const size_t n = (size_t)(ng + 3) * (ng + 3) * (ng + 3);
    hipLaunchKernelGGL(k_transpose_slab_back, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, in, halo, ng);
    return check_launch("pm_transpose_slab_back");
}

extern "C" int pm_centered_derivatives(const double* in, double* d_dx, double* d_dy, double* d_dz, int nx1,
                                       int ny1, int nz1, void* stream) {
    if (!in || !d_dx || !d_dy || !d_dz || nx1 < 3 || ny1 < 3 || nz1 < 3)
        return fail(PM_ERR_ARG, "pm_centered_derivatives: bad arguments");
    const size_t n = (size_t)(nx1 - 2) * (ny1 - 2) * (nz1 - 2);
    hipLaunchKernelGGL(k_centered, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, in, d_dx, d_dy, d_dz,
                       nx1, ny1, nz1);
    return check_launch("pm_centered_derivatives");
}

extern "C" int pm_update_particles(const double* posx, const double* posy, const double* posz,
                                   const double* d_dx, const double* d_dy, const double* d_dz, double* px,
                                   double* py, double* pz, double a, int64_t np, int nx, int ny, int nz, int x0,
                                   int y0, int z0, void* stream) {
    if (np < 0 || !d_dx || !d_dy || !d_dz || nx < 2 || ny < 2 || nz < 2 || !(a > 0.0) ||
        (np > 0 && (!posx || !posy || !posz || !px || !py || !pz)))
        return fail(PM_ERR_ARG, "pm_update_particles: bad arguments");
    if (np == 0) return PM_OK;
    const double fa_da = pm_function_a(a) * 0.001;  // Da (utils_inline.h:17)
    unsigned long long* err = range_counter();
    if (!err) return fail(PM_ERR_HIP, "pm_update_particles: range-error counter not found");
    const unsigned nb = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, (hipStream_t)stream, posx, posy, posz, d_dx, d_dy, d_dz,
                       px, py, pz, fa_da, np, nx, ny, nz, x0, y0, z0, err);
    return check_launch("pm_update_particles");
}
