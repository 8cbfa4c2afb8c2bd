// Zel'dovich initial conditions on the GPU: displaced particle lattice.
//
// The reference has no IC generator (particles are uniform random per rank,
// utils.cpp:22-36).  This builds the lattice ICs of the benchmark boxes:
//   q = (i + 1/2, j + 1/2, k + 1/2) * (ng / nrow)      (mesh cells)
//   psi_hat_a(k) = i k_a / |k|^3 * w_hat(k)              (P_delta ~ k^-2)
//   x = q + s psi,   v = vel_factor * s psi,
// with w a white Gaussian field given on the nrow^3 particle lattice,
// psi computed by the same Stockham passes in FP64, k = 0 and the Nyquist
// component of k_a removed, and s chosen so that the rms of |s psi| equals
// the requested displacement in mesh cells.
#include <vector>

#include "pm_internal.h"

namespace pm {

__global__ __launch_bounds__(256) void k_ic_component(const double2* __restrict__ spec, double2* __restrict__ out,
                                                      int N, int axis) {
    const size_t total = (size_t)(N / 2 + 1) * N * N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e % N);
        const int ky = (int)((e / N) % N);
        const int kz = (int)(e / ((size_t)N * N));
        const int sx = kx <= N / 2 ? kx : kx - N;
        const int sy = ky <= N / 2 ? ky : ky - N;
        const int sz = kz;  // kz in [0, N/2]
        const int ka = axis == 0 ? sx : axis == 1 ? sy : sz;
        const double k2 = (double)sx * sx + (double)sy * sy + (double)sz * sz;
        double f = 0.0;
        if (k2 > 0.0 && (ka < 0 ? -ka : ka) != N / 2) f = (double)ka / (k2 * sqrt(k2));
        const double2 v = spec[e];
        out[e] = double2{-f * v.y, f * v.x};  // i * f * v
    }
}

// Sum of squares of the displacement: one partial per workgroup (grid-stride
// over a fixed grid, tree in LDS), then one workgroup adds the partials in a
// fixed order, so the normalisation -- and the ICs -- are bitwise
// reproducible from run to run.
__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ a, const double* __restrict__ b,
                                               const double* __restrict__ c, size_t n, double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += a[i] * a[i] + b[i] * b[i] + c[i] * c[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int nparts,
                                                   double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// psx/psy/psz alias vx/vy/vz (the displacement is staged there): each
// element is read and then written by the same thread.
__global__ __launch_bounds__(256) void k_ic_place(const double* psx, const double* psy, const double* psz, int nrow,
                                                  double spacing, double box, const double* __restrict__ sumsq,
                                                  double rms, double vel_factor, double* __restrict__ px,
                                                  double* __restrict__ py, double* __restrict__ pz, double* vx,
                                                  double* vy, double* vz) {
    const size_t n = (size_t)nrow * nrow * nrow;
    const double cur = sqrt(*sumsq / (double)n);
    const double s = cur > 0.0 ? rms / cur : 0.0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % nrow);
        const int j = (int)((e / nrow) % nrow);
        const int i = (int)(e / ((size_t)nrow * nrow));
        const double dx = s * psx[e], dy = s * psy[e], dz = s * psz[e];
        double x = ((double)i + 0.5) * spacing + dx;
        double y = ((double)j + 0.5) * spacing + dy;
        double z = ((double)k + 0.5) * spacing + dz;
        x -= box * floor(x / box);
        y -= box * floor(y / box);
        z -= box * floor(z / box);
        px[e] = x < box ? x : 0.0;
        py[e] = y < box ? y : 0.0;
        pz[e] = z < box ? z : 0.0;
        vx[e] = vel_factor * dx;
        vy[e] = vel_factor * dy;
        vz[e] = vel_factor * dz;
    }
}

}  // namespace pm

using namespace pm;

extern "C" int pm_ic_zeldovich(int nrow, int ng, const double* noise, double rms_cells, double vel_factor,
                               double* posx, double* posy, double* posz, double* vx, double* vy, double* vz,
                               void* stream) {
    if (!valid_ng(nrow) || !valid_ng(ng) || ng % nrow != 0 || !noise || !posx || !posy || !posz || !vx || !vy ||
        !vz || !(rms_cells >= 0.0))
        return fail(PM_ERR_ARG, "pm_ic_zeldovich: bad arguments (nrow, ng powers of two in [8,1024], ng % nrow == 0)");
    hipStream_t st = (hipStream_t)stream;
    const int N = nrow;
    const size_t nspec = (size_t)(N / 2 + 1) * N * N;
    const size_t nreal = (size_t)N * N * N;
    FftTables tab;
    void *specA = nullptr, *specB = nullptr, *specC = nullptr, *acc = nullptr;
    int rc = tables_create(&tab, N, PM_FP64, st);
    auto alloc = [&](void** p, size_t bytes) {
        if (!rc && hipMalloc(p, bytes) != hipSuccess) rc = fail(PM_ERR_NOMEM, "pm_ic_zeldovich: hipMalloc");
    };
    alloc(&specA, nspec * sizeof(double2));
    alloc(&specB, nspec * sizeof(double2));
    alloc(&specC, nspec * sizeof(double2));
    const unsigned nb = (unsigned)std::min<size_t>((nreal + 255) / 256, 16384);
    alloc(&acc, sizeof(double) * (nb + 1));  // nb partial sums, then the total
    // psi components are staged in the caller's velocity arrays (nreal each)
    double* psi[3] = {vx, vy, vz};
    if (!rc) {
        const GreenArgs ga{PM_GREEN_REF_KH, 1.0, tab.cos_tab};
        int bad = fft_pass(PM_FP64, N, kPassZFwd, st, noise, specA, tab.twM, tab.twN, ga, 1.0, false);
        bad |= fft_pass(PM_FP64, N, kPassYFwd, st, specA, specB, tab.twM, tab.twN, ga, 1.0, false);
        bad |= fft_pass(PM_FP64, N, kPassXFwd, st, nullptr, specB, tab.twM, tab.twN, ga, 1.0, false);
        const unsigned nbs = (unsigned)std::min<size_t>((nspec + 255) / 256, 16384);
        for (int a = 0; a < 3 && !rc && !bad; ++a) {
            hipLaunchKernelGGL(k_ic_component, dim3(nbs), dim3(256), 0, st, (const double2*)specB, (double2*)specC, N, a);
            bad |= fft_pass(PM_FP64, N, kPassXInv, st, nullptr, specC, tab.twM, tab.twN, ga, 1.0, false);
            bad |= fft_pass(PM_FP64, N, kPassYInv, st, specC, specA, tab.twM, tab.twN, ga, 1.0, false);
            bad |= fft_pass(PM_FP64, N, kPassZInv, st, specA, psi[a], tab.twM, tab.twN, ga, 1.0, false);
            rc = check_launch("pm_ic_zeldovich: psi");
        }
        if (bad && !rc) rc = fail(PM_ERR_ARG, "pm_ic_zeldovich: unsupported FFT pass");
    }
    if (!rc) {
        double* part = (double*)acc;
        hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, st, psi[0], psi[1], psi[2], nreal, part);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (const double*)part, (int)nb, part + nb);
        hipLaunchKernelGGL(k_ic_place, dim3(nb), dim3(256), 0, st, psi[0], psi[1], psi[2], N, (double)ng / N,
                           (double)ng, (const double*)(part + nb), rms_cells, vel_factor, posx, posy, posz, vx, vy,
                           vz);
        rc = check_launch("pm_ic_zeldovich: place");
    }
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(PM_ERR_HIP, "pm_ic_zeldovich: sync");
    for (void* p : {specA, specB, specC, acc})
        if (p) (void)hipFree(p);
    tables_destroy(&tab);
    return rc;
}
