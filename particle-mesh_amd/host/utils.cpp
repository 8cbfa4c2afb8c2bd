// The reference's utility functions (src/utils/utils.cpp) for one GPU, over
// the C ABI (include/pm_api.h).
#include "utils.h"

#include <hip/hip_runtime.h>
#include <time.h>

#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "pm_api.h"

namespace {
void check(int rc, const char* what) {
    if (rc != PM_OK) throw std::runtime_error(std::string(what) + ": " + pm_last_error());
}
}  // namespace

double getTime(double* now_time) {
    struct timespec tv;
    clock_gettime(CLOCK_REALTIME, &tv);
    const double told = *now_time;
    *now_time = (tv.tv_sec + (double)tv.tv_nsec * 1e-9);
    return *now_time - told;
}

void report_max_time(double local_sec, MPI_Comm comm, const int i, const int worm_up, double& sum_time) {
    if (i < worm_up) return;
    double global_max = 0.0;
    MPI_Reduce(&local_sec, &global_max, 1, MPI_DOUBLE, MPI_MAX, 0, comm);
    int rank = 0;
    MPI_Comm_rank(comm, &rank);
    if (rank == 0) sum_time += global_max;
    MPI_Barrier(comm);
}

void init_particles(double* posx, double* posy, double* posz, size_t Np_local, int x0, int y0, int z0, int nx_loc,
                    int ny_loc, int nz_loc, int rank) {
    // the reference's generators and distributions (utils.cpp:23-35), on the host
    std::mt19937_64 gen_x(static_cast<uint64_t>(rank) * 3ULL);
    std::mt19937_64 gen_y(static_cast<uint64_t>(rank) * 3ULL + 1);
    std::mt19937_64 gen_z(static_cast<uint64_t>(rank) * 3ULL + 2);
    std::uniform_real_distribution<double> rx(x0, x0 + nx_loc), ry(y0, y0 + ny_loc), rz(z0, z0 + nz_loc);
    std::vector<double> hx(Np_local), hy(Np_local), hz(Np_local);
    for (size_t p = 0; p < Np_local; ++p) {
        hx[p] = rx(gen_x);
        hy[p] = ry(gen_y);
        hz[p] = rz(gen_z);
    }
    const size_t bytes = Np_local * sizeof(double);
    if (hipMemcpy(posx, hx.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(posy, hy.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(posz, hz.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("init_particles: hipMemcpy failed");
}

void init_buf(double* buf, int nx, int ny, int nz, size_t, int) {
    check(pm_init_buf(buf, nx, ny, nz, nullptr), "init_buf");
}

void deposit_particles(const double* posx, const double* posy, const double* posz, size_t Np_local, double* buf,
                       size_t, int nx, int ny, int nz, int x0, int y0, int z0, double mass) {
    check(pm_deposit_particles(posx, posy, posz, (int64_t)Np_local, buf, nx, ny, nz, x0, y0, z0, mass, nullptr),
          "deposit_particles");
}

void reduce_private_to_shared(double*, size_t, int) {}

void centered_derivatives(const double* in, double* d_dx, double* d_dy, double* d_dz, int nx1, int ny1, int nz1) {
    check(pm_centered_derivatives(in, d_dx, d_dy, d_dz, nx1, ny1, nz1, nullptr), "centered_derivatives");
}

void update_particles(const double* posx, const double* posy, const double* posz, const double* d_dx,
                      const double* d_dy, const double* d_dz, double* px, double* py, double* pz, const double a,
                      const size_t Np_local, const int ny, const int nz, const int x0, const int y0, const int z0) {
    check(pm_update_particles(posx, posy, posz, d_dx, d_dy, d_dz, px, py, pz, a, (int64_t)Np_local, ny, ny, nz, x0, y0,
                              z0, nullptr),
          "update_particles");
}
