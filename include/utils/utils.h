// The reference's PM utility functions (utils.h:3-10 of the reference) for one
// GPU: same names and arguments, device pointers, each forwarding to the HIP
// kernels of the C ABI (include/pm_api.h).  They run on the default stream
// and throw std::runtime_error when a kernel launch fails.
#pragma once
#include <mpi.h>

#include <cstddef>

constexpr double Omega_0 = 1.0;
constexpr double Omega_m = 0.3;
constexpr double Omega_lambda = 0.7;
constexpr double Omega_k = 1 - Omega_m - Omega_lambda;
constexpr double Da = 0.001, Start_a = 0.9;

// Wall-clock seconds since the previous call (utils.cpp:3-10).
double getTime(double* now_time);
// Sum of phase times after the warm-up (utils.cpp:12-20): the maximum over
// the ranks of comm, added on its rank 0.
void report_max_time(double local_sec, MPI_Comm comm, const int i, const int worm_up, double& sum_time);
// utils.cpp:22-36 on the host (mt19937_64 seeds 3r, 3r+1, 3r+2), copied to
// the device arrays.
void init_particles(double* posx, double* posy, double* posz, size_t Np_local, int x0, int y0, int z0, int nx_loc,
                    int ny_loc, int nz_loc, int rank);
// utils.cpp:38-68 for the single (atomically shared) grid; nthreads unused.
void init_buf(double* buf, int nx, int ny, int nz, size_t ncell_local, int nthreads);
// utils.cpp:70-102.
void deposit_particles(const double* posx, const double* posy, const double* posz, size_t Np_local, double* buf,
                       size_t ncell_local, int nx, int ny, int nz, int x0, int y0, int z0, double mass);
// utils.cpp:104-113: nothing to reduce (the deposit accumulates atomically
// into one grid).
void reduce_private_to_shared(double* buf, size_t ncell_local, int nthreads);
// utils.cpp:115-137.
void centered_derivatives(const double* in, double* d_dx, double* d_dy, double* d_dz, int nx1, int ny1, int nz1);
// utils.cpp:141-170 (the cube side along x is taken equal to ny).
void update_particles(const double* posx, const double* posy, const double* posz, const double* d_dx,
                      const double* d_dy, const double* d_dz, double* px, double* py, double* pz, const double a,
                      const size_t Np_local, const int ny, const int nz, const int x0, const int y0, const int z0);
