/* FP64 CPU oracle for the particle-mesh (PM) gravity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This restates, in plain C with OpenMP, the
 * reference's single-box step (yukiknk/Particle-Mesh) so that the HIP path
 * can be checked against it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the reference that
 * GPU results are compared against or as the timed CPU baseline -- never as
 * part of the product path.
 *
 * Conventions (all functions):
 *   - positions are in cell units, x in [0, ng); the node of index i sits at
 *     coordinate i (utils.cpp:79-88: ix = (int)x, t = x - ix);
 *   - meshes are row-major [x][y][z], z fastest;
 *   - the half spectrum is [x][y][kz], kz in [0, ng/2], interleaved re/im;
 *   - FFTs are unnormalised (forward e^{-i}, inverse e^{+i}).
 *
 * Green's functions:
 *   ORC_GREEN_REF_KH  G(n) = -3*Omega0 / (4 (3 - sum_i cos(2 pi n_i/ng)))
 *                     (fftw/fft.cpp:50-57, ffte1/fft.cpp:28-35), G(0) = 0;
 *                     the caller's `scale` carries 1/a (fftw/fft.cpp:75-82).
 *   ORC_GREEN_CONT_K2 G(n) = -ng^2 / (pi |n|^2) with n the signed integer
 *                     wavevector: the continuous -4 pi G rho_bar / k^2 of a
 *                     unit box with G = 1, M = 1, rescaled so that the
 *                     reference's centred difference (phi[i-1]-phi[i+1])/2
 *                     is the acceleration in cell units per unit time^2.
 * Both are divided by ng^3 inside orc_solve (normalisation of the inverse).
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_GREEN_REF_KH = 0, ORC_GREEN_CONT_K2 = 1 };

/* ---- random inputs -------------------------------------------------- */
/* mt19937_64 + uniform_real_distribution<double>(lo, hi) as libstdc++ draws
 * them (utils.cpp:22-36). */
void orc_mt_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
/* The first 'count' raw mt19937_64 outputs for 'seed'. */
void orc_mt_raw(uint64_t seed, int64_t count, uint64_t* out);
/* init_particles for one rank owning the whole box (seeds 0, 1, 2). */
void orc_init_particles(int64_t np, int ng, double* posx, double* posy, double* posz);

/* ---- reference-layout operators (cube with ghosts) ------------------- */
/* init_buf for one thread grid: (ng+1)^3, -1 on owned nodes, 0 on + ghosts. */
void orc_init_cube(int ng, double* cube);
/* deposit_particles + cic_deposit into the (ng+1)^3 cube (no wrap). */
void orc_deposit_cube(int64_t np, const double* posx, const double* posy, const double* posz,
                      double mass, int ng, double* cube);
/* ReorderSlab at one box: fold the + ghost planes of the cube periodically
 * into an ng^3 slab (row stride ng). */
void orc_ghost_fold(int ng, const double* cube, double* slab);
/* ReorderSlabBack + TransposeSlabBack at one box: slab -> (ng+3)^3 cube with
 * a -1/+2 periodic halo on every axis. */
void orc_halo_expand(int ng, const double* slab, double* halo);
/* centered_derivatives: (ng+3)^3 halo -> three (ng+1)^3 node meshes. */
void orc_centered_derivatives(int ng, const double* halo, double* dx, double* dy, double* dz);
/* update_particles: trilinear interpolation of the three (ng+1)^3 meshes and
 * p += g * kick (the reference's kick is f(a)*Da). */
void orc_update_particles(int64_t np, const double* posx, const double* posy, const double* posz,
                          const double* dx, const double* dy, const double* dz, int ng,
                          double kick, double* px, double* py, double* pz);
/* f(a) = 1/sqrt(Om/a + Ok + OL a^2) (utils_inline.h:51-53). */
double orc_function_a(double a);

/* ---- FFT and Poisson solve ------------------------------------------- */
/* 3-D real-to-complex forward DFT of an ng^3 real mesh into the half
 * spectrum [ng][ng][ng/2+1] (complex interleaved).  ng must be a power of
 * two >= 2.  Returns 0 on success, -1 on a bad size. */
int orc_fft_r2c(int ng, const double* in, double* spec);
/* Unnormalised inverse of orc_fft_r2c. */
int orc_fft_c2r(int ng, const double* spec, double* out);
/* Green's function value for the signed-free index (i, j, k) in [0, ng). */
double orc_green(int mode, int ng, int i, int j, int k);
/* phi = IFFT[G * scale / ng^3 * FFT[delta]]. */
int orc_solve(int ng, int mode, double scale, const double* delta, double* phi);

/* ---- periodic (one-box) step ---------------------------------------- */
/* delta = mass * CIC(x) - 1 on the periodic ng^3 mesh (deposit + fold). */
void orc_deposit_periodic(int64_t np, const double* posx, const double* posy, const double* posz,
                          double mass, int ng, double* delta);
/* Per-particle acceleration: trilinear interpolation of the centred
 * difference (phi[i-1]-phi[i+1])/2 of the periodic potential. */
void orc_interp_force(int64_t np, const double* posx, const double* posy, const double* posz,
                      const double* phi, int ng, double* gx, double* gy, double* gz);
/* v += g*kick, then x += v*drift wrapped into [0, ng). */
void orc_interp_kick_drift(int64_t np, double* posx, double* posy, double* posz,
                           double* vx, double* vy, double* vz, const double* phi, int ng,
                           double kick, double drift);
/* n KDK steps of size dt: kick dt/2, drift dt, force, kick dt/2 per step,
 * one force evaluation per step (merged kicks).  Returns 0 on success. */
int orc_run_kdk(int64_t np, double* posx, double* posy, double* posz,
                double* vx, double* vy, double* vz, int ng, int nsteps, double dt,
                int green_mode, double green_scale);
/* Cosmological KDK in a (kick f(a) da / a, drift f(a) da / a^2 at the half
 * step), nsteps + 1 force evaluations; green_scale for a = 1. */
int orc_run_cosmo(int64_t np, double* posx, double* posy, double* posz, double* px, double* py, double* pz,
                  int ng, int nsteps, double a0, double da, int green_mode, double green_scale);

/* ---- threading ------------------------------------------------------- */
int orc_max_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
