/* pm_api.h -- C ABI of the MI355X particle-mesh (PM) gravity step.
 *
 * One periodic box on one GPU.  Every pointer argument that names mesh or
 * particle data is a DEVICE pointer (hipMalloc'd, coarse-grained); sizes are
 * element counts; `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  Functions return 0 (PM_OK) or a negative PM_ERR_* code
 * and leave a message for pm_last_error() (thread-local).  Calls are
 * asynchronous with respect to the host unless stated otherwise.
 *
 * Conventions shared with the reference (yukiknk/Particle-Mesh):
 *   positions in cell units with node i at coordinate i, cell index by
 *   truncation (utils.cpp:79-88); meshes row-major [x][y][z], z fastest;
 *   density contrast delta = mass * CIC - 1 with mass = Ng^3/Np
 *   (main.cpp:23, utils.cpp:38-68); the centred difference
 *   (phi[i-1]-phi[i+1])/2 as the force (utils.cpp:115-137).
 *
 * Two groups of entry points:
 *   1. reference-layout operators, FP64, one per reference routine, taking
 *      the reference's buffers and sizes (the drop-in for src/utils,
 *      src/decomp/slab and src/fft);
 *   2. the fused one-box step on a plan (periodic indexing instead of ghost
 *      planes and halos, FP32 or FP64 mesh, FP64 positions/velocities).
 */
#ifndef PM_API_H
#define PM_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
#define PM_OK 0
#define PM_ERR_ARG (-1)      /* bad size, null pointer, unsupported mode */
#define PM_ERR_HIP (-2)      /* a HIP runtime call or kernel launch failed */
#define PM_ERR_NOMEM (-3)    /* device allocation failed */
#define PM_ERR_RANGE (-4)    /* particles outside the box (reported by a check) */

/* Precision of the mesh / FFT / force path. */
#define PM_FP32 0
#define PM_FP64 1

/* Green's functions (see pm_plan_create). */
#define PM_GREEN_REF_KH 0    /* -3 Omega0 / (4 (3 - sum cos 2 pi n_i/N)): fftw/fft.cpp:50-57 */
#define PM_GREEN_CONT_K2 1   /* -N^2 / (pi |n|^2): -4 pi G rho_bar / k^2, unit box, G = M = 1 */

/* Deposit variants. */
#define PM_DEPOSIT_ATOMIC 0  /* one thread per particle, 8 global float atomic adds */
#define PM_DEPOSIT_SORTED 1  /* particles binned by 16^3-cell tile, exact 64-bit fixed-point
                                accumulation in LDS, deterministic (no global float atomics) */
/* Flag for pm_step's variant: the positions are exactly those written by the
 * previous pm_step of this plan (same arrays, untouched since), so the
 * deposit can read the copy of them that step's interpolation staged in
 * tile order instead of binning the positions again.  Only honoured with
 * PM_DEPOSIT_SORTED (results are bit-identical with and without it);
 * pm_run_kdk sets it itself after its first step. */
#define PM_STEP_REUSE_BINS 0x100

/* Last error message of the calling thread ("" if none). */
const char* pm_last_error(void);
/* Library version string. */
const char* pm_version(void);

/* ---- 1. reference-layout operators (FP64) -------------------------------
 * The sizes mirror the reference arguments: nx = nx_loc + 1 etc. are the
 * cube sides with the + ghost plane; x0/y0/z0 the cube origin (0 for one
 * box); nx1 = nx + 2 etc. the sides of the -1/+2 halo cube. */

/* init_buf (utils.cpp:38-68) for one grid: -1 on [0,nx-1)x[0,ny-1)x[0,nz-1),
 * 0 on the + ghost planes.  The per-thread grids of the reference do not
 * exist here (the deposit uses atomics), so `nthreads` is not an argument. */
int pm_init_buf(double* buf, int nx, int ny, int nz, void* stream);

/* deposit_particles + cic_deposit + reduce_private_to_shared
 * (utils.cpp:70-113, utils_inline.h:25-49): adds mass * CIC weights of the
 * np particles to buf (nx*ny*nz).  A particle whose cell does not lie in
 * [x0, x0+nx-1) x ... is skipped and counted; pm_range_errors() reports the
 * count. */
int pm_deposit_particles(const double* posx, const double* posy, const double* posz, int64_t np,
                         double* buf, int nx, int ny, int nz, int x0, int y0, int z0,
                         double mass, void* stream);

/* TransposeSlab at one box (transpose_slab.cpp:3-39): the rank's cube goes
 * to the only slab owner, itself -- a device copy of the (ng+1)^3 cube. */
int pm_transpose_slab(const double* cube, double* recv, int ng, void* stream);

/* ReorderSlab at one box (reorder_slab.cpp:3-68): fold the + ghost planes of
 * the (ng+1)^3 cube periodically into the ng x ng x stride slab (stride >=
 * ng; padding columns untouched).  Deterministic gather, no atomics. */
int pm_reorder_slab(const double* cube, double* slab, int ng, int stride, void* stream);

/* ReorderSlabBack at one box (reorder_slab_back.cpp:3-49): slab ->
 * ng x (ng+3) x (ng+3) block with the -1/+2 periodic halo on y and z. */
int pm_reorder_slab_back(const double* slab, double* out, int ng, int stride, void* stream);

/* TransposeSlabBack at one box (transpose_slab_back.cpp:3-49): the x halo
 * (plane -1 from plane ng-1, planes ng, ng+1 from planes 0, 1) around the
 * ng planes of pm_reorder_slab_back's block -> the (ng+3)^3 halo cube. */
int pm_transpose_slab_back(const double* in, double* halo, int ng, void* stream);

/* centered_derivatives (utils.cpp:115-137): in is nx1*ny1*nz1, the outputs
 * (nx1-2)*(ny1-2)*(nz1-2); d_d?[i] = (in[i-1] - in[i+1]) * 0.5. */
int pm_centered_derivatives(const double* in, double* d_dx, double* d_dy, double* d_dz, int nx1,
                            int ny1, int nz1, void* stream);

/* update_particles + trilerp1 + function_a (utils.cpp:141-170,
 * utils_inline.h:51-70): p += trilerp(d_d?) * f(a) * Da with Da = 0.001.
 * The gradient meshes are (nx)*ny*nz with ny, nz the cube sides. Particles
 * outside the cube are skipped and counted as in pm_deposit_particles. */
int pm_update_particles(const double* posx, const double* posy, const double* posz,
                        const double* d_dx, const double* d_dy, const double* d_dz, double* px,
                        double* py, double* pz, double a, int64_t np, int nx, int ny, int nz,
                        int x0, int y0, int z0, void* stream);

/* Number of particles skipped by range checks since the last call
 * (synchronises the device; resets the counter). */
int64_t pm_range_errors(void);

/* f(a) = 1/sqrt(Om/a + Ok + OL a^2) (utils_inline.h:51-53), host-side. */
double pm_function_a(double a);

/* The FFT class (fftw/fft.h:10-48, ffte1/fft.h:12-42) at one box.  The real
 * slab is dense (stride = ng, the ffte1 layout); the half spectrum lives in
 * the handle, layout [kz][ky][kx], kz in [0, ng/2].  As in ffte1, backward()
 * is normalised by 1/ng^3 and apply_green() carries no 1/ng^3
 * (ffte1/fft.cpp:28-35, pzdfft3d.f:90). */
typedef struct pm_fft pm_fft_t;
int pm_fft_create(pm_fft_t** out, int ng, int precision, void* stream);
int pm_fft_destroy(pm_fft_t* fft);
/* slab (ng^3, mesh precision) -> spectrum */
int pm_fft_forward(pm_fft_t* fft, const void* slab);
/* spectrum *= G(k) * (1/a) with the KH Green's function of the reference */
int pm_fft_apply_green(pm_fft_t* fft, double a);
/* spectrum -> slab, times 1/ng^3 */
int pm_fft_backward(pm_fft_t* fft, void* slab);
/* Device pointer of the spectrum ((ng/2+1)*ng*ng complex, interleaved). */
void* pm_fft_spectrum(pm_fft_t* fft);

/* ---- particle snapshots (no counterpart in the reference) --------------- */
/* A versioned binary file: 64-byte header (magic "PMSNAP01", version 1,
 * header size, np, ng, step, a) then x, y, z, vx, vy, vz as float64 arrays
 * of np values.  Device arrays are copied through a pinned host buffer in
 * chunks on `stream`; the calls return when the file is complete. */
int pm_snapshot_write(const char* path, int64_t np, int ng, int step, double a, const double* posx,
                      const double* posy, const double* posz, const double* vx, const double* vy,
                      const double* vz, void* stream);
/* Header fields of a snapshot (any pointer may be NULL). */
int pm_snapshot_info(const char* path, int64_t* np, int* ng, int* step, double* a);
/* Read a snapshot of exactly np particles into device arrays. */
int pm_snapshot_read(const char* path, int64_t np, double* posx, double* posy, double* posz, double* vx,
                     double* vy, double* vz, void* stream);

/* ---- initial conditions (no counterpart in the reference) -------------- */
/* Zel'dovich ICs on a displaced nrow^3 particle lattice in an ng^3 mesh
 * (ng % nrow == 0): q = (i+1/2, j+1/2, k+1/2) * ng/nrow, psi_hat_a =
 * i k_a / |k|^3 * FFT[noise] (density spectrum ~ k^-2, k = 0 and the Nyquist
 * k_a removed), x = q + s psi wrapped into [0, ng), v = vel_factor * s psi,
 * with s set so that the rms of |s psi| is rms_cells.  noise is a white
 * Gaussian field, nrow^3 FP64, row-major.  Particle p = (i*nrow + j)*nrow + k.
 * Synchronous; allocates FFT scratch for the call. */
int pm_ic_zeldovich(int nrow, int ng, const double* noise, double rms_cells, double vel_factor,
                    double* posx, double* posy, double* posz, double* vx, double* vy, double* vz,
                    void* stream);

/* ---- 2. fused one-box step --------------------------------------------- */
typedef struct pm_plan pm_plan_t;

/* Create a plan for an ng^3 periodic mesh (ng a power of two in [8, 1024])
 * and np particles (1 <= np < 2^32).  green_mode selects the Green's function; green_scale
 * multiplies it (1/a for the reference's apply_green(a); 1 for CONT_K2).
 * The plan owns its meshes and FFT work space (about 4.2 ng^3 reals). */
int pm_plan_create(pm_plan_t** out, int ng, int64_t np, int precision, int green_mode,
                   double green_scale, void* stream);
int pm_plan_destroy(pm_plan_t* plan);
int pm_plan_set_stream(pm_plan_t* plan, void* stream);
/* Device pointers of the plan's density (after pm_deposit: mass*CIC, without
 * the -1) and potential (after pm_solve) meshes, ng^3 in the mesh precision.
 * The adjoint calls (section 3) use the density mesh as scratch for the
 * divergence sum_a D_a^T U_a, and pm_plan_tune overwrites both. */
void* pm_plan_density(pm_plan_t* plan);
void* pm_plan_potential(pm_plan_t* plan);

/* Zero the density mesh and deposit mass*CIC of the particles (periodic). */
int pm_deposit(pm_plan_t* plan, const double* posx, const double* posy, const double* posz,
               int variant);
/* potential = IFFT[G * scale / ng^3 * FFT[density - 1]] (5 passes). */
int pm_solve(pm_plan_t* plan);
/* v += g(x) * kick; x += v * drift wrapped into [0, ng), with g the
 * trilinear interpolation of the centred difference of the potential. */
int pm_interp_kick_drift(pm_plan_t* plan, double* posx, double* posy, double* posz, double* vx,
                         double* vy, double* vz, double kick, double drift);
/* Per-particle forces only (g into gx, gy, gz; FP64). */
int pm_interp_force(pm_plan_t* plan, const double* posx, const double* posy, const double* posz,
                    double* gx, double* gy, double* gz);
/* One force evaluation + kick + drift: deposit, solve, interp_kick_drift. */
int pm_step(pm_plan_t* plan, double* posx, double* posy, double* posz, double* vx, double* vy,
            double* vz, double kick, double drift, int variant);
/* nsteps kick-drift-kick steps of size dt with merged kicks (nsteps + 1
 * force evaluations: kick dt/2 first, dt between, dt/2 last; nsteps = 0
 * leaves the particles as they are). */
int pm_run_kdk(pm_plan_t* plan, double* posx, double* posy, double* posz, double* vx, double* vy,
               double* vz, int nsteps, double dt, int variant);
/* Cosmological kick-drift-kick in the expansion factor a, from a0 in nsteps
 * steps of da: the reference's kick p += g(a) f(a) da (update_particles,
 * utils.cpp:141-170, f(a) of utils_inline.h:51-53, g(a) with the 1/a of
 * apply_green(a), fftw/fft.cpp:75-82) with the drift it leaves commented
 * out (utils.cpp:145-146), x += p f(a) da / a^2 at the half step.  The plan's
 * Green's function is the one at a = 1 (e.g. PM_GREEN_REF_KH, scale 1); the
 * run divides each kick by a.  nsteps + 1 force evaluations: kick weights
 * 1/2, 1, ..., 1, 1/2.  nsteps = 0, da = 2 * 0.001, a0 = 0.9 is exactly the
 * reference's single update_particles at Start_a (main.cpp:83). */
int pm_run_cosmo(pm_plan_t* plan, double* posx, double* posy, double* posz, double* px, double* py,
                 double* pz, double a0, double da, int nsteps, int variant);

/* Split-precision particle state.  No counterpart in the reference, which
 * keeps FP64 positions and momenta (update_particles, utils.cpp:141-170):
 * this is the split representation of the 512^3 benchmark step.  The plan
 * holds the particles itself, in tile order, each position as an integer
 * cell plus a fixed-point fraction (a drift is an integer add that wraps at
 * the box edge) and each velocity in the mesh precision.  FP32-mesh plans:
 * one u32 per axis, x * 2^(32 - log2 ng), 2^-23 cells at ng = 512, and FP32
 * velocities, 24 bytes per particle against 48.  FP64-mesh plans: one u64 per
 * axis, x * 2^(64 - log2 ng), 2^-55 cells at ng = 512 (finer than an FP64
 * position), and FP64 velocities.  The deposit, solve and force are those of
 * pm_step with PM_DEPOSIT_SORTED (bit for bit from positions on the 2^-S
 * grid); the kick adds the force in the mesh precision, and the drift adds
 * v * drift rounded to the nearest 2^-S cells.
 * pm_split_load: copy the caller's FP64 particles in (positions rounded to
 * the nearest 2^-S cells, exact with an FP64 mesh; one outside [0, ng) is
 * counted as a range error and wrapped into the box).  pm_split_store: write
 * them back as FP64 (exact for the positions of an FP32 mesh; rounded to
 * FP64 with an FP64 mesh).  pm_split_step: one force evaluation + kick + drift on
 * the plan's particles.  pm_split_run_kdk: nsteps kick-drift-kick steps as
 * pm_run_kdk does.  A kick that leaves a velocity non-finite is counted as a
 * range error and leaves that particle as it was. */
int pm_split_load(pm_plan_t* plan, const double* posx, const double* posy, const double* posz,
                  const double* vx, const double* vy, const double* vz);
int pm_split_store(pm_plan_t* plan, double* posx, double* posy, double* posz, double* vx, double* vy,
                   double* vz);
int pm_split_step(pm_plan_t* plan, double kick, double drift);
int pm_split_run_kdk(pm_plan_t* plan, int nsteps, double dt);

/* ---- 3. adjoint (no counterpart in the reference: it has no gradients) ---
 * Reverse mode of the fused step.  With F(x) the per-particle force of
 * pm_interp_force, u a per-particle cotangent and m = ng^3/np:
 *   J_F^T u = sum_a u_a grad_x interp(g_a)  +  m grad_x interp(lambda),
 *   lambda  = S sum_a D_a^T U_a,  U_a = interp^T(u_a),
 * S the (symmetric) Green's-function solve and D_a the centred difference
 * (D_a^T = -D_a).  The adjoint entry points allocate their work meshes
 * (3 + 1 meshes and the binning) on first use.
 * The deposits of u (interpolation transpose, divergence deposit) sum in
 * signed 64-bit fixed point at a power-of-two scale set by max |u_a|: exact
 * and order-free, valid while no mesh node collects more than 2^22
 * particles (an overflow there wraps silently).  An inf or NaN component of
 * u has no such scale: it is counted in pm_range_errors() and the adjoint
 * meshes are then meaningless. */

/* Transposed CIC (adjoint of the deposit w.r.t. positions):
 * xbar += m * grad_x sum_c W(x, c) mesh_bar[c]; mesh_bar is ng^3 in the
 * plan's precision.  The gradient of a density loss goes through this. */
int pm_adjoint_deposit(pm_plan_t* plan, const double* posx, const double* posy, const double* posz,
                       const void* mesh_bar, double* xbar, double* ybar, double* zbar);
/* Interpolation transpose: U_a = sum_p W(x_p, .) u_p^a into the plan's
 * adjoint meshes 0..2 (LDS tiles over the tile binning of x). */
int pm_interp_transpose(pm_plan_t* plan, const double* posx, const double* posy, const double* posz,
                        const double* ux, const double* uy, const double* uz);
/* Adjoint FFT solve: lambda = S sum_a D_a^T U_a into adjoint mesh 3. */
int pm_adjoint_solve(pm_plan_t* plan);
/* Device pointer of adjoint mesh `which` (0..2: U_a, 3: lambda), or NULL.
 * pm_adjoint_force, pm_adjoint_step and pm_run_adjoint deposit the
 * divergence sum_a D_a^T U_a directly: after them mesh 0 holds it (meshes 1
 * and 2 are not written) and mesh 3 holds lambda. */
void* pm_plan_adjoint_mesh(pm_plan_t* plan, int which);
/* xbar += J_F(x)^T u (recomputes the potential at x). */
int pm_adjoint_force(pm_plan_t* plan, const double* posx, const double* posy, const double* posz,
                     const double* ux, const double* uy, const double* uz, double* xbar, double* ybar,
                     double* zbar, int variant);
/* One reverse step of pm_step(kick, drift) without a stored trajectory:
 * x -= drift v; vbar += drift xbar; v -= kick F(x); xbar += kick J_F^T vbar.
 * On entry (x, v) is the state after the forward step and (xbar, vbar) the
 * cotangent of it; on exit both refer to the state before the step. */
int pm_adjoint_step(pm_plan_t* plan, double* posx, double* posy, double* posz, double* vx, double* vy,
                    double* vz, double* xbar, double* ybar, double* zbar, double* vxbar, double* vybar,
                    double* vzbar, double kick, double drift, int variant);
/* Reverse sweep of pm_run_kdk(nsteps, dt): reverse-time leapfrog
 * recomputation, O(1) memory in the number of steps (nsteps = 0: nothing
 * changes). */
int pm_run_adjoint(pm_plan_t* plan, double* posx, double* posy, double* posz, double* vx, double* vy,
                   double* vz, double* xbar, double* ybar, double* zbar, double* vxbar, double* vybar,
                   double* vzbar, int nsteps, double dt, int variant);

/* Phase timing (hipEvents on the plan's stream).  pm_plan_enable_timing
 * reserves events for the next max_steps calls of pm_step (0 turns timing
 * off); recording needs no synchronisation.  pm_plan_phase_ms synchronises
 * and returns the accumulated milliseconds per phase
 * (0 zero+deposit, 1 z fwd, 2 y fwd, 3 x+green, 4 y inv, 5 z inv, 6 interp)
 * and the number of samples; pm_plan_reset_timing clears them. */
#define PM_NPHASES 7
int pm_plan_enable_timing(pm_plan_t* plan, int max_steps);
int pm_plan_phase_ms(pm_plan_t* plan, double* ms, int64_t* counts);
int pm_plan_reset_timing(pm_plan_t* plan);

/* ---- 4. FFT shapes and the tuning cache ----------------------------------
 * The analogue of the reference's FFTW wisdom (fftw/fft.cpp:28-49,63-70:
 * the FFT constructor imports wisdom_<n>_mpi_<size>.dat if present, plans
 * FFTW_WISDOM_ONLY, else plans FFTW_EXHAUSTIVE and exports at destruction).
 * Each of the five tunable FFT passes (slots 0 z fwd, 1 y fwd, 2 x (fused
 * forward + Green + inverse, or the forward / inverse alone), 3 y inv,
 * 4 z inv) runs in one of PM_FFT_VARIANTS shapes (lines per workgroup: the
 * default, half, a quarter; x threads per workgroup); variant 0 is the
 * built-in default, and every variant gives bit-identical results.  Meshes below 64 always use
 * variant 0.  Plans and FFT handles take the cached shape for their
 * (ng, precision) when they are created. */
#define PM_FFT_SLOTS 5
#define PM_FFT_VARIANTS 9
/* Time every variant of every slot on the plan's own buffers (`repeats`
 * launches each, after one warm-up; the density, potential and spectra are
 * overwritten, as FFTW_MEASURE overwrites its arrays), keep the fastest for
 * the plan and store it in the cache.  solve_ms (may be NULL) receives the
 * summed pass times of the chosen shape. */
int pm_plan_tune(pm_plan_t* plan, int repeats, double* solve_ms);
int pm_plan_set_fft_shape(pm_plan_t* plan, const int variants[PM_FFT_SLOTS]);
int pm_plan_get_fft_shape(const pm_plan_t* plan, int variants[PM_FFT_SLOTS]);
/* As pm_plan_tune for an FFT handle; slab (ng^3 reals, overwritten) is
 * the scratch real mesh, slot 2 is timed as forward + inverse x passes. */
int pm_fft_tune(pm_fft_t* fft, void* slab, int repeats, double* fft_ms);
int pm_fft_get_shape(const pm_fft_t* fft, int variants[PM_FFT_SLOTS]);
/* The cache as a text file: a "pm-wisdom 1" line, then one line per entry
 * "ng fp32|fp64 v0 v1 v2 v3 v4 solve_ms".  Import merges (an entry for the
 * same ng and precision is replaced); a malformed file imports nothing and
 * returns PM_ERR_ARG. */
int pm_wisdom_import(const char* path);
int pm_wisdom_export(const char* path);
/* *found = 1 and the variants if the cache holds (ng, precision), else 0. */
int pm_wisdom_lookup(int ng, int precision, int variants[PM_FFT_SLOTS], int* found);
void pm_wisdom_forget(void);
/* The device's streaming copy rate, the ceiling that the HBM-bound kernels
 * are measured against beside the 8 TB/s nameplate (no counterpart in the
 * reference, which reports times only, main.cpp:84-90 / utils.cpp:12-20):
 * `bytes` (a multiple of 16, at least 16 MiB) copied between two fresh
 * device buffers by one 16-byte load and store per thread, `repeats` timed
 * launches after two warm-ups on a private stream; *gbs receives bytes read
 * plus bytes written per second, in GB/s.  Allocates and frees the two
 * buffers. */
int pm_copy_rate(size_t bytes, int repeats, double* gbs);

/* ---- 5. deposit without device allocation -------------------------------
 * No counterpart in the reference: for callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The same sorted CIC deposit as pm_deposit(plan, ...,
 * PM_DEPOSIT_SORTED) -- rho (ng^3 reals of `precision`) = mass * CIC of the
 * np particles, mass = ng^3/np, periodic, bit for bit the plan's density --
 * without a plan: its binning buffers live in a fixed device array of
 * PM_NOALLOC_SCRATCH_MB MiB that the library holds on every device it runs
 * on until the process ends.  The call needs about 12 np + 0.14 ng^3 * (4 or
 * 8) bytes of it (e.g. up to 3.4 M particles on 256^3 FP32, 2.6 M on 256^3
 * FP64, 4.1 M on 128^3) and returns PM_ERR_ARG, touching nothing, when that
 * exceeds the array.  Out-of-box particles are skipped and counted
 * (pm_range_errors).  Asynchronous on `stream`.  Every call on one device
 * shares that device's array: calls that overlap on one device (several
 * streams or host threads at once) race and give wrong results. */
#define PM_NOALLOC_SCRATCH_MB 48
int pm_deposit_sorted_noalloc(const double* posx, const double* posy, const double* posz, int64_t np, int ng,
                              int precision, void* rho, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PM_API_H */
