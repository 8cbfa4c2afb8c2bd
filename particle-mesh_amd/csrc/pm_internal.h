// Internal declarations shared by the PM HIP sources.
#pragma once

#include <string>
#include <vector>

#include "pm_api.h"
#include "pm_common.h"

namespace pm {

enum Pass {
    kPassZFwd = 0,
    kPassYFwd,
    kPassXFused,  // forward x FFT, Green's multiply, inverse x FFT
    kPassXFwd,
    kPassXInv,
    kPassYInv,
    kPassZInv,
    kPassGreen,   // standalone spectrum *= G
};

enum XMode { kXForward = 1, kXGreen = 2, kXInverse = 4 };

struct GreenArgs {
    int mode;
    double scale;            // multiplies G (1/a, 1/N^3, ...)
    const double* cos_tab;   // device, N entries cos(2 pi n / N)
};

// Device tables for one mesh side and precision.
struct FftTables {
    int ng = 0;
    int prec = PM_FP32;
    void* twM = nullptr;     // N/2 complex: exp(-2 pi i m / (N/2))
    void* twN = nullptr;     // N complex:   exp(-2 pi i m / N)
    double* cos_tab = nullptr;
};

int tables_create(FftTables* t, int ng, int prec, hipStream_t st);
void tables_destroy(FftTables* t);

// Launch one FFT pass (see pm_fft.hip).  Returns -1 for an unsupported N.
int fft_pass(int prec, int N, int pass, hipStream_t st, const void* in, void* out, const void* twM,
             const void* twN, const GreenArgs& ga, double norm, bool sub1);

// Particle kernels (pm_particles.hip).
int launch_fill(int prec, void* p, size_t n, double v, hipStream_t st);
int launch_deposit_atomic(int prec, const double* px, const double* py, const double* pz,
                          int64_t np, int ng, double mass, void* rho, unsigned long long* err,
                          hipStream_t st);
int launch_interp(int prec, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                  int64_t np, int ng, const void* phi, double kick, double drift, int mode,
                  unsigned long long* err, hipStream_t st);

// Error reporting.
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);
unsigned long long* range_counter();

inline bool valid_ng(int ng) { return ng >= 8 && ng <= 1024 && (ng & (ng - 1)) == 0; }

}  // namespace pm

// Plan of the fused one-box step.
struct pm_plan {
    int ng = 0;
    int64_t np = 0;
    int prec = PM_FP32;
    int green_mode = PM_GREEN_REF_KH;
    double green_scale = 1.0;
    hipStream_t stream = nullptr;
    void* rho = nullptr;     // ng^3 reals
    void* phi = nullptr;     // ng^3 reals
    void* specA = nullptr;   // (ng/2+1) ng^2 complex
    void* specB = nullptr;
    pm::FftTables tab;
    // Phase timing: one set of PM_NPHASES+1 events per recorded step, read
    // back (and summed) only by pm_plan_phase_ms, so that recording needs no
    // synchronisation inside a timed loop.
    int timing_slots = 0;
    int timing_used = 0;
    std::vector<hipEvent_t> ev;
    double phase_ms[PM_NPHASES] = {};
    int64_t phase_n[PM_NPHASES] = {};
};

struct pm_fft {
    int ng = 0;
    int prec = PM_FP64;
    hipStream_t stream = nullptr;
    void* specA = nullptr;
    void* specB = nullptr;
    pm::FftTables tab;
};
