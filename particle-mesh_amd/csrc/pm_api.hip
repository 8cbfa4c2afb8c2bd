// C ABI glue: error state, FFT tables, the FFT handle and the fused-step plan.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "pm_internal.h"

namespace pm {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return PM_OK;
}

static int hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return PM_OK;
    return fail(e == hipErrorOutOfMemory ? PM_ERR_NOMEM : PM_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

// One range-error counter per device: a zero-initialised device variable of
// this module, so that resolving it allocates nothing and cannot run out of
// memory.  Its address is looked up once per device (under a lock: plans on
// one device may be driven from several host threads).
static __device__ unsigned long long range_count_dev;
static unsigned long long* g_range[64] = {};
static std::mutex g_range_mu;

unsigned long long* range_counter() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(g_range_mu);
    if (!g_range[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(range_count_dev)) != hipSuccess) return nullptr;
        g_range[dev] = (unsigned long long*)p;
    }
    return g_range[dev];
}

int tables_create(FftTables* t, int ng, int prec, hipStream_t st) {
    (void)st;
    t->ng = ng;
    t->prec = prec;
    const int M = ng / 2;
    const size_t es = prec == PM_FP32 ? sizeof(float2) : sizeof(double2);
    std::vector<double> twm(2 * (size_t)M), twn(2 * (size_t)ng), ct(ng);
    const double two_pi = 2.0 * M_PI;
    // exp(-2 pi i m / L), with tw[m + L/2] = -tw[m] exactly: the FFT stages
    // read only the first half of a table
    auto fill = [&](std::vector<double>& tw, int L) {
        for (int m = 0; m < L; ++m) {
            if (m < L / 2 || L < 2) {
                tw[2 * (size_t)m] = std::cos(two_pi * m / L);
                tw[2 * (size_t)m + 1] = -std::sin(two_pi * m / L);
            } else {
                tw[2 * (size_t)m] = -tw[2 * (size_t)(m - L / 2)];
                tw[2 * (size_t)m + 1] = -tw[2 * (size_t)(m - L / 2) + 1];
            }
        }
    };
    fill(twm, M);
    fill(twn, ng);
    for (int m = 0; m < ng; ++m) ct[m] = std::cos(two_pi / ng * m);  // same expression as ffte1/fft.cpp:29,33
    int rc = hip_ok(hipMalloc(&t->twM, es * M), "tables: hipMalloc");
    if (rc) return rc;
    rc = hip_ok(hipMalloc(&t->twN, es * ng), "tables: hipMalloc");
    if (rc) return rc;
    rc = hip_ok(hipMalloc((void**)&t->cos_tab, sizeof(double) * ng), "tables: hipMalloc");
    if (rc) return rc;
    if (prec == PM_FP32) {
        std::vector<float> fm(twm.begin(), twm.end()), fn(twn.begin(), twn.end());
        rc = hip_ok(hipMemcpy(t->twM, fm.data(), es * M, hipMemcpyHostToDevice), "tables: copy");
        if (!rc) rc = hip_ok(hipMemcpy(t->twN, fn.data(), es * ng, hipMemcpyHostToDevice), "tables: copy");
    } else {
        rc = hip_ok(hipMemcpy(t->twM, twm.data(), es * M, hipMemcpyHostToDevice), "tables: copy");
        if (!rc) rc = hip_ok(hipMemcpy(t->twN, twn.data(), es * ng, hipMemcpyHostToDevice), "tables: copy");
    }
    if (!rc) rc = hip_ok(hipMemcpy(t->cos_tab, ct.data(), sizeof(double) * ng, hipMemcpyHostToDevice), "tables: copy");
    return rc;
}

void tables_destroy(FftTables* t) {
    if (t->twM) (void)hipFree(t->twM);
    if (t->twN) (void)hipFree(t->twN);
    if (t->cos_tab) (void)hipFree(t->cos_tab);
    t->twM = t->twN = nullptr;
    t->cos_tab = nullptr;
}

static size_t spec_elems(int ng) { return (size_t)(ng / 2 + 1) * ng * ng; }

}  // namespace pm

using namespace pm;

extern "C" const char* pm_last_error(void) { return g_error.c_str(); }
extern "C" const char* pm_version(void) { return "pm-mi355x 0.1 (gfx950)"; }

extern "C" int64_t pm_range_errors(void) {
    unsigned long long* c = range_counter();
    if (!c) return fail(PM_ERR_HIP, "pm_range_errors: no counter");
    unsigned long long h = 0;
    if (hip_ok(hipDeviceSynchronize(), "pm_range_errors: sync")) return PM_ERR_HIP;
    if (hip_ok(hipMemcpy(&h, c, sizeof(h), hipMemcpyDeviceToHost), "pm_range_errors: copy")) return PM_ERR_HIP;
    if (hip_ok(hipMemset(c, 0, sizeof(h)), "pm_range_errors: reset")) return PM_ERR_HIP;
    return (int64_t)h;
}

extern "C" double pm_function_a(double a) {
    // utils_inline.h:13-16,51-53
    const double om = 0.3, ol = 0.7, ok = 1.0 - om - ol;
    return 1.0 / std::sqrt(om / a + ok + ol * a * a);
}

// ------------------------------------------------------------------ FFT API
extern "C" int pm_fft_create(pm_fft_t** out, int ng, int precision, void* stream) {
    if (!out) return fail(PM_ERR_ARG, "pm_fft_create: null handle");
    *out = nullptr;
    if (!valid_ng(ng)) return fail(PM_ERR_ARG, "pm_fft_create: ng must be a power of two in [8, 1024]");
    if (precision != PM_FP32 && precision != PM_FP64) return fail(PM_ERR_ARG, "pm_fft_create: bad precision");
    pm_fft* f = new pm_fft();
    f->ng = ng;
    f->prec = precision;
    f->stream = (hipStream_t)stream;
    wisdom_find(ng, precision, &f->shape);  // the cached shape, else the default
    const size_t cs = precision == PM_FP32 ? sizeof(float2) : sizeof(double2);
    int rc = tables_create(&f->tab, ng, precision, f->stream);
    if (!rc) rc = hip_ok(hipMalloc(&f->specA, cs * spec_elems(ng)), "pm_fft_create: hipMalloc");
    if (!rc) rc = hip_ok(hipMalloc(&f->specB, cs * spec_elems(ng)), "pm_fft_create: hipMalloc");
    if (rc) {
        pm_fft_destroy(f);
        return rc;
    }
    *out = f;
    return PM_OK;
}

extern "C" int pm_fft_destroy(pm_fft_t* f) {
    if (!f) return PM_OK;
    tables_destroy(&f->tab);
    if (f->specA) (void)hipFree(f->specA);
    if (f->specB) (void)hipFree(f->specB);
    delete f;
    return PM_OK;
}

extern "C" int pm_fft_forward(pm_fft_t* f, const void* slab) {
    if (!f || !slab) return fail(PM_ERR_ARG, "pm_fft_forward: null argument");
    const GreenArgs ga{PM_GREEN_REF_KH, 1.0, f->tab.cos_tab};
    const void *tm = f->tab.twM, *tn = f->tab.twN;
    int bad = fft_pass(f->prec, f->ng, kPassZFwd, f->stream, slab, f->specA, tm, tn, ga, 1.0, false, &f->shape);
    bad |= fft_pass(f->prec, f->ng, kPassYFwd, f->stream, f->specA, f->specB, tm, tn, ga, 1.0, false, &f->shape);
    bad |= fft_pass(f->prec, f->ng, kPassXFwd, f->stream, nullptr, f->specB, tm, tn, ga, 1.0, false, &f->shape);
    if (bad) return fail(PM_ERR_ARG, "pm_fft_forward: unsupported FFT pass");
    return check_launch("pm_fft_forward");
}

extern "C" int pm_fft_apply_green(pm_fft_t* f, double a) {
    if (!f || !(a > 0.0)) return fail(PM_ERR_ARG, "pm_fft_apply_green: bad argument");
    const GreenArgs ga{PM_GREEN_REF_KH, 1.0 / a, f->tab.cos_tab};
    if (fft_pass(f->prec, f->ng, kPassGreen, f->stream, nullptr, f->specB, f->tab.twM, f->tab.twN, ga, 1.0, false))
        return fail(PM_ERR_ARG, "pm_fft_apply_green: unsupported FFT pass");
    return check_launch("pm_fft_apply_green");
}

extern "C" int pm_fft_backward(pm_fft_t* f, void* slab) {
    if (!f || !slab) return fail(PM_ERR_ARG, "pm_fft_backward: null argument");
    const GreenArgs ga{PM_GREEN_REF_KH, 1.0, f->tab.cos_tab};
    const double n3 = (double)f->ng * f->ng * f->ng;
    const void *tm = f->tab.twM, *tn = f->tab.twN;
    int bad = fft_pass(f->prec, f->ng, kPassXInv, f->stream, nullptr, f->specB, tm, tn, ga, 1.0, false, &f->shape);
    bad |= fft_pass(f->prec, f->ng, kPassYInv, f->stream, f->specB, f->specA, tm, tn, ga, 1.0, false, &f->shape);
    bad |= fft_pass(f->prec, f->ng, kPassZInv, f->stream, f->specA, slab, tm, tn, ga, 1.0 / n3, false, &f->shape);
    if (bad) return fail(PM_ERR_ARG, "pm_fft_backward: unsupported FFT pass");
    return check_launch("pm_fft_backward");
}

extern "C" void* pm_fft_spectrum(pm_fft_t* f) { return f ? f->specB : nullptr; }

// ------------------------------------------------------------------ plan API
extern "C" int pm_plan_destroy(pm_plan_t* p) {
    if (!p) return PM_OK;
    tables_destroy(&p->tab);
    void* bufs[14] = {p->rho,          p->phi,           p->specA,        p->specB,        p->sorted.cnt,
                      p->sorted.off,   p->sorted.key,    p->sorted.slot,  p->sorted.idx,   p->sorted.shell,
                      p->sorted.fill,  p->sorted.coff,   p->sorted.spill, p->sorted.stage};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (p->split_buf) (void)hipFree(p->split_buf);
    void* abufs[7] = {p->adjU[0], p->adjU[1], p->adjU[2], p->adjLam, p->adjShell, p->adjMax, p->adjUst};
    for (void* b : abufs)
        if (b) (void)hipFree(b);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->adjFork) (void)hipEventDestroy(p->adjFork);
    if (p->adjJoin) (void)hipEventDestroy(p->adjJoin);
    if (p->adjStream) (void)hipStreamDestroy(p->adjStream);
    delete p;
    return PM_OK;
}

extern "C" int pm_plan_create(pm_plan_t** out, int ng, int64_t np, int precision, int green_mode,
                              double green_scale, void* stream) {
    if (!out) return fail(PM_ERR_ARG, "pm_plan_create: null handle");
    *out = nullptr;
    if (!valid_ng(ng)) return fail(PM_ERR_ARG, "pm_plan_create: ng must be a power of two in [8, 1024]");
    if (np <= 0) return fail(PM_ERR_ARG, "pm_plan_create: np must be positive");
    // tile keys, slots and the tile-ordered index list are 32-bit
    if (np > (int64_t)0xFFFFFFFFll) return fail(PM_ERR_ARG, "pm_plan_create: np must be below 2^32");
    if (precision != PM_FP32 && precision != PM_FP64) return fail(PM_ERR_ARG, "pm_plan_create: bad precision");
    if (green_mode != PM_GREEN_REF_KH && green_mode != PM_GREEN_CONT_K2)
        return fail(PM_ERR_ARG, "pm_plan_create: bad green mode");
    if (!std::isfinite(green_scale)) return fail(PM_ERR_ARG, "pm_plan_create: bad green scale");
    pm_plan* p = new pm_plan();
    p->ng = ng;
    p->np = np;
    p->prec = precision;
    p->green_mode = green_mode;
    p->green_scale = green_scale;
    p->stream = (hipStream_t)stream;
    wisdom_find(ng, precision, &p->fft_shape);  // the cached shape, else the default
    const size_t rs = precision == PM_FP32 ? sizeof(float) : sizeof(double);
    const size_t nc = (size_t)ng * ng * ng;
    p->err = range_counter();
    int rc = p->err ? PM_OK : fail(PM_ERR_HIP, "pm_plan_create: range-error counter not found");
    if (!rc) rc = tables_create(&p->tab, ng, precision, p->stream);
    if (!rc) rc = hip_ok(hipMalloc(&p->rho, rs * nc), "pm_plan_create: hipMalloc rho");
    if (!rc) rc = hip_ok(hipMalloc(&p->phi, rs * nc), "pm_plan_create: hipMalloc phi");
    if (!rc) rc = hip_ok(hipMalloc(&p->specA, 2 * rs * spec_elems(ng)), "pm_plan_create: hipMalloc spec");
    if (!rc) rc = hip_ok(hipMalloc(&p->specB, 2 * rs * spec_elems(ng)), "pm_plan_create: hipMalloc spec");
    if (!rc) rc = launch_fill(precision, p->phi, nc, 0.0, p->stream);
    if (rc) {
        pm_plan_destroy(p);
        return rc;
    }
    *out = p;
    return PM_OK;
}

extern "C" int pm_plan_set_stream(pm_plan_t* p, void* stream) {
    if (!p) return fail(PM_ERR_ARG, "pm_plan_set_stream: null plan");
    p->stream = (hipStream_t)stream;
    return PM_OK;
}

extern "C" void* pm_plan_density(pm_plan_t* p) { return p ? p->rho : nullptr; }
extern "C" void* pm_plan_potential(pm_plan_t* p) { return p ? p->phi : nullptr; }

// Record event i of the current timing slot (no-op when timing is off or
// every slot is used).
static inline void mark(pm_plan* p, int i) {
    if (p->timing_used < p->timing_slots)
        (void)hipEventRecord(p->ev[(size_t)p->timing_used * (PM_NPHASES + 1) + i], p->stream);
}

static inline void next_slot(pm_plan* p) {
    if (p->timing_used < p->timing_slots) ++p->timing_used;
}

// Sum the elapsed time of every recorded slot into phase_ms (synchronises).
static void collect(pm_plan* p) {
    for (int s = 0; s < p->timing_used; ++s) {
        hipEvent_t* e = &p->ev[(size_t)s * (PM_NPHASES + 1)];
        (void)hipEventSynchronize(e[PM_NPHASES]);
        for (int i = 0; i < PM_NPHASES; ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e[i], e[i + 1]) == hipSuccess) {
                p->phase_ms[i] += ms;
                p->phase_n[i] += 1;
            }
        }
    }
    p->timing_used = 0;
}

static int ensure_sorted(pm_plan* p) {
    SortedScratch& s = p->sorted;
    if (s.idx) return PM_OK;
    const int ts = sorted_tile(p->ng);
    const size_t nt = (size_t)(p->ng / ts);
    const size_t ntiles = nt * nt * nt;
    const size_t rs = p->prec == PM_FP32 ? sizeof(float) : sizeof(double);
    int rc = hip_ok(hipMalloc((void**)&s.cnt, 4 * ntiles), "sorted deposit: hipMalloc");
    if (!rc) {  // a fill kernel, as everywhere on the step path (no memset nodes if a caller captures it)
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3((unsigned)std::min<size_t>((ntiles + 255) / 256, 8192)), dim3(256),
                           0, p->stream, s.cnt, ntiles, 0u);
        rc = check_launch("sorted deposit: clear counts");
    }
    if (!rc) rc = hip_ok(hipMalloc((void**)&s.off, 4 * (ntiles + 1)), "sorted deposit: hipMalloc");
    if (!rc) rc = hip_ok(hipMalloc((void**)&s.key, 4 * (size_t)p->np), "sorted deposit: hipMalloc");
    if (!rc) rc = hip_ok(hipMalloc((void**)&s.slot, 4 * (size_t)p->np), "sorted deposit: hipMalloc");
    if (!rc) rc = hip_ok(hipMalloc(&s.shell, rs * sorted_shell_elems(p->ng)), "sorted deposit: hipMalloc");
    if (!rc) rc = hip_ok(hipMalloc((void**)&s.idx, 4 * (size_t)p->np), "sorted deposit: hipMalloc");
    // staging buckets: count + count/8 + 32 slots per tile (u32 slot numbers)
    const size_t cap_total = (size_t)p->np + (size_t)p->np / 8 + 32 * ntiles;
    if (!rc && cap_total < 0xFFFFFFF0ull) {
        rc = hip_ok(hipMalloc((void**)&s.fill, 4 * ntiles), "sorted deposit: hipMalloc");
        if (!rc) rc = hip_ok(hipMalloc((void**)&s.coff, 4 * (ntiles + 1)), "sorted deposit: hipMalloc");
        if (!rc) rc = hip_ok(hipMalloc((void**)&s.spill, 4), "sorted deposit: hipMalloc");
        const size_t rec = rs == 4 ? sizeof(StageQ) : sizeof(StageD);
        if (!rc) rc = hip_ok(hipMalloc(&s.stage, rec * cap_total), "sorted deposit: hipMalloc");
        if (!rc) s.cap_total = cap_total;
    }
    if (rc) {
        void* bufs[10] = {s.cnt, s.off, s.key, s.slot, s.shell, s.idx, s.fill, s.coff, s.spill, s.stage};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        s = SortedScratch();
    }
    return rc;
}

// Other users of the binning buffers or of x end the step-to-step reuse.
static void invalidate_reuse(pm_plan* p) { p->staged_valid = false; }

static int deposit_impl(pm_plan* p, const double* x, const double* y, const double* z, int variant) {
    const size_t nc = (size_t)p->ng * p->ng * p->ng;
    const double mass = (double)nc / (double)p->np;
    if (variant == PM_DEPOSIT_SORTED) {
        int rc = ensure_sorted(p);
        if (rc) return rc;
        invalidate_reuse(p);
        return launch_deposit_sorted(p->prec, x, y, z, p->np, p->ng, mass, p->rho, p->sorted, p->err,
                                     p->stream);
    }
    if (variant != PM_DEPOSIT_ATOMIC) return fail(PM_ERR_ARG, "pm_deposit: unknown variant");
    int rc = launch_fill(p->prec, p->rho, nc, 0.0, p->stream);
    if (rc) return rc;
    return launch_deposit_atomic(p->prec, x, y, z, p->np, p->ng, mass, p->rho, p->err, p->stream);
}

// potential = IFFT[G * scale / N^3 * FFT[in - sub1]], in -> out (5 passes)
static int solve_into(pm_plan* p, const void* in, void* out, bool sub1) {
    const int N = p->ng;
    const GreenArgs ga{p->green_mode, p->green_scale / ((double)N * N * N), p->tab.cos_tab};
    const void* tm = p->tab.twM;
    const void* tn = p->tab.twN;
    int bad = fft_pass(p->prec, N, kPassZFwd, p->stream, in, p->specA, tm, tn, ga, 1.0, sub1, &p->fft_shape);
    mark(p, 2);
    bad |= fft_pass(p->prec, N, kPassYFwd, p->stream, p->specA, p->specB, tm, tn, ga, 1.0, false, &p->fft_shape);
    mark(p, 3);
    bad |= fft_pass(p->prec, N, kPassXFused, p->stream, nullptr, p->specB, tm, tn, ga, 1.0, false, &p->fft_shape);
    mark(p, 4);
    bad |= fft_pass(p->prec, N, kPassYInv, p->stream, p->specB, p->specA, tm, tn, ga, 1.0, false, &p->fft_shape);
    mark(p, 5);
    bad |= fft_pass(p->prec, N, kPassZInv, p->stream, p->specA, out, tm, tn, ga, 1.0, false, &p->fft_shape);
    mark(p, 6);
    if (bad) return fail(PM_ERR_ARG, "pm_solve: unsupported FFT pass");
    return check_launch("pm_solve");
}

static int solve_impl(pm_plan* p) { return solve_into(p, p->rho, p->phi, true); }

extern "C" int pm_deposit(pm_plan_t* p, const double* x, const double* y, const double* z, int variant) {
    if (!p || !x || !y || !z) return fail(PM_ERR_ARG, "pm_deposit: null argument");
    return deposit_impl(p, x, y, z, variant);
}

extern "C" int pm_solve(pm_plan_t* p) {
    if (!p) return fail(PM_ERR_ARG, "pm_solve: null plan");
    const int slots = p->timing_slots;
    p->timing_slots = 0;  // only pm_step records phases
    const int rc = solve_impl(p);
    p->timing_slots = slots;
    return rc;
}

extern "C" int pm_interp_kick_drift(pm_plan_t* p, double* x, double* y, double* z, double* vx, double* vy,
                                    double* vz, double kick, double drift) {
    if (!p || !x || !y || !z || !vx || !vy || !vz) return fail(PM_ERR_ARG, "pm_interp_kick_drift: null argument");
    invalidate_reuse(p);
    return launch_interp(p->prec, x, y, z, vx, vy, vz, p->np, p->ng, p->phi, kick, drift, 0, p->err,
                         p->stream);
}

extern "C" int pm_interp_force(pm_plan_t* p, const double* x, const double* y, const double* z, double* gx,
                               double* gy, double* gz) {
    if (!p || !x || !y || !z || !gx || !gy || !gz) return fail(PM_ERR_ARG, "pm_interp_force: null argument");
    return launch_interp(p->prec, (double*)x, (double*)y, (double*)z, gx, gy, gz, p->np, p->ng, p->phi, 0.0, 0.0,
                         1, p->err, p->stream);
}

extern "C" int pm_step(pm_plan_t* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                       double kick, double drift, int variant) {
    if (!p || !x || !y || !z || !vx || !vy || !vz) return fail(PM_ERR_ARG, "pm_step: null argument");
    const bool reuse = (variant & PM_STEP_REUSE_BINS) != 0;
    variant &= ~PM_STEP_REUSE_BINS;
    if (variant != PM_DEPOSIT_SORTED && variant != PM_DEPOSIT_ATOMIC) return fail(PM_ERR_ARG, "pm_step: bad variant");
    mark(p, 0);
    int rc;
    const double mass = (double)p->ng * p->ng * p->ng / (double)p->np;
    bool staged = false;  // deposit from / interpolation into the staging buckets
    if (variant == PM_DEPOSIT_SORTED) {
        rc = ensure_sorted(p);
        staged = !rc && p->sorted.stage != nullptr;
        if (rc) return rc;
        if (staged && reuse && p->staged_valid && p->staged_x == x) {
            // positions staged in tile order by the previous step's
            // interpolation (caller's guarantee that x is unchanged since)
            rc = launch_deposit_staged(p->prec, x, y, z, p->np, p->ng, mass, p->rho, p->sorted, p->err,
                                       p->stream);
        } else {
            rc = launch_bin(x, y, z, p->np, p->ng, p->sorted, p->err, p->stream);
            if (!rc) rc = launch_deposit_binned(p->prec, x, y, z, p->ng, mass, p->rho, p->sorted, p->stream);
            if (!rc && staged) rc = launch_cap_scan(p->ng, p->sorted, true, p->stream);
        }
    } else {
        rc = deposit_impl(p, x, y, z, variant);
    }
    invalidate_reuse(p);
    if (rc) return rc;
    mark(p, 1);
    rc = solve_impl(p);
    if (rc) return rc;
    // particles in memory order (a coherent order keeps the potential's
    // gathers in L2); with the sorted deposit the drifted positions are
    // staged in tile order for the next step's deposit
    rc = launch_interp(p->prec, x, y, z, vx, vy, vz, p->np, p->ng, p->phi, kick, drift, staged ? 2 : 0,
                       p->err, p->stream, staged ? &p->sorted : nullptr);
    if (staged) {
        p->staged_valid = rc == PM_OK;
        p->staged_x = x;
    }
    if (rc) return rc;
    mark(p, 7);
    next_slot(p);
    return PM_OK;
}

extern "C" int pm_run_kdk(pm_plan_t* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                          int nsteps, double dt, int variant) {
    if (!p || nsteps < 0) return fail(PM_ERR_ARG, "pm_run_kdk: bad arguments");
    if (nsteps == 0) return PM_OK;  // no step: no kick either
    variant &= ~PM_STEP_REUSE_BINS;
    for (int s = 0; s <= nsteps; ++s) {
        const double kick = (s == 0 || s == nsteps) ? 0.5 * dt : dt;
        const double drift = s == nsteps ? 0.0 : dt;
        const int rc = pm_step(p, x, y, z, vx, vy, vz, kick, drift, s == 0 ? variant : variant | PM_STEP_REUSE_BINS);
        if (rc) return rc;
    }
    return PM_OK;
}

extern "C" int pm_run_cosmo(pm_plan_t* p, double* x, double* y, double* z, double* px, double* py, double* pz,
                            double a0, double da, int nsteps, int variant) {
    if (!p || nsteps < 0 || !(a0 > 0.0) || !(a0 + nsteps * da > 0.0))
        return fail(PM_ERR_ARG, "pm_run_cosmo: bad arguments (a must stay positive)");
    variant &= ~PM_STEP_REUSE_BINS;
    for (int s = 0; s <= nsteps; ++s) {
        const double a = a0 + s * da, ah = a + 0.5 * da;
        const double w = (s == 0 || s == nsteps) ? 0.5 : 1.0;
        const double kick = w * da * pm_function_a(a) / a;
        const double drift = s == nsteps ? 0.0 : da * pm_function_a(ah) / (ah * ah);
        const int rc = pm_step(p, x, y, z, px, py, pz, kick, drift, s == 0 ? variant : variant | PM_STEP_REUSE_BINS);
        if (rc) return rc;
    }
    return PM_OK;
}

// ------------------------------------------------- split-precision state
extern "C" int pm_split_load(pm_plan_t* p, const double* x, const double* y, const double* z, const double* vx,
                             const double* vy, const double* vz) {
    if (!p || !x || !y || !z || !vx || !vy || !vz) return fail(PM_ERR_ARG, "pm_split_load: null argument");
    int rc = ensure_sorted(p);  // the shell buffers of the deposit's column walk
    if (rc) return rc;
    SplitState& st = p->split;
    if (!p->split_buf) {
        const int ts = sorted_tile(p->ng);
        const size_t nt = (size_t)(p->ng / ts), ntiles = nt * nt * nt, n = (size_t)p->np;
        const size_t cap = split_capacity(n, ntiles);
        if (cap >= 0xFFFFFFF0ull) return fail(PM_ERR_ARG, "pm_split_load: too many particles for 32-bit slots");
        // two sets of slots (q, u, id: 3 words + 3 words + 4 B a slot: 28 B
        // with an FP32 mesh, 52 B with an FP64 mesh), the outboxes (the same
        // an entry), the hole lists (4 B an entry), the pending list (the
        // same a particle), boff, bfill, bnow, ocnt, hcnt, huse, ctl.  The
        // 8-byte words start on 8-byte boundaries: cap is a multiple of 32,
        // nout of 128, and the pending words come first in their part.
        const size_t w = p->prec == PM_FP64 ? 8 : 4, sb = 6 * w + 4;
        const size_t nout = ntiles * kOutbox;
        const size_t bytes = 2 * sb * cap + (sb + 4) * nout + sb * n + 4 * (6 * ntiles + 1) + 64;
        char* b = nullptr;
        rc = hip_ok(hipMalloc((void**)&b, bytes), "pm_split_load: hipMalloc");
        if (rc) return rc;
        p->split_buf = b;
        st.prec = p->prec;
        st.cap = (int64_t)cap;
        for (int k = 0; k < 2; ++k) {
            char* c = b + sb * cap * k;
            st.q[k] = c;
            st.u[k] = c + 3 * w * cap;
            st.id[k] = reinterpret_cast<uint32_t*>(c + 6 * w * cap);
        }
        char* c = b + 2 * sb * cap;
        st.oq = c;
        st.ou = c + 3 * w * nout;
        st.oid = reinterpret_cast<uint32_t*>(c + 6 * w * nout);
        c += sb * nout;
        st.hole = reinterpret_cast<uint32_t*>(c);
        c += 4 * nout;
        st.pq = c;
        st.pu = c + 3 * w * n;
        st.pid = reinterpret_cast<uint32_t*>(c + 6 * w * n);
        c += sb * n;
        st.boff = reinterpret_cast<uint32_t*>(c);
        st.bfill = st.boff + ntiles + 1;
        st.bnow = st.bfill + ntiles;
        st.ocnt = st.bnow + ntiles;
        st.hcnt = st.ocnt + ntiles;
        st.huse = st.hcnt + ntiles;
        st.ctl = st.huse + ntiles;
        // pending count, flags and the current set start at zero
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, p->stream, st.ctl, (size_t)16, 0u);
        rc = check_launch("pm_split_load: control words");
        if (rc) return rc;
    }
    rc = launch_state_load(x, y, z, vx, vy, vz, p->np, p->ng, st, p->err, p->stream);
    if (rc) return rc;
    p->split_loaded = true;
    return PM_OK;
}

extern "C" int pm_split_store(pm_plan_t* p, double* x, double* y, double* z, double* vx, double* vy, double* vz) {
    if (!p || !x || !y || !z || !vx || !vy || !vz) return fail(PM_ERR_ARG, "pm_split_store: null argument");
    if (!p->split_loaded) return fail(PM_ERR_ARG, "pm_split_store: no particles loaded (pm_split_load)");
    // the last step's leavers into their buckets first
    const int rc = launch_state_arrive(p->split, p->np, p->ng, p->stream);
    if (rc) return rc;
    return launch_state_store(p->split, p->np, p->ng, x, y, z, vx, vy, vz, p->stream);
}

extern "C" int pm_split_step(pm_plan_t* p, double kick, double drift) {
    if (!p) return fail(PM_ERR_ARG, "pm_split_step: null plan");
    if (!p->split_loaded) return fail(PM_ERR_ARG, "pm_split_step: no particles loaded (pm_split_load)");
    if (!std::isfinite(kick) || !std::isfinite(drift)) return fail(PM_ERR_ARG, "pm_split_step: bad kick or drift");
    mark(p, 0);
    const double mass = (double)p->ng * p->ng * p->ng / (double)p->np;
    // the last step's leavers into their buckets, the buckets rebuilt if a
    // load or a full bucket asked for it, then the deposit straight from them
    int rc = launch_state_arrive(p->split, p->np, p->ng, p->stream);
    if (!rc) rc = launch_state_compact(p->split, p->np, p->ng, p->sorted, p->stream);
    if (!rc) rc = launch_deposit_state(p->ng, mass, p->split, p->rho, p->sorted, p->stream);
    if (rc) return rc;
    mark(p, 1);
    rc = solve_impl(p);
    if (rc) return rc;
    rc = launch_state_interp(p->split, p->np, p->ng, p->phi, kick, drift, p->err, p->stream);
    if (rc) return rc;
    mark(p, 7);
    next_slot(p);
    return PM_OK;
}

extern "C" int pm_split_run_kdk(pm_plan_t* p, int nsteps, double dt) {
    if (!p || nsteps < 0) return fail(PM_ERR_ARG, "pm_split_run_kdk: bad arguments");
    if (nsteps == 0) return PM_OK;  // no step: no kick either
    for (int s = 0; s <= nsteps; ++s) {
        const double kick = (s == 0 || s == nsteps) ? 0.5 * dt : dt;
        const double drift = s == nsteps ? 0.0 : dt;
        const int rc = pm_split_step(p, kick, drift);
        if (rc) return rc;
    }
    return PM_OK;
}

extern "C" int pm_plan_enable_timing(pm_plan_t* p, int max_steps) {
    if (!p || max_steps < 0) return fail(PM_ERR_ARG, "pm_plan_enable_timing: bad argument");
    collect(p);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    p->ev.assign((size_t)max_steps * (PM_NPHASES + 1), nullptr);
    for (auto& e : p->ev)
        if (hip_ok(hipEventCreate(&e), "pm_plan_enable_timing: event")) {
            p->timing_slots = 0;
            return PM_ERR_HIP;
        }
    p->timing_slots = max_steps;
    p->timing_used = 0;
    return PM_OK;
}

extern "C" int pm_plan_phase_ms(pm_plan_t* p, double* ms, int64_t* counts) {
    if (!p || !ms) return fail(PM_ERR_ARG, "pm_plan_phase_ms: null argument");
    collect(p);
    for (int i = 0; i < PM_NPHASES; ++i) {
        ms[i] = p->phase_ms[i];
        if (counts) counts[i] = p->phase_n[i];
    }
    return PM_OK;
}

extern "C" int pm_plan_reset_timing(pm_plan_t* p) {
    if (!p) return fail(PM_ERR_ARG, "pm_plan_reset_timing: null plan");
    collect(p);
    for (int i = 0; i < PM_NPHASES; ++i) {
        p->phase_ms[i] = 0.0;
        p->phase_n[i] = 0;
    }
    return PM_OK;
}

// ------------------------------------------------------------------ adjoint
static int ensure_adjoint(pm_plan* p) {
    if (p->adjLam) return PM_OK;
    const size_t rs = p->prec == PM_FP32 ? sizeof(float) : sizeof(double);
    const size_t nc = (size_t)p->ng * p->ng * p->ng;
    int rc = ensure_sorted(p);
    for (int a = 0; a < 3 && !rc; ++a) rc = hip_ok(hipMalloc(&p->adjU[a], rs * nc), "adjoint: hipMalloc");
    // shells of the vec3 deposit (three meshes) or of the divergence deposit
    const size_t shell_elems = std::max(3 * sorted_shell_elems(p->ng), halo_shell_elems(p->ng));
    if (!rc) rc = hip_ok(hipMalloc(&p->adjShell, rs * shell_elems), "adjoint: hipMalloc");
    if (!rc) rc = hip_ok(hipMalloc(&p->adjLam, rs * nc), "adjoint: hipMalloc");
    const size_t nmax = 3 + 3 * (size_t)kMaxSlots;
    if (!rc) rc = hip_ok(hipMalloc((void**)&p->adjMax, nmax * sizeof(unsigned long long)), "adjoint: hipMalloc");
    if (!rc) rc = launch_fill(PM_FP64, p->adjMax, nmax, 0.0, p->stream);  // 0.0 has all bits zero
    // cotangent buckets for the staging of the reverse sweep (beside the staging records)
    if (!rc && p->sorted.stage)
        rc = hip_ok(hipMalloc(&p->adjUst, 3 * rs * p->sorted.cap_total), "adjoint: hipMalloc");
    if (!rc) rc = hip_ok(hipStreamCreateWithFlags(&p->adjStream, hipStreamNonBlocking), "adjoint: stream");
    if (!rc) rc = hip_ok(hipEventCreateWithFlags(&p->adjFork, hipEventDisableTiming), "adjoint: event");
    if (!rc) rc = hip_ok(hipEventCreateWithFlags(&p->adjJoin, hipEventDisableTiming), "adjoint: event");
    if (rc) {
        void* bufs[5] = {p->adjU[0], p->adjU[1], p->adjU[2], p->adjShell, p->adjLam};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (p->adjMax) (void)hipFree(p->adjMax);
        if (p->adjUst) (void)hipFree(p->adjUst);
        if (p->adjFork) (void)hipEventDestroy(p->adjFork);
        if (p->adjJoin) (void)hipEventDestroy(p->adjJoin);
        if (p->adjStream) (void)hipStreamDestroy(p->adjStream);
        p->adjU[0] = p->adjU[1] = p->adjU[2] = p->adjShell = p->adjLam = nullptr;
        p->adjMax = nullptr;
        p->adjUst = nullptr;
        p->adjStream = nullptr;
        p->adjFork = p->adjJoin = nullptr;
    }
    return rc;
}

// rho and phi at x with the given deposit variant, leaving the tile binning
// of x in p->sorted (the interpolation transpose needs it).
static int forward_field(pm_plan* p, const double* x, const double* y, const double* z, int variant) {
    const size_t nc = (size_t)p->ng * p->ng * p->ng;
    const double mass = (double)nc / (double)p->np;
    invalidate_reuse(p);
    int rc = launch_bin(x, y, z, p->np, p->ng, p->sorted, p->err, p->stream);
    if (rc) return rc;
    if (variant == PM_DEPOSIT_SORTED) {
        rc = launch_deposit_binned(p->prec, x, y, z, p->ng, mass, p->rho, p->sorted, p->stream);
    } else if (variant == PM_DEPOSIT_ATOMIC) {
        rc = launch_fill(p->prec, p->rho, nc, 0.0, p->stream);
        if (!rc) rc = launch_deposit_atomic(p->prec, x, y, z, p->np, p->ng, mass, p->rho, p->err, p->stream);
    } else {
        rc = fail(PM_ERR_ARG, "adjoint: unknown deposit variant");
    }
    if (rc) return rc;
    return solve_into(p, p->rho, p->phi, true);
}

// lambda = S sum_a D_a^T U_a with U_a the interpolation transpose of u at x
// (binning of x already in p->sorted): the divergence mu is deposited
// directly into adjoint mesh 0.
static int adjoint_field(pm_plan* p, const double* x, const double* y, const double* z, const double* ux,
                         const double* uy, const double* uz) {
    int rc = launch_deposit_div(p->prec, x, y, z, ux, uy, uz, p->np, p->ng, p->sorted, p->adjU[0], p->adjShell,
                                p->adjMax, p->err, p->stream);
    if (!rc) rc = solve_into(p, p->adjU[0], p->adjLam, false);
    return rc;
}

extern "C" int pm_adjoint_deposit(pm_plan_t* p, const double* x, const double* y, const double* z,
                                  const void* mesh_bar, double* xbar, double* ybar, double* zbar) {
    if (!p || !x || !y || !z || !mesh_bar || !xbar || !ybar || !zbar)
        return fail(PM_ERR_ARG, "pm_adjoint_deposit: null argument");
    const double mass = (double)p->ng * p->ng * p->ng / (double)p->np;
    return launch_adj_deposit(p->prec, x, y, z, p->np, p->ng, mesh_bar, mass, xbar, ybar, zbar, p->err,
                              p->stream);
}

extern "C" int pm_interp_transpose(pm_plan_t* p, const double* x, const double* y, const double* z,
                                   const double* ux, const double* uy, const double* uz) {
    if (!p || !x || !y || !z || !ux || !uy || !uz) return fail(PM_ERR_ARG, "pm_interp_transpose: null argument");
    int rc = ensure_adjoint(p);
    invalidate_reuse(p);
    if (!rc) rc = launch_bin(x, y, z, p->np, p->ng, p->sorted, p->err, p->stream);
    if (!rc)
        rc = launch_deposit_vec3(p->prec, x, y, z, ux, uy, uz, p->np, p->ng, p->sorted, p->adjU, p->adjShell,
                                 p->adjMax, p->err, p->stream);
    return rc;
}

extern "C" int pm_adjoint_solve(pm_plan_t* p) {
    if (!p) return fail(PM_ERR_ARG, "pm_adjoint_solve: null plan");
    int rc = ensure_adjoint(p);
    if (rc) return rc;
    const int slots = p->timing_slots;
    p->timing_slots = 0;
    rc = launch_div3(p->prec, p->adjU, p->ng, p->rho, p->stream);
    if (!rc) rc = solve_into(p, p->rho, p->adjLam, false);
    p->timing_slots = slots;
    return rc;
}

extern "C" void* pm_plan_adjoint_mesh(pm_plan_t* p, int which) {
    if (!p || which < 0 || which > 3) return nullptr;
    return which < 3 ? p->adjU[which] : p->adjLam;
}

extern "C" int pm_adjoint_force(pm_plan_t* p, const double* x, const double* y, const double* z, const double* ux,
                                const double* uy, const double* uz, double* xbar, double* ybar, double* zbar,
                                int variant) {
    if (!p || !x || !y || !z || !ux || !uy || !uz || !xbar || !ybar || !zbar)
        return fail(PM_ERR_ARG, "pm_adjoint_force: null argument");
    int rc = ensure_adjoint(p);
    if (rc) return rc;
    const int slots = p->timing_slots;
    p->timing_slots = 0;
    const double mass = (double)p->ng * p->ng * p->ng / (double)p->np;
    rc = forward_field(p, x, y, z, variant);
    if (!rc) rc = adjoint_field(p, x, y, z, ux, uy, uz);
    if (!rc)
        rc = launch_adj_gather(p->prec, 1, const_cast<double*>(x), const_cast<double*>(y), const_cast<double*>(z),
                               nullptr, nullptr, nullptr, const_cast<double*>(ux), const_cast<double*>(uy),
                               const_cast<double*>(uz), xbar, ybar, zbar, p->np, p->ng, p->phi, p->adjLam, mass, 0.0,
                               1.0, p->err, p->stream);  // mode 1 writes only xbar
    p->timing_slots = slots;
    return rc;
}

// One reverse step once its undrift is done (x and u = vbar + drift xbar in
// place): forward and adjoint fields at x, then the gather that undoes the
// kick and adds kick * J_F^T u to xbar; gather mode 2 also applies the next
// reverse step's undrift (next_drift) and leaves max |u| for its deposit.
//
// Staging (pm_run_adjoint with the sorted deposit): from_stage = the previous
// reverse step's gather staged x in the tile buckets (in-tile records and
// cotangents), so the forward deposit reads the records and the
// divergence deposit records and cotangents, with the full binning behind the
// overflow flag;
// stage_next = this step's gather stages the undrifted positions, into
// buckets sized from this step's counts.
static int adjoint_core(pm_plan* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                        double* xbar, double* ybar, double* zbar, double* ux, double* uy, double* uz, double kick,
                        int variant, int gmode, double next_drift, bool umax_ready, bool from_stage = false,
                        bool stage_next = false) {
    const size_t nc = (size_t)p->ng * p->ng * p->ng;
    const double mass = (double)nc / (double)p->np;
    if (variant != PM_DEPOSIT_SORTED && variant != PM_DEPOSIT_ATOMIC)
        return fail(PM_ERR_ARG, "adjoint: unknown deposit variant");
    // 1. the tile binning of x: from scratch, or (staged) only behind the
    //    overflow flag of the buckets
    int rc;
    if (from_stage) {
        rc = launch_bin(x, y, z, p->np, p->ng, p->sorted, p->err, p->stream, p->sorted.spill);
    } else {
        invalidate_reuse(p);
        rc = launch_bin(x, y, z, p->np, p->ng, p->sorted, p->err, p->stream);
    }
    // 2. the divergence deposit mu = sum_a D_a^T interp^T(u_a) on the second
    //    stream, beside the forward deposit and solve: it reads x, u and the
    //    tile lists, writes only adjoint mesh 0 and its shells
    if (!rc) rc = hip_ok(hipEventRecord(p->adjFork, p->stream), "adjoint: fork");
    if (!rc) rc = hip_ok(hipStreamWaitEvent(p->adjStream, p->adjFork, 0), "adjoint: fork");
    if (!rc)
        rc = launch_deposit_div(p->prec, x, y, z, ux, uy, uz, p->np, p->ng, p->sorted, p->adjU[0], p->adjShell,
                                p->adjMax, p->err, p->adjStream, umax_ready, from_stage ? p->adjUst : nullptr);
    if (!rc) rc = hip_ok(hipEventRecord(p->adjJoin, p->adjStream), "adjoint: join");
    // 3. rho and phi at x
    if (!rc) {
        if (from_stage)
            rc = launch_deposit_staged(p->prec, x, y, z, p->np, p->ng, mass, p->rho, p->sorted, p->err,
                                       p->stream, false, false);
        else if (variant == PM_DEPOSIT_SORTED)
            rc = launch_deposit_binned(p->prec, x, y, z, p->ng, mass, p->rho, p->sorted, p->stream);
        else if (!(rc = launch_fill(p->prec, p->rho, nc, 0.0, p->stream)))
            rc = launch_deposit_atomic(p->prec, x, y, z, p->np, p->ng, mass, p->rho, p->err, p->stream);
    }
    if (!rc) rc = solve_into(p, p->rho, p->phi, true);
    // 4. join; lambda = S mu
    if (!rc) rc = hip_ok(hipStreamWaitEvent(p->stream, p->adjJoin, 0), "adjoint: join");
    if (!rc) rc = solve_into(p, p->adjU[0], p->adjLam, false);
    // 5. the gather (undo the kick, VJP; mode 2: the next step's undrift,
    //    staged into buckets sized from this step's counts)
    AdjStage stg;
    if (!rc && stage_next) {
        rc = launch_cap_scan(p->ng, p->sorted, !from_stage, p->stream);
        stg.ts = sorted_tile(p->ng);
        stg.fill = p->sorted.fill;
        stg.coff = p->sorted.coff;
        stg.spill = p->sorted.spill;
        stg.rec = p->sorted.stage;
        stg.u = p->adjUst;
    }
    if (!rc)
        rc = launch_adj_gather(p->prec, gmode, x, y, z, vx, vy, vz, ux, uy, uz, xbar, ybar, zbar, p->np, p->ng, p->phi,
                               p->adjLam, mass, kick, kick, p->err, p->stream, next_drift, p->adjMax,
                               stage_next ? &stg : nullptr);
    return rc;
}

extern "C" int pm_adjoint_step(pm_plan_t* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                               double* xbar, double* ybar, double* zbar, double* vxbar, double* vybar,
                               double* vzbar, double kick, double drift, int variant) {
    if (!p || !x || !y || !z || !vx || !vy || !vz || !xbar || !ybar || !zbar || !vxbar || !vybar || !vzbar)
        return fail(PM_ERR_ARG, "pm_adjoint_step: null argument");
    int rc = ensure_adjoint(p);
    if (rc) return rc;
    const int slots = p->timing_slots;
    p->timing_slots = 0;
    invalidate_reuse(p);
    // undo the drift; the velocity cotangent becomes u = vbar + drift * xbar
    rc = launch_undrift(x, y, z, vx, vy, vz, xbar, ybar, zbar, vxbar, vybar, vzbar, p->np, p->ng, drift, p->stream);
    if (!rc) rc = adjoint_core(p, x, y, z, vx, vy, vz, xbar, ybar, zbar, vxbar, vybar, vzbar, kick, variant, 0, 0.0, false);
    p->timing_slots = slots;
    return rc;
}

extern "C" int pm_run_adjoint(pm_plan_t* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                              double* xbar, double* ybar, double* zbar, double* vxbar, double* vybar, double* vzbar,
                              int nsteps, double dt, int variant) {
    if (!p || nsteps < 0 || !x || !y || !z || !vx || !vy || !vz || !xbar || !ybar || !zbar || !vxbar || !vybar ||
        !vzbar)
        return fail(PM_ERR_ARG, "pm_run_adjoint: bad arguments");
    if (nsteps == 0) return PM_OK;  // the adjoint of zero steps (the identity)
    int rc = ensure_adjoint(p);
    if (rc) return rc;
    const int slots = p->timing_slots;
    p->timing_slots = 0;
    invalidate_reuse(p);
    // Reverse steps s = nsteps..0 of pm_run_kdk (kick dt/2 at both ends,
    // no drift after the last).  The last step has no drift to undo; every
    // other step's undrift is done by the gather of the step after it.
    // With the sorted deposit every gather but the last stages the positions
    // it undrifts for the next reverse step (no rebinning from scratch).
    const bool stage = variant == PM_DEPOSIT_SORTED && p->adjUst && p->sorted.stage;
    for (int s = nsteps; s >= 0 && !rc; --s) {
        const double kick = (s == 0 || s == nsteps) ? 0.5 * dt : dt;
        rc = adjoint_core(p, x, y, z, vx, vy, vz, xbar, ybar, zbar, vxbar, vybar, vzbar, kick, variant, s > 0 ? 2 : 0,
                          s > 0 ? dt : 0.0, s < nsteps, stage && s < nsteps, stage && s > 0);
    }
    p->timing_slots = slots;
    return rc;
}
