// FFT pass shapes and the process-wide tuning cache (host code).
//
// The reference's FFT constructor imports FFTW wisdom from
// wisdom_<n>_mpi_<size>.dat when the file exists and plans FFTW_WISDOM_ONLY,
// else plans FFTW_EXHAUSTIVE (measuring on its own arrays) and exports the
// wisdom when it is destroyed (fftw/fft.cpp:28-49,63-70).  Here the choice
// is the shape (lines per workgroup, threads per workgroup) of each of the
// five FFT passes: pm_plan_tune / pm_fft_tune time every variant of every
// pass on the handle's buffers and cache the fastest per (ng, precision);
// plans and FFT handles created later start from the cached shape; the
// cache is read from and written to a small text file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pm_internal.h"

namespace pm {

namespace {

struct WisdomEntry {
    int ng, prec;
    FftShape shape;
    double ms;
};

std::mutex g_mu;
std::vector<WisdomEntry> g_wisdom;

bool valid_side(int ng) { return ng >= 8 && ng <= 1024 && (ng & (ng - 1)) == 0; }

void store_locked(const WisdomEntry& e) {
    for (auto& w : g_wisdom)
        if (w.ng == e.ng && w.prec == e.prec) {
            w = e;
            return;
        }
    g_wisdom.push_back(e);
}

int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return PM_OK;
    return fail(PM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Events of one tuning run.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    int create() {
        int rc = hip_status(hipEventCreate(&a), "tune: hipEventCreate");
        if (!rc) rc = hip_status(hipEventCreate(&b), "tune: hipEventCreate");
        return rc;
    }
};

// One pass launch of a tuning run (the x slot of an FFT handle runs its
// forward and inverse passes, as pm_fft_forward / pm_fft_backward do).
struct PassCall {
    int pass, pass2;  // pass2 >= 0: launched right after pass
    const void* in;
    void* out;
    double norm;
    bool sub1;
};

// Mean milliseconds of `reps` launches of `call` in variant v of `slot`
// (after one untimed launch).
int time_call(int prec, int N, const PassCall& call, int slot, int v, hipStream_t st, const FftTables& tab,
              const GreenArgs& ga, int reps, EventPair& ev, double* ms) {
    FftShape sh;
    sh.v[slot] = v;
    auto launch = [&]() {
        int r = fft_pass(prec, N, call.pass, st, call.in, call.out, tab.twM, tab.twN, ga, call.norm, call.sub1, &sh);
        if (!r && call.pass2 >= 0)
            r = fft_pass(prec, N, call.pass2, st, call.in, call.out, tab.twM, tab.twN, ga, call.norm, call.sub1, &sh);
        return r;
    };
    if (launch()) return fail(PM_ERR_ARG, "tune: unsupported pass shape");
    int rc = hip_status(hipEventRecord(ev.a, st), "tune: hipEventRecord");
    for (int r = 0; !rc && r < reps; ++r) rc = launch() ? fail(PM_ERR_ARG, "tune: unsupported pass shape") : PM_OK;
    if (!rc) rc = hip_status(hipEventRecord(ev.b, st), "tune: hipEventRecord");
    if (!rc) rc = hip_status(hipEventSynchronize(ev.b), "tune: hipEventSynchronize");
    float t = 0.0f;
    if (!rc) rc = hip_status(hipEventElapsedTime(&t, ev.a, ev.b), "tune: hipEventElapsedTime");
    if (!rc) rc = check_launch("tune");
    *ms = (double)t / reps;
    return rc;
}

// Times all variants of the five slots (two interleaved rounds, the faster
// of each kept) and picks per slot the fastest; a variant replaces the
// default only when it is at least 2% faster.
int tune_calls(int prec, int N, const PassCall (&calls)[kFftSlots], hipStream_t st, const FftTables& tab,
               const GreenArgs& ga, int reps, FftShape* shape, double* total_ms) {
    EventPair ev;
    int rc = ev.create();
    if (rc) return rc;
    FftShape best;
    double total = 0.0;
    const int nv = N >= kTuneMinN ? kFftVariants : 1;
    for (int slot = 0; slot < kFftSlots; ++slot) {
        double t[kFftVariants];
        std::fill(t, t + kFftVariants, INFINITY);
        for (int round = 0; round < 2; ++round)
            for (int v = 0; v < nv; ++v) {
                double ms = 0.0;
                rc = time_call(prec, N, calls[slot], slot, v, st, tab, ga, reps, ev, &ms);
                if (rc) return rc;
                t[v] = std::min(t[v], ms);
            }
        int pick = 0;
        for (int v = 1; v < nv; ++v)
            if (t[v] < t[pick] && t[v] < 0.98 * t[0]) pick = v;
        best.v[slot] = pick;
        total += t[pick];
    }
    *shape = best;
    if (total_ms) *total_ms = total;
    return PM_OK;
}

std::vector<std::string> split_fields(const char* s) {
    std::vector<std::string> out;
    std::string cur;
    for (; *s; ++s) {
        if (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += *s;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// Whole-token parses: "1.0" is not an integer, "1x" not a number.
bool parse_int(const std::string& t, int* v) {
    char* end = nullptr;
    const long x = std::strtol(t.c_str(), &end, 10);
    if (end == t.c_str() || *end != '\0' || x < -(1L << 30) || x > (1L << 30)) return false;
    *v = (int)x;
    return true;
}

bool parse_double(const std::string& t, double* v) {
    char* end = nullptr;
    *v = std::strtod(t.c_str(), &end);
    return end != t.c_str() && *end == '\0';
}

bool valid_shape(const int* v) {
    for (int s = 0; s < kFftSlots; ++s)
        if (v[s] < 0 || v[s] >= kFftVariants) return false;
    return true;
}

}  // namespace

bool wisdom_find(int ng, int prec, FftShape* shape) {
    std::lock_guard<std::mutex> lock(g_mu);
    for (const auto& w : g_wisdom)
        if (w.ng == ng && w.prec == prec) {
            *shape = w.shape;
            return true;
        }
    return false;
}

void wisdom_store(int ng, int prec, const FftShape& shape, double solve_ms) {
    std::lock_guard<std::mutex> lock(g_mu);
    store_locked(WisdomEntry{ng, prec, shape, solve_ms});
}

}  // namespace pm

using namespace pm;

extern "C" int pm_plan_tune(pm_plan_t* p, int repeats, double* solve_ms) {
    if (!p || repeats < 1) return fail(PM_ERR_ARG, "pm_plan_tune: bad argument");
    const int N = p->ng;
    const size_t nc = (size_t)N * N * N;
    // finite data everywhere: a zero density through one default solve
    int rc = launch_fill(p->prec, p->rho, nc, 0.0, p->stream);
    const GreenArgs ga{p->green_mode, p->green_scale / ((double)N * N * N), p->tab.cos_tab};
    const PassCall calls[kFftSlots] = {
        {kPassZFwd, -1, p->rho, p->specA, 1.0, true},
        {kPassYFwd, -1, p->specA, p->specB, 1.0, false},
        {kPassXFused, -1, nullptr, p->specB, 1.0, false},
        {kPassYInv, -1, p->specB, p->specA, 1.0, false},
        {kPassZInv, -1, p->specA, p->phi, 1.0, false},
    };
    const FftShape dflt;
    for (int s = 0; !rc && s < kFftSlots; ++s)
        if (fft_pass(p->prec, N, calls[s].pass, p->stream, calls[s].in, calls[s].out, p->tab.twM, p->tab.twN, ga,
                     calls[s].norm, calls[s].sub1, &dflt))
            rc = fail(PM_ERR_ARG, "pm_plan_tune: unsupported mesh");
    FftShape shape;
    double ms = 0.0;
    if (!rc) rc = tune_calls(p->prec, N, calls, p->stream, p->tab, ga, repeats, &shape, &ms);
    if (rc) return rc;
    p->fft_shape = shape;
    wisdom_store(N, p->prec, shape, ms);
    if (solve_ms) *solve_ms = ms;
    return PM_OK;
}

extern "C" int pm_plan_set_fft_shape(pm_plan_t* p, const int variants[PM_FFT_SLOTS]) {
    if (!p || !variants || !valid_shape(variants)) return fail(PM_ERR_ARG, "pm_plan_set_fft_shape: bad argument");
    for (int s = 0; s < kFftSlots; ++s) p->fft_shape.v[s] = p->ng >= kTuneMinN ? variants[s] : 0;
    return PM_OK;
}

extern "C" int pm_plan_get_fft_shape(const pm_plan_t* p, int variants[PM_FFT_SLOTS]) {
    if (!p || !variants) return fail(PM_ERR_ARG, "pm_plan_get_fft_shape: null argument");
    for (int s = 0; s < kFftSlots; ++s) variants[s] = p->fft_shape.v[s];
    return PM_OK;
}

extern "C" int pm_fft_tune(pm_fft_t* f, void* slab, int repeats, double* fft_ms) {
    if (!f || !slab || repeats < 1) return fail(PM_ERR_ARG, "pm_fft_tune: bad argument");
    const int N = f->ng;
    const size_t nc = (size_t)N * N * N;
    int rc = launch_fill(f->prec, slab, nc, 0.0, f->stream);
    const GreenArgs ga{PM_GREEN_REF_KH, 1.0, f->tab.cos_tab};
    const double n3 = (double)N * N * N;
    const PassCall calls[kFftSlots] = {
        {kPassZFwd, -1, slab, f->specA, 1.0, false},
        {kPassYFwd, -1, f->specA, f->specB, 1.0, false},
        {kPassXFwd, kPassXInv, nullptr, f->specB, 1.0, false},
        {kPassYInv, -1, f->specB, f->specA, 1.0, false},
        {kPassZInv, -1, f->specA, slab, 1.0 / n3, false},
    };
    const FftShape dflt;
    for (int s = 0; !rc && s < kFftSlots; ++s)
        if (fft_pass(f->prec, N, calls[s].pass, f->stream, calls[s].in, calls[s].out, f->tab.twM, f->tab.twN, ga,
                     calls[s].norm, false, &dflt))
            rc = fail(PM_ERR_ARG, "pm_fft_tune: unsupported mesh");
    FftShape shape;
    double ms = 0.0;
    if (!rc) rc = tune_calls(f->prec, N, calls, f->stream, f->tab, ga, repeats, &shape, &ms);
    if (rc) return rc;
    f->shape = shape;
    wisdom_store(N, f->prec, shape, ms);
    if (fft_ms) *fft_ms = ms;
    return PM_OK;
}

extern "C" int pm_fft_get_shape(const pm_fft_t* f, int variants[PM_FFT_SLOTS]) {
    if (!f || !variants) return fail(PM_ERR_ARG, "pm_fft_get_shape: null argument");
    for (int s = 0; s < kFftSlots; ++s) variants[s] = f->shape.v[s];
    return PM_OK;
}

extern "C" int pm_wisdom_import(const char* path) {
    if (!path) return fail(PM_ERR_ARG, "pm_wisdom_import: null path");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(PM_ERR_ARG, std::string("pm_wisdom_import: cannot open ") + path);
    std::vector<WisdomEntry> got;
    char line[512];
    int lineno = 0;
    bool header = false, bad = false;
    while (!bad && std::fgets(line, sizeof line, fp)) {
        ++lineno;
        if (!std::strchr(line, '\n') && !std::feof(fp)) {
            bad = true;  // longer than any valid line
            break;
        }
        char* s = line;
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '\n' || *s == '\r' || *s == '\0' || *s == '#') continue;
        if (!header) {
            int version = 0;
            char tag[16] = {0};
            header = std::sscanf(s, "%15s %d", tag, &version) == 2 && !std::strcmp(tag, "pm-wisdom") && version == 1;
            if (!header) bad = true;
            continue;
        }
        WisdomEntry e{};
        const std::vector<std::string> tok = split_fields(s);
        if (tok.size() != 8 || !parse_int(tok[0], &e.ng) || !parse_double(tok[7], &e.ms)) {
            bad = true;
            break;
        }
        for (int k = 0; k < kFftSlots; ++k)
            if (!parse_int(tok[2 + k], &e.shape.v[k])) bad = true;
        if (tok[1] == "fp32")
            e.prec = PM_FP32;
        else if (tok[1] == "fp64")
            e.prec = PM_FP64;
        else
            bad = true;
        if (!valid_side(e.ng) || !valid_shape(e.shape.v) || !std::isfinite(e.ms) || e.ms < 0.0) bad = true;
        if (!bad && e.ng < kTuneMinN)
            for (int& v : e.shape.v) v = 0;
        if (!bad) got.push_back(e);
    }
    std::fclose(fp);
    if (bad || !header)
        return fail(PM_ERR_ARG, std::string("pm_wisdom_import: malformed file ") + path + " at line " +
                                    std::to_string(lineno));
    std::lock_guard<std::mutex> lock(g_mu);
    for (const auto& e : got) store_locked(e);
    return PM_OK;
}

extern "C" int pm_wisdom_export(const char* path) {
    if (!path) return fail(PM_ERR_ARG, "pm_wisdom_export: null path");
    std::vector<WisdomEntry> all;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        all = g_wisdom;
    }
    std::sort(all.begin(), all.end(), [](const WisdomEntry& a, const WisdomEntry& b) {
        return a.ng != b.ng ? a.ng < b.ng : a.prec < b.prec;
    });
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(PM_ERR_ARG, std::string("pm_wisdom_export: cannot open ") + path);
    bool ok = std::fprintf(fp, "pm-wisdom 1\n# ng precision zfwd yfwd x yinv zinv solve_ms\n") > 0;
    for (const auto& e : all)
        ok = ok && std::fprintf(fp, "%d %s %d %d %d %d %d %.6f\n", e.ng, e.prec == PM_FP32 ? "fp32" : "fp64",
                                e.shape.v[0], e.shape.v[1], e.shape.v[2], e.shape.v[3], e.shape.v[4], e.ms) > 0;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? PM_OK : fail(PM_ERR_ARG, std::string("pm_wisdom_export: write failed for ") + path);
}

extern "C" int pm_wisdom_lookup(int ng, int precision, int variants[PM_FFT_SLOTS], int* found) {
    if (!found) return fail(PM_ERR_ARG, "pm_wisdom_lookup: null argument");
    FftShape s;
    *found = wisdom_find(ng, precision, &s) ? 1 : 0;
    if (*found && variants)
        for (int i = 0; i < kFftSlots; ++i) variants[i] = s.v[i];
    return PM_OK;
}

extern "C" void pm_wisdom_forget(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_wisdom.clear();
}

// ------------------------------------------------------------ copy ceiling
namespace {
typedef float v4f __attribute__((ext_vector_type(4)));
// one 16-byte element a thread (many short workgroups: measured on MI355X,
// 6.2 TB/s read + write against 4.8-5.3 TB/s for grid-stride loops of 16 to
// 64 elements a thread)
__global__ __launch_bounds__(256) void k_copy16(const v4f* __restrict__ a, v4f* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
}  // namespace

extern "C" int pm_copy_rate(size_t bytes, int repeats, double* gbs) {
    using namespace pm;
    if (!gbs || repeats < 1 || bytes < ((size_t)16 << 20) || bytes % 16)
        return fail(PM_ERR_ARG, "pm_copy_rate: bad arguments");
    const size_t n = bytes / 16;
    if (n / 256 + 1 > 0x7FFFFFFFull) return fail(PM_ERR_ARG, "pm_copy_rate: too many bytes");
    v4f *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = hip_status(hipMalloc((void**)&a, bytes), "pm_copy_rate: hipMalloc");
    if (!rc) rc = hip_status(hipMalloc((void**)&b, bytes), "pm_copy_rate: hipMalloc");
    if (!rc) rc = hip_status(hipStreamCreate(&st), "pm_copy_rate: stream");
    if (!rc) rc = hip_status(hipEventCreate(&e0), "pm_copy_rate: event");
    if (!rc) rc = hip_status(hipEventCreate(&e1), "pm_copy_rate: event");
    if (!rc) {
        const dim3 grid((unsigned)((n + 255) / 256));
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3(4096), dim3(256), 0, st, reinterpret_cast<uint32_t*>(a), n * 4, 1u);
        for (int w = 0; w < 2; ++w) hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, st, a, b, n);
        rc = hip_status(hipEventRecord(e0, st), "pm_copy_rate: record");
        for (int r = 0; !rc && r < repeats; ++r) hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, st, a, b, n);
        if (!rc) rc = hip_status(hipEventRecord(e1, st), "pm_copy_rate: record");
        if (!rc) rc = hip_status(hipEventSynchronize(e1), "pm_copy_rate: synchronize");
        float ms = 0.f;
        if (!rc) rc = hip_status(hipEventElapsedTime(&ms, e0, e1), "pm_copy_rate: elapsed");
        if (!rc) rc = check_launch("pm_copy_rate");
        if (!rc) *gbs = 2.0 * (double)bytes * repeats / ((double)ms * 1e6);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
