// Internal declarations shared by the PM HIP sources.
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "pm_api.h"
#include "pm_common.h"

namespace pm {

// Lattice rows per band of the particle kernels' block order (band_remap);
// 0 sweeps the particles in plain order.
#ifndef PM_BAND_ROWS
#define PM_BAND_ROWS 32
#endif
// The band map for np particles in nb blocks of per_block: a band map when
// np = L^3 and the XCD ranges, lattice planes and bands fall on whole blocks,
// else the plain order.
inline BandMap band_map(int64_t np, unsigned nb, int per_block) {
    const BandMap plain{0u, 0u};
    if (PM_BAND_ROWS <= 0 || np <= 0 || nb % kXcds) return plain;
    const int64_t l = std::llround(std::cbrt((double)np));
    if (l * l * l != np || PM_BAND_ROWS >= l) return plain;
    const int64_t plane = l * l, band = (int64_t)PM_BAND_ROWS * l;
    if (plane % per_block || band % per_block) return plain;
    const int64_t bpp = plane / per_block, bb = band / per_block;
    if (bpp % bb || (nb / kXcds) % bpp) return plain;
    return BandMap{(uint32_t)bpp, (uint32_t)bb};
}

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

// Shape choice of the five tunable FFT passes (slots zfwd, yfwd, x, yinv,
// zinv): a variant index each, 0 = the default shape (pm_fft.hip).
constexpr int kFftSlots = PM_FFT_SLOTS;
constexpr int kFftVariants = PM_FFT_VARIANTS;
constexpr int kTuneMinN = 64;  // smaller meshes always take variant 0
struct FftShape {
    int v[kFftSlots] = {0, 0, 0, 0, 0};
};

// Slot of a pass (-1 for the standalone Green's multiply).
int fft_slot(int pass);

// Launch one FFT pass (see pm_fft.hip) in the shape that 'shape' selects for
// its slot (nullptr: the default).  Returns -1 for an unsupported N or variant.
int fft_pass(int prec, int N, int pass, hipStream_t st, const void* in, void* out, const void* twM,
             const void* twN, const GreenArgs& ga, double norm, bool sub1, const FftShape* shape = nullptr);

// Process-wide tuning cache (pm_wisdom.hip): the shape chosen for a mesh
// side and precision by pm_plan_tune / pm_fft_tune or read by
// pm_wisdom_import.
bool wisdom_find(int ng, int prec, FftShape* shape);
void wisdom_store(int ng, int prec, const FftShape& shape, double solve_ms);

// Grid-stride fill (used instead of hipMemsetAsync so that the step can be
// captured in a graph).
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, size_t n, T v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// A staged particle: its offsets inside its tile (one record = one vector
// store).  FP32 mesh: u32 at 2^-28 cells (cell = q >> 28).  FP64 mesh: 42
// bits each at 2^-38 cells (cell = q >> 38) packed in 16 bytes: x in bits
// 0..41 of lo, y in bits 42..63 of lo and 0..19 of hi, z in bits 20..61 of hi.
struct StageQ {
    uint32_t x, y, z;
};
struct StageD {
    uint64_t lo, hi;
};
constexpr double kQScaleD = 274877906944.0;  // 2^38
constexpr uint64_t kQMaskD = (1ull << 38) - 1;
__host__ __device__ inline StageD pack_stage_d(uint64_t qx, uint64_t qy, uint64_t qz) {
    return StageD{qx | (qy << 42), (qy >> 22) | (qz << 20)};
}
__host__ __device__ inline void unpack_stage_d(const StageD& r, uint64_t& qx, uint64_t& qy, uint64_t& qz) {
    constexpr uint64_t m42 = (1ull << 42) - 1;
    qx = r.lo & m42;
    qy = ((r.lo >> 42) | (r.hi << 22)) & m42;
    qz = (r.hi >> 20) & m42;
}

// Work space of the sorted deposit (pm_deposit_sorted.hip).
struct SortedScratch {
    uint32_t* cnt = nullptr;   // ntiles
    uint32_t* off = nullptr;   // ntiles + 1
    uint32_t* key = nullptr;   // np
    uint32_t* slot = nullptr;  // np
    uint32_t* idx = nullptr;   // np
    void* shell = nullptr;     // sorted_shell_elems reals: xy shells, then z seams
    // Staging in tile order, filled by the interpolation for the next
    // step's deposit: bucket t holds fill[t] in-tile offsets from coff[t]
    // (capacity coff[t+1] - coff[t]); spill != 0 if a bucket overflowed.
    uint32_t* fill = nullptr;   // ntiles
    uint32_t* coff = nullptr;   // ntiles + 1
    uint32_t* spill = nullptr;  // 1
    void* stage = nullptr;      // cap_total records of in-tile offsets (x, y, z):
                                // StageQ (FP32 mesh) or StageD (FP64 mesh)
    size_t cap_total = 0;
};
int sorted_tile(int ng);
size_t sorted_shell_elems(int ng);
// Halo planes of the divergence deposit's column walk (its stencil reaches
// two nodes above a cell's upper corner) and the reals of its shells.
constexpr int kDivHalo = 3;
size_t halo_shell_elems(int ng);
int launch_fold_halo(int prec, int ts, int ng, const void* shell, void* mesh, hipStream_t st);
int launch_deposit_sorted(int prec, const double* px, const double* py, const double* pz, int64_t np, int ng,
                          double mass, void* rho, SortedScratch& s, unsigned long long* err, hipStream_t st);
// binning only (count, scan, scatter) and the per-tile deposit of a binning
// (a non-null gate makes every kernel a no-op unless *gate != 0)
int launch_bin(const double* px, const double* py, const double* pz, int64_t np, int ng, SortedScratch& s,
               unsigned long long* err, hipStream_t st, const uint32_t* gate = nullptr);
int launch_deposit_binned(int prec, const double* px, const double* py, const double* pz, int ng, double mass,
                          void* rho, SortedScratch& s, hipStream_t st, const uint32_t* gate = nullptr,
                          bool fold = true);
// staging capacities for the next step (from the binning offsets, or from
// the fill counts of a staged step)
int launch_cap_scan(int ng, SortedScratch& s, bool from_off, hipStream_t st);
// the same from the fill counters, the binning offsets or the tile counts
// (s.cnt, cleared on the way), behind an optional gate
enum CapFrom { kCapFromFill = 0, kCapFromOff = 1, kCapFromCnt = 2 };
int launch_cap_scan_from(int ng, SortedScratch& s, int from, const uint32_t* gate, hipStream_t st);
// deposit from the staging arrays (binned fallback from px, py, pz if a
// bucket overflowed; px == nullptr: no fallback), then the next capacities
// (rescan = false leaves the buckets and their fill counts for a second
// reader, the reverse sweep's divergence deposit; launch_cap_scan follows it.
// bin = false: the fallback binning was launched already)
int launch_deposit_staged(int prec, const double* px, const double* py, const double* pz, int64_t np, int ng,
                          double mass, void* rho, SortedScratch& s, unsigned long long* err, hipStream_t st,
                          bool rescan = true, bool bin = true);
int launch_fold_shells(int prec, int ts, int ng, const void* shell, void* rho, hipStream_t st);

// Adjoint kernels (pm_adjoint.hip).
int launch_adj_deposit(int prec, const double* px, const double* py, const double* pz, int64_t np, int ng,
                       const void* mbar, double scale, double* bx, double* by, double* bz, unsigned long long* err,
                       hipStream_t st);
// The interpolation transpose U_a = interp^T(u_a) over the tile binning of
// x in s (pm_interp_transpose), with its own max |u_a| and shell folds.  A
// non-finite cotangent component is counted in err (the range-error counter).
int launch_deposit_vec3(int prec, const double* px, const double* py, const double* pz, const double* ux,
                        const double* uy, const double* uz, int64_t np, int ng, const SortedScratch& s,
                        void* const U[3], void* shell3, unsigned long long* umax, unsigned long long* err,
                        hipStream_t st);
int launch_div3(int prec, void* const U[3], int ng, void* mu, hipStream_t st);
// mu = sum_a D_a^T interp^T(u_a) in one deposit (the mesh that
// launch_deposit_vec3 + launch_div3 form), folded; shell holds
// halo_shell_elems(ng) reals.  umax_ready: umax already holds max |u_a|,
// left by a mode-2 gather; bucket_u: the particles come from the buckets
// that a staging mode-2 gather filled (coff / fill of s), or, if a
// bucket overflowed (s.spill), from the binning that the staged deposit's
// fallback made.  A non-finite cotangent component is counted in err.
// bucket_u: the cotangents staged beside the records (Real3 of the mesh
// precision); the binned pass reads u by particle index and rounds it to the
// mesh precision the same way, so both passes deposit the same values.
int launch_deposit_div(int prec, const double* px, const double* py, const double* pz, const double* ux,
                       const double* uy, const double* uz, int64_t np, int ng, const SortedScratch& s, void* mu,
                       void* shell, unsigned long long* umax, unsigned long long* err, hipStream_t st,
                       bool umax_ready = false, const void* bucket_u = nullptr);
// mode 1: xbar += VJP; 0: also undo the kick; 2: also the next reverse
// step's undrift (next_drift) with max |u| into umax[0..3) (umax holds
// 3 + 3 * kMaxSlots entries, the slots zero between calls)
constexpr int kMaxSlots = 256;
// Staging of a mode-2 gather: the undrifted positions go to the tile
// buckets of s (coff, fill, spill) as in-tile records (s.stage) and their
// cotangents u, rounded to the mesh precision, at the same slots (u), for
// the next reverse step's divergence deposit.
template <typename Rl> struct Real3 {
    Rl x, y, z;
};
struct AdjStage {
    int ts = 0;
    uint32_t* fill = nullptr;
    const uint32_t* coff = nullptr;
    uint32_t* spill = nullptr;
    void* rec = nullptr;
    void* u = nullptr;  // the next step's cotangent u, 3 mesh reals per slot (Real3)
};
int launch_adj_gather(int prec, int mode, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                      double* ux, double* uy, double* uz, double* bx, double* by, double* bz, int64_t np, int ng,
                      const void* phi, const void* lam, double mass, double kick, double vjp_scale,
                      unsigned long long* err, hipStream_t st, double next_drift = 0.0,
                      unsigned long long* umax = nullptr, const AdjStage* stage = nullptr);
int launch_undrift(double* px, double* py, double* pz, const double* vx, const double* vy, const double* vz,
                   const double* bx, const double* by, const double* bz, double* ubx, double* uby, double* ubz,
                   int64_t np, int ng, double drift, hipStream_t st);

// Split-precision particle state (pm_split.hip), resident in tile order:
// positions as fixed point over the box (q = x * 2^(bits - log2 ng): cell and
// fraction; u32 for an FP32 mesh, u64 for an FP64 mesh) and velocities in the
// mesh precision, a slot's three values side by side
// (sel() of pm_tiles.h), with the particle index of every slot (id, kNoId
// for a free slot or a hole).  Two sets of slots: ctl[3] names the current
// one, a compaction rebuilds the buckets into the other and switches.  Tile
// t owns slots [boff[t], boff[t+1]) and uses bfill[t] of them; bnow[t]: the
// slots the current step's interpolation takes.  pq, pu, pid: the pending
// list (np entries: the particles of a load, and the arrivals that found
// their bucket full); ctl[0]: pending count, ctl[1]: compaction flag.
struct SplitState {
    int prec = PM_FP32;  // the mesh precision: u32 positions and FP32 velocities, or u64 and FP64
    int64_t cap = 0;
    void* q[2] = {};
    void* u[2] = {};
    uint32_t* id[2] = {};
    void* pq = nullptr;
    void* pu = nullptr;
    uint32_t* pid = nullptr;
    uint32_t* boff = nullptr;
    uint32_t* bfill = nullptr;
    uint32_t* bnow = nullptr;
    uint32_t* ctl = nullptr;
    // per tile, the particles that left it in the last interpolation:
    // kOutbox entries of position and velocity triples and the particle
    // index, and their count
    void* oq = nullptr;
    void* ou = nullptr;
    uint32_t* oid = nullptr;
    uint32_t* ocnt = nullptr;
    // per tile, up to kOutbox holes of its range (slots whose particle left)
    // that the last interpolation listed, their count (hcnt, right after
    // ocnt) and how many arrivals took one (huse, right after hcnt)
    uint32_t* hole = nullptr;
    uint32_t* hcnt = nullptr;
    uint32_t* huse = nullptr;
};
constexpr int kOutbox = 128;
// slots per tile bucket of the split state: count + count/2 + 64, rounded up
// to 32 slots (a bucket starts on a 128-byte line of every slot array, so a
// wave's 64 slots are whole lines); the total is a multiple of 32 as well, so
// that every array of both sets starts on a line.  The slack takes the
// arrivals that find no hole: tiles that accrete (collapsing regions of a
// clustered run) grow by several percent a step, and one full bucket makes
// the next step rebuild every bucket.
inline size_t split_capacity(size_t n, size_t ntiles) { return (n + n / 2 + 96 * ntiles + 31) / 32 * 32; }
int launch_state_load(const double* x, const double* y, const double* z, const double* vx, const double* vy,
                      const double* vz, int64_t np, int ng, SplitState& st, unsigned long long* err, hipStream_t s);
int launch_state_store(const SplitState& st, int64_t np, int ng, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, hipStream_t s);
// rebuild the buckets from the occupied slots and the pending list, if the
// compaction flag is set (clears it)
int launch_state_compact(SplitState& st, int64_t np, int ng, SortedScratch& s, hipStream_t hs);
// the sorted deposit of the buckets' particles (+ fold), leaving bnow
int launch_deposit_state(int ng, double mass, const SplitState& st, void* rho, SortedScratch& s, hipStream_t hs);
// interpolation + kick + drift tile by tile (FP32 mesh)
// the leavers that the last interpolation put in the outboxes, into their
// new buckets (a no-op if the outboxes are empty)
int launch_state_arrive(SplitState& st, int64_t np, int ng, hipStream_t hs);
int launch_state_interp(SplitState& st, int64_t np, int ng, const void* phi, double kick, double drift,
                        unsigned long long* err, hipStream_t hs);
// the capacity scan on explicit arrays, from the tile counts cnt (cleared)
int launch_cap_scan_raw(int n, uint32_t* fill, uint32_t* cnt, uint32_t* coff, uint32_t* flag, const uint32_t* gate,
                        hipStream_t st);

// Particle kernels (pm_particles.hip).
int launch_fill(int prec, void* p, size_t n, double v, hipStream_t st);
int launch_deposit_atomic(int prec, const double* px, const double* py, const double* pz,
                          int64_t np, int ng, double mass, void* rho, unsigned long long* err,
                          hipStream_t st);
// mode 0: kick + drift, 1: forces only (into vx, vy, vz), 2: kick + drift and
// the drifted positions staged in tile order into s (fill, coff, stage)
int launch_interp(int prec, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                  int64_t np, int ng, const void* phi, double kick, double drift, int mode,
                  unsigned long long* err, hipStream_t st, SortedScratch* s = nullptr);

// Error reporting.
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);
// The current device's range-error counter (a device variable; nullptr only
// if the library's code object cannot be found on the device).
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
    unsigned long long* err = nullptr;  // the device's range-error counter (range_counter())
    void* rho = nullptr;     // ng^3 reals
    void* phi = nullptr;     // ng^3 reals
    void* specA = nullptr;   // (ng/2+1) ng^2 complex
    void* specB = nullptr;
    pm::FftTables tab;
    pm::FftShape fft_shape;  // the FFT pass shapes (default, cached or tuned)
    pm::SortedScratch sorted;  // allocated on the first sorted deposit
    // the interpolation of the last pm_step staged positions staged_x in tile
    // order (sorted.stage) for the next deposit (PM_STEP_REUSE_BINS)
    bool staged_valid = false;
    const double* staged_x = nullptr;
    pm::SplitState split;       // allocated by the first pm_split_load
    void* split_buf = nullptr;  // its one allocation
    bool split_loaded = false;  // split holds the particles of a pm_split_load
    void* adjU[3] = {};        // adjoint meshes U_a (interpolation transpose); the
                               // fused paths put the divergence in adjU[0]
    void* adjLam = nullptr;    // lambda = S D^T U
    void* adjShell = nullptr;  // shells of the vec3 deposit (3) or the divergence deposit
    unsigned long long* adjMax = nullptr;  // max |u_a| (fixed-point scales of the adjoint deposits) + slots
    void* adjUst = nullptr;  // staged cotangents of the reverse sweep (cap_total Real3), beside sorted.stage
    // Second stream of an adjoint step (default priority): the divergence
    // deposit runs there beside the forward deposit and solve; fork / join events.
    hipStream_t adjStream = nullptr;
    hipEvent_t adjFork = nullptr, adjJoin = nullptr;
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
    pm::FftShape shape;
};
