// The sorted CIC deposit for callers that cannot allocate device memory.
//
// pm_deposit(plan, ..., PM_DEPOSIT_SORTED) bins the particles by tile in
// buffers that the plan allocates on its first sorted deposit.  This entry
// point runs the same kernels (count, scan, scatter, per-tile LDS deposit,
// shell fold) with every work buffer carved out of one fixed device array,
// `scratch`, that the library's code object holds on each device from the
// moment it is loaded until the process ends.  The array's device address is
// looked up once per device (hipGetSymbolAddress, with that device current);
// after that only the looked-up address is used.  No hipMalloc, hipFree or
// allocating library routine is called on this path.
#include <algorithm>
#include <mutex>

#include "pm_internal.h"

namespace pm {

constexpr size_t kScratchBytes = (size_t)PM_NOALLOC_SCRATCH_MB << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static void* g_scratch[64] = {};
static std::mutex g_scratch_mu;

// Device address of this device's scratch array (looked up on the first call
// per device).
static int scratch_address(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return fail(PM_ERR_HIP, "pm_deposit_sorted_noalloc: no current device");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PM_ERR_HIP, std::string("pm_deposit_sorted_noalloc: scratch address: ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return PM_OK;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Layout of the binning inside the scratch array: per-tile counts and
// offsets, per-particle tile key, slot and tile-ordered index, then the
// deposit's xy shells and z seams; every part starts on 256 bytes.
struct NoAllocLayout {
    size_t cnt, off, key, slot, idx, shell, total;
};

static NoAllocLayout layout_for(int64_t np, int ng, int prec) {
    const int ts = sorted_tile(ng);
    const size_t nt = (size_t)(ng / ts);
    const size_t ntiles = nt * nt * nt;
    const size_t rs = prec == PM_FP32 ? sizeof(float) : sizeof(double);
    NoAllocLayout l;
    l.cnt = 0;
    l.off = l.cnt + round256(4 * ntiles);
    l.key = l.off + round256(4 * (ntiles + 1));
    l.slot = l.key + round256(4 * (size_t)np);
    l.idx = l.slot + round256(4 * (size_t)np);
    l.shell = l.idx + round256(4 * (size_t)np);
    l.total = l.shell + round256(rs * sorted_shell_elems(ng));
    return l;
}

}  // namespace pm

using namespace pm;

extern "C" int pm_deposit_sorted_noalloc(const double* posx, const double* posy, const double* posz, int64_t np,
                                         int ng, int precision, void* rho, void* stream) {
    if (!posx || !posy || !posz || !rho) return fail(PM_ERR_ARG, "pm_deposit_sorted_noalloc: null argument");
    if (!valid_ng(ng)) return fail(PM_ERR_ARG, "pm_deposit_sorted_noalloc: ng must be a power of two in [8, 1024]");
    if (precision != PM_FP32 && precision != PM_FP64)
        return fail(PM_ERR_ARG, "pm_deposit_sorted_noalloc: bad precision");
    if (np <= 0) return fail(PM_ERR_ARG, "pm_deposit_sorted_noalloc: np must be positive");
    const NoAllocLayout l = layout_for(np, ng, precision);
    if (l.total > kScratchBytes)
        return fail(PM_ERR_ARG, "pm_deposit_sorted_noalloc: np = " + std::to_string(np) + " on " + std::to_string(ng) +
                                    "^3 needs " + std::to_string(l.total) + " bytes of scratch, more than the " +
                                    std::to_string(kScratchBytes) + " the library holds");
    void* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    unsigned long long* err = range_counter();
    if (!err) return fail(PM_ERR_HIP, "pm_deposit_sorted_noalloc: range-error counter not found");
    hipStream_t st = (hipStream_t)stream;
    unsigned char* b = (unsigned char*)base;
    SortedScratch s;
    s.cnt = (uint32_t*)(b + l.cnt);
    s.off = (uint32_t*)(b + l.off);
    s.key = (uint32_t*)(b + l.key);
    s.slot = (uint32_t*)(b + l.slot);
    s.idx = (uint32_t*)(b + l.idx);
    s.shell = b + l.shell;
    // The binning counts must start at zero; the array keeps whatever the
    // last call on this device left, so clear them (a fill kernel, not a
    // memset node, in case the caller captures the call in a graph).
    const int ts = sorted_tile(ng);
    const size_t nt = (size_t)(ng / ts);
    const size_t ntiles = nt * nt * nt;
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3((unsigned)std::min<size_t>((ntiles + 255) / 256, 8192)), dim3(256), 0,
                       st, s.cnt, ntiles, 0u);
    rc = check_launch("pm_deposit_sorted_noalloc: clear counts");
    if (rc) return rc;
    const double mass = (double)ng * ng * ng / (double)np;
    return launch_deposit_sorted(precision, posx, posy, posz, np, ng, mass, rho, s, err, st);
}
