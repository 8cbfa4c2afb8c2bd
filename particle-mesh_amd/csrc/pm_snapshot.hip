// Particle snapshots: a small versioned binary file of the particle state,
// written from and read into device arrays (host code only).  The
// reference keeps its particles in memory and has no snapshot format.
//
// Layout (little endian):
//   bytes  0..7   magic "PMSNAP01"
//          8..11  u32 version (1)
//         12..15  u32 header size (64)
//         16..23  i64 np
//         24..27  i32 ng (mesh side; positions in cell units, [0, ng))
//         28..31  i32 step
//         32..39  f64 a (expansion factor or time)
//         40..63  reserved (0)
//   then six float64 arrays of np values: x, y, z, vx, vy, vz.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pm_internal.h"

namespace pm {

constexpr char kSnapMagic[8] = {'P', 'M', 'S', 'N', 'A', 'P', '0', '1'};
constexpr uint32_t kSnapVersion = 1;
constexpr uint32_t kSnapHeader = 64;
constexpr size_t kSnapChunk = (size_t)8 << 20;  // doubles per staged copy (64 MiB)

struct SnapHeader {
    char magic[8];
    uint32_t version, header;
    int64_t np;
    int32_t ng, step;
    double a;
    uint8_t reserved[24];
};
static_assert(sizeof(SnapHeader) == kSnapHeader, "snapshot header is 64 bytes");

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

struct Pinned {
    double* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

static int read_header(FILE* f, SnapHeader& h, const char* path) {
    if (fread(&h, sizeof(h), 1, f) != 1) return fail(PM_ERR_ARG, std::string("snapshot: short header in ") + path);
    if (memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != kSnapVersion || h.header != kSnapHeader || h.np < 0)
        return fail(PM_ERR_ARG, std::string("snapshot: not a version-1 PM snapshot: ") + path);
    return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" int pm_snapshot_write(const char* path, int64_t np, int ng, int step, double a, const double* x,
                                 const double* y, const double* z, const double* vx, const double* vy,
                                 const double* vz, void* stream) {
    if (!path || np < 0 || (np > 0 && (!x || !y || !z || !vx || !vy || !vz)))
        return fail(PM_ERR_ARG, "pm_snapshot_write: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    File file;
    file.f = fopen(path, "wb");
    if (!file.f) return fail(PM_ERR_ARG, std::string("pm_snapshot_write: cannot open ") + path);
    SnapHeader h{};
    memcpy(h.magic, kSnapMagic, 8);
    h.version = kSnapVersion;
    h.header = kSnapHeader;
    h.np = np;
    h.ng = ng;
    h.step = step;
    h.a = a;
    if (fwrite(&h, sizeof(h), 1, file.f) != 1) return fail(PM_ERR_ARG, "pm_snapshot_write: write failed");
    Pinned buf;
    const size_t chunk = std::min<size_t>(kSnapChunk, (size_t)std::max<int64_t>(np, 1));
    if (hipHostMalloc((void**)&buf.p, chunk * sizeof(double)) != hipSuccess)
        return fail(PM_ERR_NOMEM, "pm_snapshot_write: hipHostMalloc");
    const double* arr[6] = {x, y, z, vx, vy, vz};
    for (const double* src : arr) {
        for (size_t o = 0; o < (size_t)np; o += chunk) {
            const size_t n = std::min(chunk, (size_t)np - o);
            if (hipMemcpyAsync(buf.p, src + o, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail(PM_ERR_HIP, "pm_snapshot_write: device copy failed");
            if (fwrite(buf.p, sizeof(double), n, file.f) != n) return fail(PM_ERR_ARG, "pm_snapshot_write: write failed");
        }
    }
    if (fclose(file.f) != 0) {
        file.f = nullptr;
        return fail(PM_ERR_ARG, "pm_snapshot_write: close failed");
    }
    file.f = nullptr;
    return PM_OK;
}

extern "C" int pm_snapshot_info(const char* path, int64_t* np, int* ng, int* step, double* a) {
    if (!path) return fail(PM_ERR_ARG, "pm_snapshot_info: null path");
    File file;
    file.f = fopen(path, "rb");
    if (!file.f) return fail(PM_ERR_ARG, std::string("pm_snapshot_info: cannot open ") + path);
    SnapHeader h;
    const int rc = read_header(file.f, h, path);
    if (rc) return rc;
    if (np) *np = h.np;
    if (ng) *ng = h.ng;
    if (step) *step = h.step;
    if (a) *a = h.a;
    return PM_OK;
}

extern "C" int pm_snapshot_read(const char* path, int64_t np, double* x, double* y, double* z, double* vx,
                                double* vy, double* vz, void* stream) {
    if (!path || np < 0 || (np > 0 && (!x || !y || !z || !vx || !vy || !vz)))
        return fail(PM_ERR_ARG, "pm_snapshot_read: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    File file;
    file.f = fopen(path, "rb");
    if (!file.f) return fail(PM_ERR_ARG, std::string("pm_snapshot_read: cannot open ") + path);
    SnapHeader h;
    int rc = read_header(file.f, h, path);
    if (rc) return rc;
    if (h.np != np) return fail(PM_ERR_ARG, "pm_snapshot_read: particle count differs from the file's");
    Pinned buf;
    const size_t chunk = std::min<size_t>(kSnapChunk, (size_t)std::max<int64_t>(np, 1));
    if (hipHostMalloc((void**)&buf.p, chunk * sizeof(double)) != hipSuccess)
        return fail(PM_ERR_NOMEM, "pm_snapshot_read: hipHostMalloc");
    double* arr[6] = {x, y, z, vx, vy, vz};
    for (double* dst : arr) {
        for (size_t o = 0; o < (size_t)np; o += chunk) {
            const size_t n = std::min(chunk, (size_t)np - o);
            if (fread(buf.p, sizeof(double), n, file.f) != n) return fail(PM_ERR_ARG, "pm_snapshot_read: short file");
            if (hipMemcpyAsync(dst + o, buf.p, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail(PM_ERR_HIP, "pm_snapshot_read: device copy failed");
        }
    }
    return PM_OK;
}
