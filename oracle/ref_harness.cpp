// Test infrastructure only: a ctypes-callable harness around the reference's
// own, unmodified single-rank operators (src/utils/utils.cpp,
// src/decomp/slab/reorder_slab*.cpp, src/fft/ffte1/fft.cpp + FFTE).  It is
// linked by oracle/Makefile into oracle/_ref/libpmref.so and used only by the
// tests and the golden-fixture script to pin the FP64 oracle (pm_oracle.c).
//
// At one rank the reference's MPI_Alltoallv transposes are defective (the
// wrap override at transpose_slab.cpp:23-27 and transpose_slab_back.cpp:22-28
// replaces the whole send).  Their intended one-rank semantics are plain
// copies, so this harness performs them itself:
//   * cube -> recv buffer: the (Ng+1)^3 cube is sent as-is; its trailing x
//     ghost plane forms a second "peer" block so that ReorderSlab folds it
//     onto slab plane 0 (reorder_slab.cpp:15-22 computes xi per block).
//   * slab halo -> cube: x planes {Ng-1, 0..Ng-1, 0, 1} of the y/z-haloed
//     slab from ReorderSlabBack, i.e. the -1/+2 periodic x halo that
//     transpose_slab_back.cpp:22-28 assembles across ranks.
#include <mpi.h>
#include <omp.h>
#include <cstring>
#include <vector>
#include <array>
#include <cstdlib>

#include "utils.h"
#include "fft.h"
#include "reorder_slab.h"
#include "reorder_slab_back.h"

namespace {
void ensure_mpi() {
    int ini = 0;
    MPI_Initialized(&ini);
    if (!ini) {
        int provided = 0;
        MPI_Init_thread(nullptr, nullptr, MPI_THREAD_FUNNELED, &provided);
    }
}
}  // namespace

extern "C" {

// init_particles for one rank owning the whole box (utils.cpp:22-36).
void ref_init_particles(long long np, int ng, double* posx, double* posy, double* posz) {
    init_particles(posx, posy, posz, (size_t)np, 0, 0, 0, ng, ng, ng, 0);
}

// One reference step at one rank: main.cpp:92-137 with a = Start_a.
// Outputs (any may be null):
//   cube   (Ng+1)^3 density contrast after init_buf+deposit+reduce
//   slab   Ng^3 folded density contrast (ReorderSlab output, stride Ng)
//   phi    Ng^3 potential after forward/apply_green/backward
//   halo   (Ng+3)^3 potential cube with -1/+2 halo
//   dx,dy,dz (Ng+1)^3 centred differences
//   px,py,pz updated in place by update_particles (one kick)
int ref_step(int ng, long long np, double a,
             const double* posx, const double* posy, const double* posz,
             double* px, double* py, double* pz,
             double* cube_out, double* slab_out, double* phi_out, double* halo_out,
             double* dx_out, double* dy_out, double* dz_out) {
    ensure_mpi();
    const int nthreads = omp_get_max_threads();
    const int nx = ng + 1, ny = ng + 1, nz = ng + 1;
    const int nx1 = nx + 2, ny1 = ny + 2, nz1 = nz + 2;
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t ncell1 = (size_t)nx1 * ny1 * nz1;
    const double mass = (double)ng * ng * ng / (double)np;
    const int stride = ng;  // ffte1 layout (main.cpp:24-28)

    std::vector<double> priv(ncell * nthreads);
    init_buf(priv.data(), nx, ny, nz, ncell, nthreads);
    deposit_particles(posx, posy, posz, (size_t)np, priv.data(), ncell, nx, ny, nz, 0, 0, 0, mass);
    reduce_private_to_shared(priv.data(), ncell, nthreads);
    if (cube_out) std::memcpy(cube_out, priv.data(), ncell * sizeof(double));

    // FFT work area: real slab + in-place complex + FFTE scratch.
    const size_t local_alloc = (size_t)ng * ng * (ng / 2 + 1);
    double* fbuf = static_cast<double*>(std::aligned_alloc(64, ((4 * local_alloc + 64) * sizeof(double) + 63) / 64 * 64));
    std::memset(fbuf, 0, (4 * local_alloc + 64) * sizeof(double));
    {
        FFT fft(ng, fbuf);
        std::array<int, 3> dims{1, 1, 1};
        std::vector<int> rc{(int)((size_t)ng * ny * nz), ny * nz};
        std::vector<int> rd{0, (int)((size_t)ng * ny * nz)};
        ReorderSlab reorder(ng, ny, nz, ng, stride, rc, rd, dims);
        reorder.execute(priv.data(), fbuf);
        if (slab_out) std::memcpy(slab_out, fbuf, (size_t)ng * ng * ng * sizeof(double));
        fft.forward();
        fft.apply_green(a);
        fft.backward();
        if (phi_out) std::memcpy(phi_out, fbuf, (size_t)ng * ng * ng * sizeof(double));

        ReorderSlabBack reorder_back(ng, ng, ng, stride, dims, ng);
        const size_t plane = (size_t)ny1 * nz1;
        std::vector<double> yz((size_t)ng * plane);
        reorder_back.execute(fbuf, yz.data());
        std::vector<double> halo(ncell1);
        const int src_plane[3] = {ng - 1, 0, 1};
        std::memcpy(halo.data(), yz.data() + (size_t)src_plane[0] * plane, plane * sizeof(double));
        std::memcpy(halo.data() + plane, yz.data(), (size_t)ng * plane * sizeof(double));
        std::memcpy(halo.data() + (size_t)(ng + 1) * plane, yz.data() + (size_t)src_plane[1] * plane, plane * sizeof(double));
        std::memcpy(halo.data() + (size_t)(ng + 2) * plane, yz.data() + (size_t)src_plane[2] * plane, plane * sizeof(double));
        if (halo_out) std::memcpy(halo_out, halo.data(), ncell1 * sizeof(double));

        std::vector<double> gx(ncell), gy(ncell), gz(ncell);
        centered_derivatives(halo.data(), gx.data(), gy.data(), gz.data(), nx1, ny1, nz1);
        if (dx_out) std::memcpy(dx_out, gx.data(), ncell * sizeof(double));
        if (dy_out) std::memcpy(dy_out, gy.data(), ncell * sizeof(double));
        if (dz_out) std::memcpy(dz_out, gz.data(), ncell * sizeof(double));
        if (px && py && pz)
            update_particles(posx, posy, posz, gx.data(), gy.data(), gz.data(), px, py, pz, a, (size_t)np, ny, nz, 0, 0, 0);
    }
    std::free(fbuf);
    return 0;
}

// The FFTE forward + inverse round trip alone, for FFT-convention checks:
// data is the Ng^3 real slab (stride Ng); spec_out receives local_alloc
// complex values (interleaved) in the reference's transposed-out order.
int ref_fft_roundtrip(int ng, double* data, double* spec_out) {
    ensure_mpi();
    const size_t local_alloc = (size_t)ng * ng * (ng / 2 + 1);
    double* fbuf = static_cast<double*>(std::aligned_alloc(64, ((4 * local_alloc + 64) * sizeof(double) + 63) / 64 * 64));
    std::memset(fbuf, 0, (4 * local_alloc + 64) * sizeof(double));
    {
        FFT fft(ng, fbuf);
        std::memcpy(fbuf, data, (size_t)ng * ng * ng * sizeof(double));
        fft.forward();
        if (spec_out) std::memcpy(spec_out, fbuf, 2 * local_alloc * sizeof(double));
        fft.backward();
        std::memcpy(data, fbuf, (size_t)ng * ng * ng * sizeof(double));
    }
    std::free(fbuf);
    return 0;
}

}  // extern "C"
