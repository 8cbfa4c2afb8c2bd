// One-box PM step driver on one MI355X: the reference's src/core/main.cpp
// (main.cpp:9-152) with its 9-phase loop, rebuilt on the drop-in classes
// (FFT, TransposeSlab, ReorderSlab, ReorderSlabBack, TransposeSlabBack and
// the utils functions) over device buffers.  It runs as one MPI rank that
// owns the whole box (MPIEnv, the communicators of main.cpp:46,61-62 and
// report_max_time as in the reference; more ranks are rejected); every phase
// is timed as wall clock around a device synchronisation.
//
//   pm_main Ng Nrow [--warmup W] [--loop L] [--dump FILE]
//
// prints the mean seconds of the 9 phases (main.cpp:141-143), one per line;
// --dump writes the final momenta px, py, pz (3 * Nrow^3 doubles).
#include <hip/hip_runtime.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fft.h"
#include "mpi_env.h"
#include "reorder_slab.h"
#include "reorder_slab_back.h"
#include "transpose_slab.h"
#include "transpose_slab_back.h"
#include "utils.h"

namespace {
double* dalloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    return static_cast<double*>(p);
}
void sync() {
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
}
}  // namespace

int run(int argc, char** argv);

int main(int argc, char** argv) {
    MPIEnv mpi(argc, argv);  // main.cpp:10
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "pm_main: %s\n", e.what());
        return 1;
    }
}

int run(int argc, char** argv) {
    int world_rank = 0, world_size = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &world_rank);
    MPI_Comm_size(MPI_COMM_WORLD, &world_size);
    if (world_size != 1) throw std::invalid_argument("one GPU owns the whole box: run one MPI rank");
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s Ng Nrow [--warmup W] [--loop L] [--dump FILE]\n", argv[0]);
        return 2;
    }
    const int Ng = std::atoi(argv[1]);
    const int Nrow = std::atoi(argv[2]);
    int worm_up = 10, loop = 20;
    std::string dump;
    for (int i = 3; i + 1 < argc; i += 2) {
        if (!std::strcmp(argv[i], "--warmup")) worm_up = std::atoi(argv[i + 1]);
        else if (!std::strcmp(argv[i], "--loop")) loop = std::atoi(argv[i + 1]);
        else if (!std::strcmp(argv[i], "--dump")) dump = argv[i + 1];
    }
    if (Ng < 8 || Nrow < 1) throw std::invalid_argument("Ng >= 8 and Nrow >= 1 required");
    const size_t Np = (size_t)Nrow * Nrow * Nrow;
    const double mass = (double)Ng * Ng * Ng / (double)Np;  // main.cpp:23
    const int stride = Ng;                                   // the dense (ffte1) slab layout
    const std::array<int, 3> dims3 = {1, 1, 1};
    const int nx_loc = Ng, ny_loc = Ng, nz_loc = Ng;
    const int nx = nx_loc + 1, ny = ny_loc + 1, nz = nz_loc + 1;
    const int nx1 = nx + 2, ny1 = ny + 2, nz1 = nz + 2;
    const size_t ncell_local = (size_t)nx * ny * nz, ncell_local1 = (size_t)nx1 * ny1 * nz1;

    double* real_buf = dalloc(ncell_local);                   // cube with + ghosts
    double* recv_buf = dalloc(ncell_local);                   // after the transpose
    double* slab = dalloc((size_t)Ng * Ng * Ng);               // FFT slab
    double* yz_buf = dalloc((size_t)Ng * ny1 * nz1);           // after ReorderSlabBack
    double* halo = dalloc(ncell_local1);                      // after TransposeSlabBack
    double* real_buf1 = dalloc(ncell_local);
    double* real_buf2 = dalloc(ncell_local);
    double* real_buf3 = dalloc(ncell_local);

    FFT fft(Ng, slab);
    MPI_Comm fft_comm = fft.comm();
    const int color = fft.color();
    const ptrdiff_t local_n0 = fft.local_n0(), local_0_start = fft.local_0_start();
    std::vector<ptrdiff_t> all_i0{local_0_start}, all_n0{local_n0};
    TransposeSlab trans(Ng, nx, ny, nz, 0, dims3[0], all_i0, all_n0, MPI_COMM_WORLD);
    TransposeSlabBack trans_back(Ng, color, nx_loc, ny1, nz1, dims3, local_0_start, local_n0, MPI_COMM_WORLD);
    auto reorder = std::make_unique<ReorderSlab>(Ng, ny, nz, local_n0, stride, trans.recvcounts(), trans.rdispls(),
                                                 dims3);
    auto reorder_back = std::make_unique<ReorderSlabBack>(Ng, ny_loc, nz_loc, stride, dims3, local_n0);

    double* posx = dalloc(Np);
    double* posy = dalloc(Np);
    double* posz = dalloc(Np);
    double* px = dalloc(Np);
    double* py = dalloc(Np);
    double* pz = dalloc(Np);
    init_particles(posx, posy, posz, Np, 0, 0, 0, nx_loc, ny_loc, nz_loc, world_rank);
    if (hipMemset(px, 0, Np * sizeof(double)) != hipSuccess || hipMemset(py, 0, Np * sizeof(double)) != hipSuccess ||
        hipMemset(pz, 0, Np * sizeof(double)) != hipSuccess)
        throw std::runtime_error("hipMemset failed");
    sync();

    const double a = Start_a;
    const int all_loop = worm_up + loop;
    double now_time = 0.0;
    std::array<double, 9> exe_time;
    exe_time.fill(0.0);
    for (int i = 0; i < all_loop; ++i) {
        getTime(&now_time);
        init_buf(real_buf, nx, ny, nz, ncell_local, 1);
        deposit_particles(posx, posy, posz, Np, real_buf, ncell_local, nx, ny, nz, 0, 0, 0, mass);
        reduce_private_to_shared(real_buf, ncell_local, 1);
        sync();
        report_max_time(getTime(&now_time), MPI_COMM_WORLD, i, worm_up, exe_time[0]);

        trans.execute(real_buf, recv_buf);
        sync();
        report_max_time(getTime(&now_time), MPI_COMM_WORLD, i, worm_up, exe_time[1]);

        if (color == 1) {
            reorder->execute(recv_buf, slab);
            sync();
            report_max_time(getTime(&now_time), fft_comm, i, worm_up, exe_time[2]);
            fft.forward();
            sync();
            report_max_time(getTime(&now_time), fft_comm, i, worm_up, exe_time[3]);
            fft.apply_green(a);
            sync();
            report_max_time(getTime(&now_time), fft_comm, i, worm_up, exe_time[4]);
            fft.backward();
            sync();
            report_max_time(getTime(&now_time), fft_comm, i, worm_up, exe_time[5]);
            reorder_back->execute(slab, yz_buf);
            sync();
            report_max_time(getTime(&now_time), fft_comm, i, worm_up, exe_time[6]);
        }
        trans_back.execute(yz_buf, halo);
        sync();
        report_max_time(getTime(&now_time), MPI_COMM_WORLD, i, worm_up, exe_time[7]);

        centered_derivatives(halo, real_buf1, real_buf2, real_buf3, nx1, ny1, nz1);
        update_particles(posx, posy, posz, real_buf1, real_buf2, real_buf3, px, py, pz, a, Np, ny, nz, 0, 0, 0);
        sync();
        report_max_time(getTime(&now_time), MPI_COMM_WORLD, i, worm_up, exe_time[8]);
    }
    if (world_rank == 0)
        for (double t : exe_time) std::cout << (loop > 0 ? t / loop : 0.0) << std::endl;

    if (!dump.empty()) {
        std::vector<double> h(3 * Np);
        if (hipMemcpy(h.data(), px, Np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(h.data() + Np, py, Np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(h.data() + 2 * Np, pz, Np * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        FILE* f = std::fopen(dump.c_str(), "wb");
        if (!f || std::fwrite(h.data(), sizeof(double), h.size(), f) != h.size()) throw std::runtime_error("dump failed");
        std::fclose(f);
    }
    for (double* p : {real_buf, recv_buf, slab, yz_buf, halo, real_buf1, real_buf2, real_buf3, posx, posy, posz, px,
                      py, pz})
        (void)hipFree(p);
    return 0;
}
