// TransposeSlabBack at one box (transpose_slab_back.h:8-17,
// transpose_slab_back.cpp:3-49 of the reference): adds the x halo -- plane
// -1 from the last slab plane, planes Ng and Ng+1 from the first two -- to
// the y/z-haloed block of ReorderSlabBack, giving the (Ng+3)^3 cube that
// centered_derivatives reads (pm_transpose_slab_back).  The communicator is
// the reference's (transpose_slab_back.h:10); it must hold one rank.
#pragma once
#include <mpi.h>

#include <array>
#include <cstddef>

class TransposeSlabBack {
public:
    TransposeSlabBack(int Ng, int color, int nx_loc, int ny1, int nz1, std::array<int, 3> dims,
                      ptrdiff_t local_0_start, ptrdiff_t local_n0, MPI_Comm comm = MPI_COMM_WORLD);

    void execute(const double* sendbuf, double* recvbuf) const;

private:
    int ng_;
    [[maybe_unused]] MPI_Comm comm_;  // the caller's communicator, of one rank (checked at construction)
};
