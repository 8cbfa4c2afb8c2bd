// TransposeSlab at one box (transpose_slab.h:4-12, transpose_slab.cpp:3-39 of
// the reference).  The rank owns the whole (Ng+1)^3 cube with its + ghost
// planes and is the only slab owner, so the Alltoallv becomes a device copy
// of the cube: the Ng owned x planes followed by the x ghost plane, which
// recvcounts()/rdispls() describe as two blocks (the ghost plane is the
// "wrap" send to slab owner 0, transpose_slab.cpp:23-27).  The communicator
// is the reference's (transpose_slab.h:6); it must hold one rank.
#pragma once
#include <mpi.h>

#include <cstddef>
#include <vector>

#include "transpose_plan.h"

class TransposeSlab : public TransposePlan {
public:
    // nx, ny, nz: cube sides with the + ghost plane (Ng+1); x0: cube origin
    // (0); dimsx: ranks along x (1); all_i0/all_n0: slab starts and sizes
    // of the slab owners ({0}, {Ng}).  Throws std::invalid_argument when
    // the arguments do not describe one box or comm has more than one rank.
    TransposeSlab(int Ng, int nx, int ny, int nz, int x0, int dimsx, const std::vector<ptrdiff_t>& all_i0,
                  const std::vector<ptrdiff_t>& all_n0, MPI_Comm comm);

    void execute(const double* sendbuf, double* recvbuf) const override;

private:
    size_t bytes_;
    [[maybe_unused]] MPI_Comm comm_;  // the caller's communicator, of one rank (checked at construction)
};
