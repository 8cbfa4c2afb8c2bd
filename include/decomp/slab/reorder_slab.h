// ReorderSlab at one box (reorder_slab.h:10-20, reorder_slab.cpp:3-68 of the
// reference): the received cube blocks are gathered into the padded slab
// [x][y][z] of row stride `stride`, summing the periodic ghost duplicates --
// a deterministic HIP gather kernel (pm_reorder_slab) instead of the
// reference's sorted index plan.
#pragma once
#include <array>
#include <cstddef>
#include <vector>

class ReorderSlab {
public:
    ReorderSlab(int Ng, int ny, int nz, ptrdiff_t local_n0, int stride, const std::vector<int>& recvcounts,
                const std::vector<int>& rdispls, const std::array<int, 3>& dims);

    void execute(const double* cube_in, double* slab_out) const;

private:
    int ng_, stride_;
};
