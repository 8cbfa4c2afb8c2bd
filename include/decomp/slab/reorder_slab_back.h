// ReorderSlabBack at one box (reorder_slab_back.h:8-22,
// reorder_slab_back.cpp:3-49 of the reference): slab -> Ng x (Ng+3) x (Ng+3)
// block with the -1/+2 periodic halo on y and z (pm_reorder_slab_back).
#pragma once
#include <array>
#include <cstddef>

class ReorderSlabBack {
public:
    ReorderSlabBack(int Ng, int ny_loc, int nz_loc, int stride, const std::array<int, 3>& dims, ptrdiff_t local_n0);

    void execute(const double* slab_in, double* cube_out) const;

private:
    int ng_, stride_;
};
