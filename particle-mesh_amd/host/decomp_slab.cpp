// One-box decomposition classes over the C ABI (include/pm_api.h).
// Reference: src/decomp/slab/{transpose_slab,reorder_slab,reorder_slab_back,
// transpose_slab_back}.cpp.
#include <stdexcept>
#include <string>

#include "pm_api.h"
#include "reorder_slab.h"
#include "reorder_slab_back.h"
#include "transpose_slab.h"
#include "transpose_slab_back.h"

namespace {
void check(int rc, const char* what) {
    if (rc != PM_OK) throw std::runtime_error(std::string(what) + ": " + pm_last_error());
}
void require(bool ok, const char* msg) {
    if (!ok) throw std::invalid_argument(msg);
}
// The transposes exchange between the ranks of comm; one GPU owns the whole
// box, so comm must be a single rank.
void require_one_rank(MPI_Comm comm, const char* who) {
    int init = 0;
    MPI_Initialized(&init);
    if (!init) throw std::logic_error(std::string(who) + ": MPI is not initialised");
    int size = 0;
    if (comm == MPI_COMM_NULL || MPI_Comm_size(comm, &size) != MPI_SUCCESS || size != 1)
        throw std::invalid_argument(std::string(who) + ": one GPU owns the whole box; comm must hold one rank");
}
}  // namespace

TransposeSlab::TransposeSlab(int Ng, int nx, int ny, int nz, int x0, int dimsx, const std::vector<ptrdiff_t>& all_i0,
                             const std::vector<ptrdiff_t>& all_n0, MPI_Comm comm)
    : comm_(comm) {
    require_one_rank(comm, "TransposeSlab");
    require(Ng >= 2 && nx == Ng + 1 && ny == Ng + 1 && nz == Ng + 1 && x0 == 0 && dimsx == 1,
            "TransposeSlab: one box expects nx = ny = nz = Ng + 1, x0 = 0, dimsx = 1");
    require(all_i0.size() == 1 && all_n0.size() == 1 && all_i0[0] == 0 && all_n0[0] == Ng,
            "TransposeSlab: one slab owner with local_0_start = 0, local_n0 = Ng");
    // transpose_slab.cpp:10-27: the owned planes, then the x ghost plane as
    // the wrap send to slab owner 0 (two blocks for ReorderSlab).
    const int plane = ny * nz;
    sendcounts_ = {Ng * plane, plane};
    sdispls_ = {0, Ng * plane};
    recvcounts_ = sendcounts_;
    rdispls_ = sdispls_;
    recv_total_ = (size_t)nx * plane;
    bytes_ = recv_total_ * sizeof(double);
}

void TransposeSlab::execute(const double* sendbuf, double* recvbuf) const {
    const int ng = (int)(recvcounts_[1] > 0 ? recvcounts_[0] / recvcounts_[1] : 0);
    check(pm_transpose_slab(sendbuf, recvbuf, ng, nullptr), "TransposeSlab::execute");
}

ReorderSlab::ReorderSlab(int Ng, int ny, int nz, ptrdiff_t local_n0, int stride, const std::vector<int>& recvcounts,
                         const std::vector<int>& rdispls, const std::array<int, 3>& dims)
    : ng_(Ng), stride_(stride) {
    require(Ng >= 2 && ny == Ng + 1 && nz == Ng + 1 && local_n0 == Ng && stride >= Ng,
            "ReorderSlab: one box expects ny = nz = Ng + 1, local_n0 = Ng, stride >= Ng");
    require(dims[0] == 1 && dims[1] == 1 && dims[2] == 1, "ReorderSlab: one box expects dims = {1, 1, 1}");
    const size_t total = rdispls.empty() ? 0 : (size_t)rdispls.back() + recvcounts.back();
    require(total == (size_t)(Ng + 1) * ny * nz, "ReorderSlab: recvcounts must cover the whole (Ng+1)^3 cube");
}

void ReorderSlab::execute(const double* cube_in, double* slab_out) const {
    check(pm_reorder_slab(cube_in, slab_out, ng_, stride_, nullptr), "ReorderSlab::execute");
}

ReorderSlabBack::ReorderSlabBack(int Ng, int ny_loc, int nz_loc, int stride, const std::array<int, 3>& dims,
                                 ptrdiff_t local_n0)
    : ng_(Ng), stride_(stride) {
    require(Ng >= 2 && ny_loc == Ng && nz_loc == Ng && local_n0 == Ng && stride >= Ng,
            "ReorderSlabBack: one box expects ny_loc = nz_loc = local_n0 = Ng");
    require(dims[0] == 1 && dims[1] == 1 && dims[2] == 1, "ReorderSlabBack: one box expects dims = {1, 1, 1}");
}

void ReorderSlabBack::execute(const double* slab_in, double* cube_out) const {
    check(pm_reorder_slab_back(slab_in, cube_out, ng_, stride_, nullptr), "ReorderSlabBack::execute");
}

TransposeSlabBack::TransposeSlabBack(int Ng, int color, int nx_loc, int ny1, int nz1, std::array<int, 3> dims,
                                     ptrdiff_t local_0_start, ptrdiff_t local_n0, MPI_Comm comm)
    : ng_(Ng), comm_(comm) {
    require_one_rank(comm, "TransposeSlabBack");
    require(Ng >= 2 && color == 1 && nx_loc == Ng && ny1 == Ng + 3 && nz1 == Ng + 3 && local_0_start == 0 &&
                local_n0 == Ng,
            "TransposeSlabBack: one box expects color 1, nx_loc = Ng, ny1 = nz1 = Ng + 3");
    require(dims[0] == 1 && dims[1] == 1 && dims[2] == 1, "TransposeSlabBack: one box expects dims = {1, 1, 1}");
}

void TransposeSlabBack::execute(const double* sendbuf, double* recvbuf) const {
    check(pm_transpose_slab_back(sendbuf, recvbuf, ng_, nullptr), "TransposeSlabBack::execute");
}
