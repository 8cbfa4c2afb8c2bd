// Interface of the cube<->slab transposes (transpose_plan.h:6-19 of the
// reference), for one GPU owning the whole periodic box.  Buffers are
// device pointers (hipMalloc'd).
#pragma once
#include <cstddef>
#include <vector>

class TransposePlan {
public:
    virtual ~TransposePlan() = default;

    virtual void execute(const double* sendbuf, double* recvbuf) const = 0;

    size_t recv_size() const { return recv_total_; }
    const std::vector<int>& recvcounts() const { return recvcounts_; }
    const std::vector<int>& rdispls() const { return rdispls_; }

protected:
    std::vector<int> sendcounts_, sdispls_, recvcounts_, rdispls_;
    size_t recv_total_{0};
};
