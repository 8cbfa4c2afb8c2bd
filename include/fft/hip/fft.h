// The FFT class of the reference (fftw/fft.h:10-48, ffte1/fft.h:12-42) for
// one GPU owning the whole box, over the C ABI pm_fft_* (include/pm_api.h).
// buf is the caller-owned DEVICE slab of Ng^3 doubles (row stride Ng, the
// ffte1 layout); forward() and backward() work in place on it; the half
// spectrum lives in the handle.  As in ffte1, backward() is normalised by
// 1/Ng^3 and apply_green() multiplies by the reference's KH Green's function
// times 1/a (ffte1/fft.cpp:28-35,46-52).  As in the reference, the
// constructor splits MPI_COMM_WORLD by color into comm() (ffte1/fft.cpp:4-8);
// one GPU owns the whole box, so the world must hold exactly one rank
// (std::invalid_argument otherwise) and color() is 1.  As the reference's FFTW
// class does with its wisdom file (fftw/fft.cpp:28-49,63-70), the
// constructor imports wisdom_<n>_hip.dat (from $PM_WISDOM_DIR, else the
// working directory) when it exists, else times the FFT pass shapes on buf
// (overwriting it) and the destructor exports the cache to that file.
#pragma once
#include <mpi.h>

#include <cstddef>
#include <string>

struct pm_fft;

class FFT {
public:
    FFT(int n, double* buf);
    ~FFT();
    FFT(const FFT&) = delete;
    FFT& operator=(const FFT&) = delete;

    MPI_Comm comm() const;
    int color() const;

    ptrdiff_t local_alloc() const;
    ptrdiff_t local_n0() const;
    ptrdiff_t local_0_start() const;

    void forward();
    void backward();

    void apply_green(const double a) const;

private:
    int n_;
    double* real_;
    pm_fft* h_;
    MPI_Comm comm_;
    std::string filename_;  // wisdom file (fftw/fft.cpp:28)
    bool need_export_;
};
