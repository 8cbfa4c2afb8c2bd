// The FFT class at one box over pm_fft_* (reference: src/fft/ffte1/fft.cpp).
#include "fft.h"

#include <stdexcept>
#include <string>

#include "pm_api.h"

namespace {
void check(int rc, const char* what) {
    if (rc != PM_OK) throw std::runtime_error(std::string(what) + ": " + pm_last_error());
}
}  // namespace

FFT::FFT(int n, double* buf) : n_(n), real_(buf), h_(nullptr) {
    if (!buf) throw std::invalid_argument("FFT: null buffer");
    check(pm_fft_create(&h_, n, PM_FP64, nullptr), "FFT::FFT");
}

FFT::~FFT() { pm_fft_destroy(h_); }

int FFT::color() const { return 1; }

// complex elements of the half spectrum (ffte1/fft.cpp:13-19 at one rank)
ptrdiff_t FFT::local_alloc() const { return (ptrdiff_t)n_ * n_ * (n_ / 2 + 1); }
ptrdiff_t FFT::local_n0() const { return n_; }
ptrdiff_t FFT::local_0_start() const { return 0; }

void FFT::forward() { check(pm_fft_forward(h_, real_), "FFT::forward"); }
void FFT::backward() { check(pm_fft_backward(h_, real_), "FFT::backward"); }
void FFT::apply_green(const double a) const { check(pm_fft_apply_green(h_, a), "FFT::apply_green"); }
