// The FFT class at one box over pm_fft_* (reference: src/fft/ffte1/fft.cpp).
#include "fft.h"

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "pm_api.h"

namespace {
void check(int rc, const char* what) {
    if (rc != PM_OK) throw std::runtime_error(std::string(what) + ": " + pm_last_error());
}

// wisdom_<n>_mpi_<size>.dat of fftw/fft.cpp:28 for the one-GPU build, in
// $PM_WISDOM_DIR if set, else in the working directory as the reference does.
std::string wisdom_file(int n) {
    const std::string name = "wisdom_" + std::to_string(n) + "_hip.dat";
    const char* dir = std::getenv("PM_WISDOM_DIR");
    return dir && *dir ? std::string(dir) + "/" + name : name;
}

// The one-rank world the one-GPU build runs in.
int world_size() {
    int init = 0;
    MPI_Initialized(&init);
    if (!init) throw std::logic_error("FFT: MPI is not initialised (construct MPIEnv first, as main.cpp:10 does)");
    int size = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    return size;
}

bool readable(const std::string& path) {
    std::FILE* fp = std::fopen(path.c_str(), "r");
    if (fp) std::fclose(fp);
    return fp != nullptr;
}
}  // namespace

// fftw/fft.cpp:28-49: import the wisdom file if it is there; otherwise
// measure (here: time every FFT pass shape on buf, which is overwritten as
// FFTW_EXHAUSTIVE planning overwrites its arrays) and export at destruction.
FFT::FFT(int n, double* buf)
    : n_(n), real_(buf), h_(nullptr), comm_(MPI_COMM_NULL), filename_(wisdom_file(n)), need_export_(false) {
    if (!buf) throw std::invalid_argument("FFT: null buffer");
    // ffte1/fft.cpp:4-8: ranks below n/2 form the FFT communicator; one GPU
    // holds the whole box, so the world is that one rank
    if (world_size() != 1) throw std::invalid_argument("FFT: one GPU owns the whole box; run one MPI rank");
    int rank = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_split(MPI_COMM_WORLD, 1, rank, &comm_);
    const bool loaded = readable(filename_) && pm_wisdom_import(filename_.c_str()) == PM_OK;
    const int rc = pm_fft_create(&h_, n, PM_FP64, nullptr);
    if (rc != PM_OK) {
        MPI_Comm_free(&comm_);
        check(rc, "FFT::FFT");
    }
    if (!loaded) {
        const int rc = pm_fft_tune(h_, buf, 5, nullptr);
        if (rc != PM_OK) {
            pm_fft_destroy(h_);
            MPI_Comm_free(&comm_);
            check(rc, "FFT::FFT tune");
        }
        need_export_ = true;
    }
}

// fftw/fft.cpp:63-70
FFT::~FFT() {
    if (need_export_) (void)pm_wisdom_export(filename_.c_str());
    pm_fft_destroy(h_);
    MPI_Comm_free(&comm_);  // ffte1/fft.cpp:39-44
}

MPI_Comm FFT::comm() const { return comm_; }
int FFT::color() const { return 1; }

// complex elements of the half spectrum (ffte1/fft.cpp:13-19 at one rank)
ptrdiff_t FFT::local_alloc() const { return (ptrdiff_t)n_ * n_ * (n_ / 2 + 1); }
ptrdiff_t FFT::local_n0() const { return n_; }
ptrdiff_t FFT::local_0_start() const { return 0; }

void FFT::forward() { check(pm_fft_forward(h_, real_), "FFT::forward"); }
void FFT::backward() { check(pm_fft_backward(h_, real_), "FFT::backward"); }
void FFT::apply_green(const double a) const { check(pm_fft_apply_green(h_, a), "FFT::apply_green"); }
