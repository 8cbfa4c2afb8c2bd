// Batched Stockham FFT passes for the 3-D real-to-complex Poisson solve.
//
// Layouts (N = mesh side, M = N/2, all row-major, last index fastest):
//   real mesh   R[x][y][z]            N^3 reals
//   spectrum A  C1[kz][x][y]          (M+1) * N * N complex   (after the z pass)
//   spectrum B  C2[kz][ky][kx]        (M+1) * N * N complex   (after the y pass)
// Every pass reads or writes whole lines on one side and 128-byte segments
// of T consecutive lines on the other, staging a tile of T lines in LDS
// (the x pass, and the contiguous side of the z passes, reads or writes its
// lines in the first or last Stockham stage instead):
//   zfwd : R  -> C1  real FFT along z (M-point complex FFT + split)
//   yfwd : C1 -> C2  complex FFT along y
//   x    : C2 -> C2  FFT along x, Green's multiply, inverse FFT along x
//   yinv : C2 -> C1
//   zinv : C1 -> R   real inverse FFT along z, times a normalisation
// The reference computes the same transform with FFTW-MPI (fftw/fft.cpp:46-47,
// 96-97) or FFTE (pdzfft3d.f:77-265, pzdfft3d.f:76-301), both unnormalised
// with e^{-i} forward; the Green's function follows fftw/fft.cpp:50-57 and
// ffte1/fft.cpp:28-35.
#include "pm_internal.h"

namespace pm {

// Workgroup size of each pass at the largest meshes (measured at 512^3):
// the transposing y passes gain from more waves in flight than the z
// passes, whose LDS tile already limits the workgroups per CU; the x pass
// runs small tiles (4 lines) with 256 threads.
#ifndef PM_FFT_NT_ZF
#define PM_FFT_NT_ZF 512
#endif
#ifndef PM_FFT_NT_YF
#define PM_FFT_NT_YF 1024
#endif
#ifndef PM_FFT_NT_X
#define PM_FFT_NT_X 256
#endif
#ifndef PM_FFT_NT_YI
#define PM_FFT_NT_YI 1024
#endif
#ifndef PM_FFT_NT_ZI
#define PM_FFT_NT_ZI 512
#endif
// at most one thread per 4 elements of the tile (small meshes), at least a wave
__host__ __device__ constexpr int fft_threads(int want, int tile_elems) {
    int nt = want;
    while (nt > 64 && nt * 4 > tile_elems) nt /= 2;
    return nt;
}

// Global loads and stores of the passes: every element is read once and
// written once per pass, with non-temporal hints (bit 0 loads, bit 1
// stores; measured at 512^3: the five passes 1.44 -> 1.39 ms FP32, 2.79 ->
// 2.74 ms FP64, and the interpolation that follows gains as well).
#ifndef PM_FFT_NT
#define PM_FFT_NT 3
#endif
template <typename Rl> __device__ __forceinline__ cplx_t<Rl> gload(const cplx_t<Rl>* p) {
    if constexpr (PM_FFT_NT & 1) {
        typedef Rl v2 __attribute__((ext_vector_type(2)));
        const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return cplx_t<Rl>{t.x, t.y};
    }
    return *p;
}
template <typename Rl> __device__ __forceinline__ void gstore(cplx_t<Rl>* p, cplx_t<Rl> v) {
    if constexpr (PM_FFT_NT & 2) {
        typedef Rl v2 __attribute__((ext_vector_type(2)));
        v2 t;
        t.x = v.x;
        t.y = v.y;
        __builtin_nontemporal_store(t, reinterpret_cast<v2*>(p));
    } else {
        *p = v;
    }
}

// LDS layout of a line: element i sits at lpos(i) = i + i/8, so that the
// stride-8 and stride-64 writes of the Stockham stages spread over the
// banks; the line pitch adds one more complex so that the same element of
// consecutive lines falls in different banks.
__host__ __device__ constexpr int lds_pitch(int L) { return L + (L >> 3) + 1; }
__device__ __forceinline__ int lpos(int i) { return i + (i >> 3); }

template <int R, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<2, DIR, C> {
    static __device__ __forceinline__ void run(C* u) {
        const C a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    }
};

template <int DIR, typename C> struct Dft<4, DIR, C> {
    static __device__ __forceinline__ void run(C* u) {
        const C s02 = cadd(u[0], u[2]), d02 = csub(u[0], u[2]);
        const C s13 = cadd(u[1], u[3]), d13 = mul_i<DIR>(csub(u[1], u[3]));
        u[0] = cadd(s02, s13);
        u[2] = csub(s02, s13);
        u[1] = cadd(d02, d13);
        u[3] = csub(d02, d13);
    }
};

template <int DIR, typename C> struct Dft<8, DIR, C> {
    static __device__ __forceinline__ void run(C* u) {
        using Rl = decltype(u[0].x);
        C e[4] = {u[0], u[2], u[4], u[6]};
        C o[4] = {u[1], u[3], u[5], u[7]};
        Dft<4, DIR, C>::run(e);
        Dft<4, DIR, C>::run(o);
        const Rl c = Rl(0.70710678118654752440);
        // W8^1 = c + i*DIR*c, W8^2 = DIR*i, W8^3 = -c + i*DIR*c
        const C o1 = C{c * (o[1].x - Rl(DIR) * o[1].y), c * (o[1].y + Rl(DIR) * o[1].x)};
        const C o2 = mul_i<DIR>(o[2]);
        const C o3 = C{-c * (o[3].x + Rl(DIR) * o[3].y), c * (Rl(DIR) * o[3].x - o[3].y)};
        u[0] = cadd(e[0], o[0]);
        u[4] = csub(e[0], o[0]);
        u[1] = cadd(e[1], o1);
        u[5] = csub(e[1], o1);
        u[2] = cadd(e[2], o2);
        u[6] = csub(e[2], o2);
        u[3] = cadd(e[3], o3);
        u[7] = csub(e[3], o3);
    }
};

// One radix-R Stockham autosort stage over T lines of length L held in LDS
// (line pitch lds_pitch(L)).  p is the length of the sub-transforms already
// done; tw[m] = exp(-2 pi i m / L) for m < L/2, and the stage takes
// exp(-2 pi i m / L) = -tw[m - L/2] above that (the tables are built with
// exactly this symmetry, so only their first half is ever read).
// The stage's inputs come from src(line, i) and its outputs go to
// dst(line, i, v): element i of a line in LDS for every stage but the first
// and the last of a pass, which may read or write global memory instead (a
// stage reads b + r Q and writes b + r p of a line for the Q = L/R lanes b,
// i.e. consecutive lanes take consecutive elements when p = 1 on input and
// p = L/R on output: whole segments).  GIN: the input is not the LDS tile,
// so no barrier is needed between the reads and the writes; GOUT: the
// output is not the LDS tile, so none after the writes.
template <typename C, int L, int T, int R, int DIR, int NT, bool GIN, bool GOUT, typename Src, typename Dst>
__device__ __forceinline__ void stockham_stage_io(int p, const C* __restrict__ tw, const Src& src, const Dst& dst) {
    constexpr int Q = L / R;
    constexpr int NBF = T * Q;
    constexpr int PER = (NBF + NT - 1) / NT;
    C u[PER][R];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int g = threadIdx.x + q * NT;
        if (NBF % NT == 0 || g < NBF) {
            const int line = g / Q, b = g % Q;
#pragma unroll
            for (int r = 0; r < R; ++r) u[q][r] = src(line, b + r * Q);
        }
    }
    if constexpr (!GIN) __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int g = threadIdx.x + q * NT;
        if (NBF % NT == 0 || g < NBF) {
            const int line = g / Q, b = g % Q;
            const int k = b & (p - 1);
            if (p > 1) {
                const int step = L / (R * p);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const int m = r * k * step;
                    const C t = tw[m & (L / 2 - 1)];
                    const C w = (m & (L / 2)) ? C{-t.x, -t.y} : t;
                    u[q][r] = DIR < 0 ? cmul(u[q][r], w) : cmulc(u[q][r], w);
                }
            }
            Dft<R, DIR, C>::run(u[q]);
            const int j = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst(line, j + r * p, u[q][r]);
        }
    }
    if constexpr (!GOUT) __syncthreads();
}
// a product rounded on its own (never contracted into an FMA with the add
// that consumes it)
template <typename Rl> __device__ __forceinline__ Rl rmul(Rl a, Rl b) {
#pragma clang fp contract(off)
    return a * b;
}
// element i of line `line` of the LDS tile s
template <typename C, int L> struct LdsLine {
    C* s;
    __device__ __forceinline__ C operator()(int line, int i) const { return s[line * lds_pitch(L) + lpos(i)]; }
    __device__ __forceinline__ void operator()(int line, int i, C v) const { s[line * lds_pitch(L) + lpos(i)] = v; }
};
template <typename C, int L, int T, int R, int DIR, int NT>
__device__ __forceinline__ void stockham_stage(C* s, int p, const C* __restrict__ tw) {
    const LdsLine<C, L> l{s};
    stockham_stage_io<C, L, T, R, DIR, NT, false, false>(p, tw, l, l);
}
// Full length-L FFT of T lines: radix-8 stages, then one radix-4 or
// radix-2 stage for the remaining factor; the first stage reads src, the
// last writes dst (GIN / GOUT: global memory), the others the LDS tile s.
template <typename C, int L, int T, int DIR, int NT, bool GIN, bool GOUT, int ST, typename Src, typename Dst>
__device__ __forceinline__ void fft_stages_io(C* s, int p, const C* __restrict__ tw, const Src& src, const Dst& dst) {
    constexpr int lg = ilog2(L), n8 = lg / 3, NST = n8 + (lg % 3 ? 1 : 0);
    if constexpr (ST < NST) {
        constexpr int R = ST < n8 ? 8 : (lg % 3 == 2 ? 4 : 2);
        const LdsLine<C, L> l{s};
        if constexpr (ST == 0 && ST == NST - 1)
            stockham_stage_io<C, L, T, R, DIR, NT, GIN, GOUT>(p, tw, src, dst);
        else if constexpr (ST == 0)
            stockham_stage_io<C, L, T, R, DIR, NT, GIN, false>(p, tw, src, l);
        else if constexpr (ST == NST - 1)
            stockham_stage_io<C, L, T, R, DIR, NT, false, GOUT>(p, tw, l, dst);
        else
            stockham_stage_io<C, L, T, R, DIR, NT, false, false>(p, tw, l, l);
        fft_stages_io<C, L, T, DIR, NT, GIN, GOUT, ST + 1>(s, p * R, tw, src, dst);
    }
}
template <typename C, int L, int T, int DIR, int NT, bool GIN, bool GOUT, typename Src, typename Dst>
__device__ __forceinline__ void fft_lines_io(C* s, const C* __restrict__ tw, const Src& src, const Dst& dst) {
    fft_stages_io<C, L, T, DIR, NT, GIN, GOUT, 0>(s, 1, tw, src, dst);
}
// The same on the LDS tile alone.
template <typename C, int L, int T, int DIR, int NT>
__device__ __forceinline__ void fft_lines(C* s, const C* __restrict__ tw) {
    const LdsLine<C, L> l{s};
    fft_lines_io<C, L, T, DIR, NT, false, false>(s, tw, l, l);
}
// Twiddle table of a pass's L-point FFTs: for the passes in the mask of
// its precision (bit = slot: 0 z fwd, 1 y fwd, 2 x, 3 y inv, 4 z inv) a
// copy in LDS (the L/2 complex that the stages read), which the workgroup
// reads after its first barrier, instead of the global table through L1
// (the same values, so the same results bit for bit).  Interleaved A/B at
// 512^3, global -> LDS, ms:
//   FP32  z fwd 0.250 -> 0.214, y fwd 0.227 -> 0.180, x 0.355 -> 0.247,
//         y inv 0.226 -> 0.179, z inv 0.261 -> 0.241
//   FP64  z fwd 0.476 -> 0.450, y fwd 0.405 -> 0.359, x 0.593 -> 0.533,
//         y inv 0.412 -> 0.371, z inv 0.463 -> 0.431 (a full L-entry copy
//         instead of the half cost the y passes a workgroup per CU: 0.53 ms)
#ifndef PM_FFT_TW_LDS32
#define PM_FFT_TW_LDS32 31
#endif
#ifndef PM_FFT_TW_LDS64
#define PM_FFT_TW_LDS64 31
#endif
template <typename C, int SLOT> __host__ __device__ constexpr bool tw_in_lds() {
    return ((sizeof(C) == 8 ? PM_FFT_TW_LDS32 : PM_FFT_TW_LDS64) >> SLOT) & 1;
}
template <typename C, int L, int SLOT> using tw_lds_t = C[tw_in_lds<C, SLOT>() ? L / 2 : 1];
template <typename C, int L, int NT, int SLOT>
__device__ __forceinline__ const C* stage_twiddles(C* tws, const C* __restrict__ tw) {
    if constexpr (tw_in_lds<C, SLOT>()) {
        for (int i = threadIdx.x; i < L / 2; i += NT) tws[i] = tw[i];
        return tws;
    }
    return tw;
}

// Lines per workgroup: 128-byte global segments (16 float2 / 8 double2),
// halved at N >= 1024 to keep the LDS tile at 64 KB.
#ifndef PM_FFT_LINES32
#define PM_FFT_LINES32 16
#endif
#ifndef PM_FFT_LINES64
#define PM_FFT_LINES64 8
#endif
// The fused x pass (contiguous lines on both sides, two FFTs per tile)
// prefers 4 lines with 256 threads: a small LDS tile and many workgroups per
// CU (the shape pm_plan_tune picks at 512^3 in both precisions).
#ifndef PM_FFT_XLINES32
#define PM_FFT_XLINES32 4
#endif
#ifndef PM_FFT_XLINES64
#define PM_FFT_XLINES64 4
#endif
template <typename Rl, bool XPASS = false> __host__ __device__ constexpr int tile_lines(int N) {
    const int base = sizeof(Rl) == 4 ? (XPASS ? PM_FFT_XLINES32 : PM_FFT_LINES32)
                                     : (XPASS ? PM_FFT_XLINES64 : PM_FFT_LINES64);
    const int t = N >= 1024 ? base / 2 : base;
    return t < N ? t : N;
}

// ---------------------------------------------------------------- z forward
template <typename Rl, int N, bool SUB1, int T, int NT>
__global__ __launch_bounds__(NT) void k_zfwd(const Rl* __restrict__ rho,
                                                      cplx_t<Rl>* __restrict__ out,
                                                      const cplx_t<Rl>* __restrict__ twM,
                                                      const cplx_t<Rl>* __restrict__ twN) {
    using C = cplx_t<Rl>;
    constexpr int M = N / 2, LP = lds_pitch(M), YT = N / T;
    __shared__ C s[T * LP];
    __shared__ tw_lds_t<C, M, 0> tws;
    const C* tw = stage_twiddles<C, M, NT, 0>(tws, twM);
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int x = tile / YT, y0 = (tile % YT) * T;
    const C* src = reinterpret_cast<const C*>(rho + ((size_t)x * N + y0) * N);
    // the first stage reads the lines (pairs of reals) from global memory
    const auto gsrc = [src](int line, int i) {
        C v = gload<Rl>(src + (size_t)line * M + i);
        if (SUB1) {
            v.x -= Rl(1);
            v.y -= Rl(1);
        }
        return v;
    };
    fft_lines_io<C, M, T, -1, NT, true, false>(s, tw, gsrc, LdsLine<C, M>{s});
    // X[k] = (Z[k] + conj Z[M-k])/2 - i W^k (Z[k] - conj Z[M-k])/2, Z[M] = Z[0]
    for (int g = threadIdx.x; g < T * (M + 1); g += NT) {
        const int kz = g / T, line = g % T;
        const C* row = s + line * LP;
        const C zk = row[lpos(kz == M ? 0 : kz)];
        const C zm = row[lpos(kz == 0 ? 0 : M - kz)];
        const C e = C{Rl(0.5) * (zk.x + zm.x), Rl(0.5) * (zk.y - zm.y)};
        const C d = C{Rl(0.5) * (zk.x - zm.x), Rl(0.5) * (zk.y + zm.y)};
        const C wd = cmul(twN[kz], d);
        gstore<Rl>(out + ((size_t)kz * N + x) * N + y0 + line, C{e.x + wd.y, e.y - wd.x});
    }
}

// ---------------------------------------------------------------- y forward
template <typename Rl, int N, int T, int NT>
__global__ __launch_bounds__(NT) void k_yfwd(const cplx_t<Rl>* __restrict__ in,
                                                      cplx_t<Rl>* __restrict__ out,
                                                      const cplx_t<Rl>* __restrict__ twN) {
    using C = cplx_t<Rl>;
    constexpr int LP = lds_pitch(N), XT = N / T;
    __shared__ C s[T * LP];
    __shared__ tw_lds_t<C, N, 1> tws;
    const C* tw = stage_twiddles<C, N, NT, 1>(tws, twN);
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int kz = tile / XT, x0 = (tile % XT) * T;
    const C* src = in + ((size_t)kz * N + x0) * N;
    // (a first stage reading global memory measured slower here: 0.184 -> 0.187 ms)
    for (int g = threadIdx.x; g < T * N; g += NT) s[(g / N) * LP + lpos(g % N)] = gload<Rl>(src + g);
    __syncthreads();
    fft_lines<C, N, T, -1, NT>(s, tw);
    C* dst = out + (size_t)kz * N * N + x0;
    for (int g = threadIdx.x; g < T * N; g += NT) {
        const int ky = g / T, line = g % T;
        gstore<Rl>(dst + (size_t)ky * N + line, s[line * LP + lpos(ky)]);
    }
}

// Green's function times 'scale' at integer wavevector (i, j, k), i, j in
// [0, N), k in [0, N/2]; cos_tab[n] = cos(2 pi n / N).
__device__ __forceinline__ double green_value(int mode, int N, int i, int j, int k,
                                              const double* __restrict__ cos_tab) {
    if ((i | j | k) == 0) return 0.0;
    if (mode == PM_GREEN_REF_KH) {
        // ffte1/fft.cpp:33 (fftw/fft.cpp:55 carries the 1/N^3 separately)
        return -3.0 * 1.0 / (4.0 * (3.0 - cos_tab[i] - cos_tab[j] - cos_tab[k]));
    }
    const int si = i <= N / 2 ? i : i - N;
    const int sj = j <= N / 2 ? j : j - N;
    const int sk = k <= N / 2 ? k : k - N;
    const double n2 = (double)si * si + (double)sj * sj + (double)sk * sk;
    return -((double)N * N) / (3.14159265358979323846 * n2);
}

// ------------------------------------------------- x: forward/Green/inverse
template <typename Rl, int N, int MODE, int T, int NT>
__global__ __launch_bounds__(NT) void k_x(cplx_t<Rl>* __restrict__ data,
                                                   const cplx_t<Rl>* __restrict__ twN,
                                                   GreenArgs ga) {
    using C = cplx_t<Rl>;
    constexpr int LP = lds_pitch(N), YT = N / T;
    __shared__ C s[T * LP];
    __shared__ tw_lds_t<C, N, 2> tws;
    const C* tw = stage_twiddles<C, N, NT, 2>(tws, twN);
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int kz = tile / YT, ky0 = (tile % YT) * T;
    C* base = data + ((size_t)kz * N + ky0) * N;
    if constexpr (MODE == (kXForward | kXGreen | kXInverse)) {
        // the forward FFT's first stage reads the lines from global memory,
        // Green's multiply rides on the inverse FFT's first reads, and its
        // last stage writes the lines back: three LDS round trips and five
        // barriers fewer than load, transform, multiply, transform, store
        const auto gsrc = [base](int line, int i) { return gload<Rl>(base + (size_t)line * N + i); };
        const auto gdst = [base](int line, int i, C v) { gstore<Rl>(base + (size_t)line * N + i, v); };
        const LdsLine<C, N> l{s};
        fft_lines_io<C, N, T, -1, NT, true, false>(s, tw, gsrc, l);
        // (rounded products, never fused into the butterflies' adds: the
        // values of a separate multiply pass, bit for bit)
        const auto green = [&](int line, int kx) {
            const C v = l(line, kx);
            const Rl gv = (Rl)(green_value(ga.mode, N, kx, ky0 + line, kz, ga.cos_tab) * ga.scale);
            return C{rmul(v.x, gv), rmul(v.y, gv)};
        };
        fft_lines_io<C, N, T, +1, NT, false, true>(s, tw, green, gdst);
        return;
    }
    for (int g = threadIdx.x; g < T * N; g += NT) s[(g / N) * LP + lpos(g % N)] = gload<Rl>(base + g);
    __syncthreads();
    if constexpr (MODE & kXForward) fft_lines<C, N, T, -1, NT>(s, tw);
    if constexpr (MODE & kXGreen) {
        for (int g = threadIdx.x; g < T * N; g += NT) {
            const int line = g / N, kx = g % N;
            const Rl gv = (Rl)(green_value(ga.mode, N, kx, ky0 + line, kz, ga.cos_tab) * ga.scale);
            C& v = s[line * LP + lpos(kx)];
            v.x *= gv;
            v.y *= gv;
        }
        __syncthreads();
    }
    if constexpr (MODE & kXInverse) fft_lines<C, N, T, +1, NT>(s, tw);
    for (int g = threadIdx.x; g < T * N; g += NT) gstore<Rl>(base + g, s[(g / N) * LP + lpos(g % N)]);
}

// ---------------------------------------------------------------- y inverse
template <typename Rl, int N, int T, int NT>
__global__ __launch_bounds__(NT) void k_yinv(const cplx_t<Rl>* __restrict__ in,
                                                      cplx_t<Rl>* __restrict__ out,
                                                      const cplx_t<Rl>* __restrict__ twN) {
    using C = cplx_t<Rl>;
    constexpr int LP = lds_pitch(N), XT = N / T;
    __shared__ C s[T * LP];
    __shared__ tw_lds_t<C, N, 3> tws;
    const C* tw = stage_twiddles<C, N, NT, 3>(tws, twN);
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int kz = tile / XT, x0 = (tile % XT) * T;
    const C* src = in + (size_t)kz * N * N + x0;
    for (int g = threadIdx.x; g < T * N; g += NT) {
        const int ky = g / T, line = g % T;
        s[line * LP + lpos(ky)] = gload<Rl>(src + (size_t)ky * N + line);
    }
    __syncthreads();
    // (a last stage writing global memory measured slower here: 0.186 -> 0.188 ms)
    fft_lines<C, N, T, +1, NT>(s, tw);
    C* dst = out + ((size_t)kz * N + x0) * N;
    for (int g = threadIdx.x; g < T * N; g += NT) gstore<Rl>(dst + g, s[(g / N) * LP + lpos(g % N)]);
}

// ---------------------------------------------------------------- z inverse
template <typename Rl, int N, int T, int NT>
__global__ __launch_bounds__(NT) void k_zinv(const cplx_t<Rl>* __restrict__ in,
                                                      Rl* __restrict__ out,
                                                      const cplx_t<Rl>* __restrict__ twM,
                                                      const cplx_t<Rl>* __restrict__ twN, Rl norm) {
    using C = cplx_t<Rl>;
    constexpr int M = N / 2, LP = lds_pitch(M), YT = N / T;
    __shared__ C s[T * LP];
    __shared__ tw_lds_t<C, M, 4> tws;
    const C* tw = stage_twiddles<C, M, NT, 4>(tws, twM);
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int x = tile / YT, y0 = (tile % YT) * T;
    for (int g = threadIdx.x; g < T * (M + 1); g += NT) {
        const int kz = g / T, line = g % T;
        s[line * LP + lpos(kz)] = gload<Rl>(in + ((size_t)kz * N + x) * N + y0 + line);
    }
    __syncthreads();
    // Z[k] = (X[k] + conj X[M-k]) + i W^{-k} (X[k] - conj X[M-k]), k in [0, M),
    // formed as the first stage reads it (the imaginary parts of X[0] and
    // X[M] are dropped: Hermitian input); the last stage writes the lines,
    // normalised, to global memory
    const auto zsrc = [&](int line, int k) {
        const C* row = s + line * LP;
        C xk = row[lpos(k)], xm = row[lpos(M - k)];
        if (k == 0) {
            xk.y = Rl(0);
            xm.y = Rl(0);
        }
        const C a = C{xk.x + xm.x, xk.y - xm.y};
        const C b = cmulc(C{xk.x - xm.x, xk.y + xm.y}, twN[k]);
        return C{a.x - b.y, a.y + b.x};
    };
    C* dst = reinterpret_cast<C*>(out + ((size_t)x * N + y0) * N);
    const auto gdst = [dst, norm](int line, int i, C v) {
        gstore<Rl>(dst + (size_t)line * M + i, C{rmul(v.x, norm), rmul(v.y, norm)});
    };
    fft_lines_io<C, M, T, +1, NT, false, true>(s, tw, zsrc, gdst);
}

// ---------------------------------------------------------- apply_green
template <typename Rl>
__global__ __launch_bounds__(256) void k_apply_green(cplx_t<Rl>* __restrict__ spec, int N,
                                                     GreenArgs ga) {
    const size_t total = (size_t)(N / 2 + 1) * N * N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e % N);
        const int ky = (int)((e / N) % N);
        const int kz = (int)(e / ((size_t)N * N));
        const Rl gv = (Rl)(green_value(ga.mode, N, kx, ky, kz, ga.cos_tab) * ga.scale);
        cplx_t<Rl> v = spec[e];
        v.x *= gv;
        v.y *= gv;
        spec[e] = v;
    }
}

// ================================================================ launchers
// Shape variants of the five tunable passes (slots zfwd, yfwd, x, yinv,
// zinv): variant v takes the default lines per tile (v < 3), half of them
// (3 <= v < 6) or a quarter (v >= 6), at least one, and threads per
// workgroup: the pass's default (v % 3 == 0), then the smaller and the larger
// of the other two of 256 / 512 / 1024, clamped by fft_threads.  Variant 0 is the default shape; every variant computes the
// same transform bit for bit (only the work split differs).  Meshes below
// kTuneMinN (pm_internal.h) always take variant 0.

// FP64 tiles hold half the lines of FP32 ones (the same 128-byte segments):
// the z forward and the y passes take half the threads (measured at 512^3
// and 1024^3 by pm_plan_tune).
template <typename Rl> __host__ __device__ constexpr int slot_default_threads(int slot) {
    constexpr bool d = sizeof(Rl) == 8;
    return slot == 0 ? (d ? PM_FFT_NT_ZF / 2 : PM_FFT_NT_ZF) : slot == 1 ? (d ? PM_FFT_NT_YF / 2 : PM_FFT_NT_YF)
         : slot == 2 ? PM_FFT_NT_X : slot == 3 ? (d ? PM_FFT_NT_YI / 2 : PM_FFT_NT_YI) : PM_FFT_NT_ZI;
}
template <typename Rl> __host__ __device__ constexpr int variant_threads(int slot, int v) {
    const int d = slot_default_threads<Rl>(slot), i = v % 3;
    const int lo = d == 256 ? 512 : 256, hi = d == 1024 ? 512 : 1024;
    return i == 0 ? d : i == 1 ? lo : hi;
}
template <typename Rl> __host__ __device__ constexpr int variant_lines(int slot, int N, int v) {
    const int base = slot == 2 ? tile_lines<Rl, true>(N) : tile_lines<Rl>(N);
    const int t = v < 3 ? base : v < 6 ? base / 2 : base / 4;
    return t < 1 ? 1 : t;
}

template <typename Rl, int N, int V>
static void launch_variant(int pass, hipStream_t st, const void* in, void* out, const void* twM, const void* twN,
                           const GreenArgs& ga, double norm, bool sub1) {
    using C = cplx_t<Rl>;
    constexpr int M = N / 2;
    constexpr int TZF = variant_lines<Rl>(0, N, V), TYF = variant_lines<Rl>(1, N, V);
    constexpr int TX = variant_lines<Rl>(2, N, V), TYI = variant_lines<Rl>(3, N, V);
    constexpr int TZI = variant_lines<Rl>(4, N, V);
    constexpr int NZF = fft_threads(variant_threads<Rl>(0, V), TZF * M);
    constexpr int NYF = fft_threads(variant_threads<Rl>(1, V), TYF * N);
    constexpr int NX = fft_threads(variant_threads<Rl>(2, V), TX * N);
    constexpr int NYI = fft_threads(variant_threads<Rl>(3, V), TYI * N);
    constexpr int NZI = fft_threads(variant_threads<Rl>(4, V), TZI * M);
    switch (pass) {
        case kPassZFwd: {
            const unsigned g = (unsigned)((size_t)N * N / TZF);  // (x, y-tile)
            if (sub1)
                hipLaunchKernelGGL((k_zfwd<Rl, N, true, TZF, NZF>), dim3(g), dim3(NZF), 0, st, (const Rl*)in, (C*)out,
                                   (const C*)twM, (const C*)twN);
            else
                hipLaunchKernelGGL((k_zfwd<Rl, N, false, TZF, NZF>), dim3(g), dim3(NZF), 0, st, (const Rl*)in,
                                   (C*)out, (const C*)twM, (const C*)twN);
            break;
        }
        case kPassYFwd:
            hipLaunchKernelGGL((k_yfwd<Rl, N, TYF, NYF>), dim3((unsigned)((size_t)(M + 1) * N / TYF)), dim3(NYF), 0,
                               st, (const C*)in, (C*)out, (const C*)twN);
            break;
        case kPassXFused:
        case kPassXFwd:
        case kPassXInv: {
            const dim3 g((unsigned)((size_t)(M + 1) * N / TX));
            if (pass == kPassXFused)
                hipLaunchKernelGGL((k_x<Rl, N, kXForward | kXGreen | kXInverse, TX, NX>), g, dim3(NX), 0, st, (C*)out,
                                   (const C*)twN, ga);
            else if (pass == kPassXFwd)
                hipLaunchKernelGGL((k_x<Rl, N, kXForward, TX, NX>), g, dim3(NX), 0, st, (C*)out, (const C*)twN, ga);
            else
                hipLaunchKernelGGL((k_x<Rl, N, kXInverse, TX, NX>), g, dim3(NX), 0, st, (C*)out, (const C*)twN, ga);
            break;
        }
        case kPassYInv:
            hipLaunchKernelGGL((k_yinv<Rl, N, TYI, NYI>), dim3((unsigned)((size_t)(M + 1) * N / TYI)), dim3(NYI), 0,
                               st, (const C*)in, (C*)out, (const C*)twN);
            break;
        case kPassZInv:
            hipLaunchKernelGGL((k_zinv<Rl, N, TZI, NZI>), dim3((unsigned)((size_t)N * N / TZI)), dim3(NZI), 0, st,
                               (const C*)in, (Rl*)out, (const C*)twM, (const C*)twN, (Rl)norm);
            break;
        case kPassGreen: {
            const size_t total = (size_t)(M + 1) * N * N;
            const unsigned nb = (unsigned)((total + 255) / 256 < 65536 * 4 ? (total + 255) / 256 : 65536 * 4);
            hipLaunchKernelGGL((k_apply_green<Rl>), dim3(nb), dim3(256), 0, st, (C*)out, N, ga);
            break;
        }
        default:
            break;
    }
}

template <typename Rl, int N>
static void launch_pass(int pass, int v, hipStream_t st, const void* in, void* out, const void* twM,
                        const void* twN, const GreenArgs& ga, double norm, bool sub1) {
    if constexpr (N < kTuneMinN) {
        launch_variant<Rl, N, 0>(pass, st, in, out, twM, twN, ga, norm, sub1);
    } else {
        switch (v) {
            case 1: launch_variant<Rl, N, 1>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 2: launch_variant<Rl, N, 2>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 3: launch_variant<Rl, N, 3>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 4: launch_variant<Rl, N, 4>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 5: launch_variant<Rl, N, 5>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 6: launch_variant<Rl, N, 6>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 7: launch_variant<Rl, N, 7>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            case 8: launch_variant<Rl, N, 8>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
            default: launch_variant<Rl, N, 0>(pass, st, in, out, twM, twN, ga, norm, sub1); break;
        }
    }
}

template <typename Rl>
static int dispatch_pass(int N, int pass, int v, hipStream_t st, const void* in, void* out, const void* twM,
                         const void* twN, const GreenArgs& ga, double norm, bool sub1) {
    switch (N) {
        case 8: launch_pass<Rl, 8>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 16: launch_pass<Rl, 16>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 32: launch_pass<Rl, 32>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 64: launch_pass<Rl, 64>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 128: launch_pass<Rl, 128>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 256: launch_pass<Rl, 256>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 512: launch_pass<Rl, 512>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        case 1024: launch_pass<Rl, 1024>(pass, v, st, in, out, twM, twN, ga, norm, sub1); break;
        default: return -1;
    }
    return 0;
}

int fft_slot(int pass) {
    switch (pass) {
        case kPassZFwd: return 0;
        case kPassYFwd: return 1;
        case kPassXFused:
        case kPassXFwd:
        case kPassXInv: return 2;
        case kPassYInv: return 3;
        case kPassZInv: return 4;
        default: return -1;
    }
}

int fft_pass(int prec, int N, int pass, hipStream_t st, const void* in, void* out, const void* twM,
             const void* twN, const GreenArgs& ga, double norm, bool sub1, const FftShape* shape) {
    const int slot = fft_slot(pass);
    const int v = shape && slot >= 0 ? shape->v[slot] : 0;
    if (v < 0 || v >= kFftVariants) return -1;
    return prec == PM_FP32 ? dispatch_pass<float>(N, pass, v, st, in, out, twM, twN, ga, norm, sub1)
                           : dispatch_pass<double>(N, pass, v, st, in, out, twM, twN, ga, norm, sub1);
}

}  // namespace pm
