/* FP64 CPU oracle for the PM gravity step -- TEST INFRASTRUCTURE ONLY.
 * See pm_oracle.h for the conventions.  Every function cites the reference
 * file:line whose semantics it restates.  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark load this library. */
#define _GNU_SOURCE
#include "pm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Constants of utils_inline.h:13-18. */
static const double kOmega0 = 1.0;
static const double kOmegaM = 0.3;
static const double kOmegaL = 0.7;

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

/* ======================================================================
 * mt19937_64 (Matsumoto & Nishimura, "mt19937-64.c", 2004) and libstdc++'s
 * generate_canonical<double,53> + uniform_real_distribution, as used by
 * init_particles (utils.cpp:22-36).
 * ==================================================================== */
#define MT_NN 312
#define MT_MM 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

typedef struct {
    uint64_t mt[MT_NN];
    int mti;
} mt64_t;

static void mt64_seed(mt64_t* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < MT_NN; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = MT_NN;
}

static uint64_t mt64_next(mt64_t* s) {
    static const uint64_t mag01[2] = {0ULL, MT_MATRIX_A};
    uint64_t x;
    if (s->mti >= MT_NN) {
        int i;
        for (i = 0; i < MT_NN - MT_MM; ++i) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + MT_MM] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        for (; i < MT_NN - 1; ++i) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        x = (s->mt[MT_NN - 1] & MT_UM) | (s->mt[0] & MT_LM);
        s->mt[MT_NN - 1] = s->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        s->mti = 0;
    }
    x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

void orc_mt_raw(uint64_t seed, int64_t count, uint64_t* out) {
    mt64_t s;
    mt64_seed(&s, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = mt64_next(&s);
}

void orc_mt_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
    mt64_t s;
    mt64_seed(&s, seed);
    const double range = hi - lo;
    for (int64_t i = 0; i < n; ++i) {
        /* one 64-bit draw covers 53 bits: sum/2^64, clamped below 1 */
        double u = (double)mt64_next(&s) / 18446744073709551616.0;
        if (u >= 1.0) u = nextafter(1.0, 0.0);
        out[i] = u * range + lo;
    }
}

void orc_init_particles(int64_t np, int ng, double* posx, double* posy, double* posz) {
    /* utils.cpp:23-35 with rank 0, x0=y0=z0=0 and n*_loc = ng */
    orc_mt_uniform(0ULL, np, 0.0, (double)ng, posx);
    orc_mt_uniform(1ULL, np, 0.0, (double)ng, posy);
    orc_mt_uniform(2ULL, np, 0.0, (double)ng, posz);
}

double orc_function_a(double a) {
    /* utils_inline.h:51-53 */
    const double ok = 1.0 - kOmegaM - kOmegaL;
    return 1.0 / sqrt(kOmegaM / a + ok + kOmegaL * a * a);
}

/* ======================================================================
 * Reference-layout operators.
 * ==================================================================== */
void orc_init_cube(int ng, double* cube) {
    /* utils.cpp:38-68 for the first thread grid: -1 on [0,n-1)^3, 0 on the
     * + ghost planes. */
    const int n = ng + 1;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const int owned = (i < ng) && (j < ng) && (k < ng);
                cube[((size_t)i * n + j) * n + k] = owned ? -1.0 : 0.0;
            }
}

static inline void cic_add(double* a, size_t base, size_t ys, size_t xs, double tx, double ty,
                           double tz, double m) {
    /* utils_inline.h:25-49: same products, same order */
    const double sx0 = 1.0 - tx, sx1 = tx;
    const double sy0 = 1.0 - ty, sy1 = ty;
    const double sz0 = 1.0 - tz, sz1 = tz;
    const double sxy00 = sx0 * sy0, sxy10 = sx1 * sy0, sxy01 = sx0 * sy1, sxy11 = sx1 * sy1;
    const double m0 = m * sz0, m1 = m * sz1;
    double* p = a + base;
    p[0] += sxy00 * m0;
    p[xs] += sxy10 * m0;
    p[ys] += sxy01 * m0;
    p[xs + ys] += sxy11 * m0;
    p[1] += sxy00 * m1;
    p[xs + 1] += sxy10 * m1;
    p[ys + 1] += sxy01 * m1;
    p[xs + ys + 1] += sxy11 * m1;
}

void orc_deposit_cube(int64_t np, const double* posx, const double* posy, const double* posz,
                      double mass, int ng, double* cube) {
    /* utils.cpp:70-102 (serial: one private grid) */
    const size_t n = (size_t)ng + 1;
    for (int64_t p = 0; p < np; ++p) {
        const int ix = (int)posx[p], iy = (int)posy[p], iz = (int)posz[p];
        const size_t base = ((size_t)ix * n + (size_t)iy) * n + (size_t)iz;
        cic_add(cube, base, n, n * n, posx[p] - (double)ix, posy[p] - (double)iy,
                posz[p] - (double)iz, mass);
    }
}

void orc_ghost_fold(int ng, const double* cube, double* slab) {
    /* reorder_slab.cpp:15-22 (+ the one-box transpose): every cube node
     * (i,j,k), i,j,k in [0,ng], lands on (i%ng, j%ng, k%ng); duplicates are
     * summed (reorder_slab.cpp:43-67).  Images are added in increasing cube
     * index; the reference's sorted gather may order ties differently, which
     * changes the sum only by rounding. */
    const size_t n = (size_t)ng + 1;
    memset(slab, 0, sizeof(double) * (size_t)ng * ng * ng);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
            for (size_t k = 0; k < n; ++k) {
                const size_t d = ((i % ng) * ng + (j % ng)) * ng + (k % ng);
                slab[d] += cube[(i * n + j) * n + k];
            }
}

void orc_halo_expand(int ng, const double* slab, double* halo) {
    /* reorder_slab_back.cpp:11-21 (y/z halo: index -1 + Ng wrap) and
     * transpose_slab_back.cpp:22-28 (x halo: plane -1 from the last slab,
     * planes ng, ng+1 from the first). */
    const int m = ng + 3;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            for (int k = 0; k < m; ++k) {
                const int si = (i - 1 + ng) % ng, sj = (j - 1 + ng) % ng, sk = (k - 1 + ng) % ng;
                halo[((size_t)i * m + j) * m + k] = slab[((size_t)si * ng + sj) * ng + sk];
            }
}

void orc_centered_derivatives(int ng, const double* in, double* dx, double* dy, double* dz) {
    /* utils.cpp:115-137 with nx1 = ny1 = nz1 = ng+3 */
    const int n1 = ng + 3, n = ng + 1;
    const size_t yz1 = (size_t)n1 * n1;
    for (int i = 1; i < n1 - 1; ++i)
        for (int j = 1; j < n1 - 1; ++j) {
            const double* row_i = in + i * yz1 + (size_t)j * n1;
            const double* row_im1 = in + (i - 1) * yz1 + (size_t)j * n1;
            const double* row_ip1 = in + (i + 1) * yz1 + (size_t)j * n1;
            const double* row_jm1 = in + i * yz1 + (size_t)(j - 1) * n1;
            const double* row_jp1 = in + i * yz1 + (size_t)(j + 1) * n1;
            const size_t o = ((size_t)(i - 1) * n + (j - 1)) * n;
            for (int kk = 0; kk < n; ++kk) {
                const int k = kk + 1;
                dx[o + kk] = (row_im1[k] - row_ip1[k]) * 0.5;
                dy[o + kk] = (row_jm1[k] - row_jp1[k]) * 0.5;
                dz[o + kk] = (row_i[k - 1] - row_i[k + 1]) * 0.5;
            }
        }
}

static inline double trilerp(const double* f, size_t b, size_t xs, size_t ys, double tx, double ty,
                             double tz) {
    /* utils_inline.h:55-70: same products, same order */
    const double mx = 1.0 - tx, my = 1.0 - ty, mz = 1.0 - tz;
    const double f00 = f[b] * mx + f[b + xs] * tx;
    const double f10 = f[b + ys] * mx + f[b + xs + ys] * tx;
    const double fz0 = f00 * my + f10 * ty;
    const double g00 = f[b + 1] * mx + f[b + xs + 1] * tx;
    const double g10 = f[b + ys + 1] * mx + f[b + xs + ys + 1] * tx;
    const double fz1 = g00 * my + g10 * ty;
    return fz0 * mz + fz1 * tz;
}

void orc_update_particles(int64_t np, const double* posx, const double* posy, const double* posz,
                          const double* dx, const double* dy, const double* dz, int ng,
                          double kick, double* px, double* py, double* pz) {
    /* utils.cpp:141-170 */
    const size_t n = (size_t)ng + 1;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        const int ix = (int)posx[p], iy = (int)posy[p], iz = (int)posz[p];
        const double tx = posx[p] - ix, ty = posy[p] - iy, tz = posz[p] - iz;
        const size_t b = ((size_t)ix * n + iy) * n + iz;
        const double gx = trilerp(dx, b, n * n, n, tx, ty, tz);
        const double gy = trilerp(dy, b, n * n, n, tx, ty, tz);
        const double gz = trilerp(dz, b, n * n, n, tx, ty, tz);
        px[p] += gx * kick;
        py[p] += gy * kick;
        pz[p] += gz * kick;
    }
}

/* ======================================================================
 * FFT: iterative radix-2 complex FFT (textbook Cooley-Tukey with bit
 * reversal) applied line by line; the 3-D real transform keeps kz in
 * [0, ng/2].  The reference uses FFTW or FFTE (pdzfft3d.f:77-265,
 * pzdfft3d.f:76-301); both are unnormalised DFTs with e^{-i} forward.
 * ==================================================================== */
static int is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

typedef struct {
    int n;
    int logn;
    double* wr; /* cos(2 pi k/n), k < n/2 */
    double* wi; /* -sin(2 pi k/n) */
    int* rev;
} fft_plan1;

static fft_plan1 plan1_make(int n) {
    fft_plan1 p;
    p.n = n;
    p.logn = 0;
    while ((1 << p.logn) < n) ++p.logn;
    p.wr = (double*)malloc(sizeof(double) * (size_t)(n / 2 > 0 ? n / 2 : 1));
    p.wi = (double*)malloc(sizeof(double) * (size_t)(n / 2 > 0 ? n / 2 : 1));
    p.rev = (int*)malloc(sizeof(int) * (size_t)n);
    for (int k = 0; k < n / 2; ++k) {
        const double th = 2.0 * M_PI * (double)k / (double)n;
        p.wr[k] = cos(th);
        p.wi[k] = -sin(th);
    }
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < p.logn; ++b)
            if (i & (1 << b)) r |= 1 << (p.logn - 1 - b);
        p.rev[i] = r;
    }
    return p;
}

static void plan1_free(fft_plan1* p) {
    free(p->wr);
    free(p->wi);
    free(p->rev);
}

/* In-place complex FFT of re/im (length n); sign = -1 forward, +1 inverse. */
static void fft1(const fft_plan1* p, double* re, double* im, int sign) {
    const int n = p->n;
    for (int i = 0; i < n; ++i) {
        const int r = p->rev[i];
        if (r > i) {
            double t = re[i];
            re[i] = re[r];
            re[r] = t;
            t = im[i];
            im[i] = im[r];
            im[r] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; ++k) {
                const double wr = p->wr[k * step];
                const double wi = sign < 0 ? p->wi[k * step] : -p->wi[k * step];
                const int a = s + k, b = a + half;
                const double xr = re[b] * wr - im[b] * wi;
                const double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
}

/* Complex FFT along x or y of the half spectrum spec[ng][ng][nh] (in place). */
static void fft_axis_xy(const fft_plan1* p, double* spec, int ng, int nh, int axis, int sign) {
    const size_t sx = (size_t)ng * nh * 2, sy = (size_t)nh * 2;
#pragma omp parallel
    {
        double* re = (double*)malloc(sizeof(double) * (size_t)ng);
        double* im = (double*)malloc(sizeof(double) * (size_t)ng);
#pragma omp for collapse(2) schedule(static)
        for (int a = 0; a < ng; ++a)
            for (int k = 0; k < nh; ++k) {
                size_t base, stride;
                if (axis == 0) { /* lines along x at fixed (y=a, kz=k) */
                    base = (size_t)a * sy + (size_t)k * 2;
                    stride = sx;
                } else { /* lines along y at fixed (x=a, kz=k) */
                    base = (size_t)a * sx + (size_t)k * 2;
                    stride = sy;
                }
                for (int t = 0; t < ng; ++t) {
                    re[t] = spec[base + t * stride];
                    im[t] = spec[base + t * stride + 1];
                }
                fft1(p, re, im, sign);
                for (int t = 0; t < ng; ++t) {
                    spec[base + t * stride] = re[t];
                    spec[base + t * stride + 1] = im[t];
                }
            }
        free(re);
        free(im);
    }
}

int orc_fft_r2c(int ng, const double* in, double* spec) {
    if (!is_pow2(ng)) return -1;
    const int nh = ng / 2 + 1;
    fft_plan1 p = plan1_make(ng);
#pragma omp parallel
    {
        double* re = (double*)malloc(sizeof(double) * (size_t)ng);
        double* im = (double*)malloc(sizeof(double) * (size_t)ng);
#pragma omp for collapse(2) schedule(static)
        for (int x = 0; x < ng; ++x)
            for (int y = 0; y < ng; ++y) {
                const double* row = in + ((size_t)x * ng + y) * ng;
                for (int t = 0; t < ng; ++t) {
                    re[t] = row[t];
                    im[t] = 0.0;
                }
                fft1(&p, re, im, -1);
                double* o = spec + ((size_t)x * ng + y) * nh * 2;
                for (int k = 0; k < nh; ++k) {
                    o[2 * k] = re[k];
                    o[2 * k + 1] = im[k];
                }
            }
        free(re);
        free(im);
    }
    fft_axis_xy(&p, spec, ng, nh, 1, -1);
    fft_axis_xy(&p, spec, ng, nh, 0, -1);
    plan1_free(&p);
    return 0;
}

int orc_fft_c2r(int ng, const double* spec_in, double* out) {
    if (!is_pow2(ng)) return -1;
    const int nh = ng / 2 + 1;
    const size_t nspec = (size_t)ng * ng * nh * 2;
    double* spec = (double*)malloc(sizeof(double) * nspec);
    if (!spec) return -1;
    memcpy(spec, spec_in, sizeof(double) * nspec);
    fft_plan1 p = plan1_make(ng);
    fft_axis_xy(&p, spec, ng, nh, 0, +1);
    fft_axis_xy(&p, spec, ng, nh, 1, +1);
#pragma omp parallel
    {
        double* re = (double*)malloc(sizeof(double) * (size_t)ng);
        double* im = (double*)malloc(sizeof(double) * (size_t)ng);
#pragma omp for collapse(2) schedule(static)
        for (int x = 0; x < ng; ++x)
            for (int y = 0; y < ng; ++y) {
                const double* s = spec + ((size_t)x * ng + y) * nh * 2;
                for (int k = 0; k < nh; ++k) {
                    re[k] = s[2 * k];
                    im[k] = s[2 * k + 1];
                }
                for (int k = 1; k < ng / 2; ++k) { /* Hermitian extension */
                    re[ng - k] = s[2 * k];
                    im[ng - k] = -s[2 * k + 1];
                }
                fft1(&p, re, im, +1);
                double* row = out + ((size_t)x * ng + y) * ng;
                for (int t = 0; t < ng; ++t) row[t] = re[t];
            }
        free(re);
        free(im);
    }
    plan1_free(&p);
    free(spec);
    return 0;
}

double orc_green(int mode, int ng, int i, int j, int k) {
    if (i == 0 && j == 0 && k == 0) return 0.0;
    if (mode == ORC_GREEN_REF_KH) {
        /* ffte1/fft.cpp:29-33 (fftw/fft.cpp:51-55 without the 1/N^3) */
        const double pi_n = 2.0 * M_PI / ng;
        return -3 * kOmega0 / (4.0 * (3 - cos(pi_n * i) - cos(pi_n * j) - cos(pi_n * k)));
    }
    const int si = i <= ng / 2 ? i : i - ng;
    const int sj = j <= ng / 2 ? j : j - ng;
    const int sk = k <= ng / 2 ? k : k - ng;
    const double n2 = (double)si * si + (double)sj * sj + (double)sk * sk;
    return -((double)ng * ng) / (M_PI * n2);
}

int orc_solve(int ng, int mode, double scale, const double* delta, double* phi) {
    if (!is_pow2(ng)) return -1;
    const int nh = ng / 2 + 1;
    const size_t nspec = (size_t)ng * ng * nh;
    double* spec = (double*)malloc(sizeof(double) * 2 * nspec);
    if (!spec) return -1;
    orc_fft_r2c(ng, delta, spec);
    const double inv_n3 = 1.0 / ((double)ng * ng * ng);
#pragma omp parallel for collapse(2) schedule(static)
    for (int x = 0; x < ng; ++x)
        for (int y = 0; y < ng; ++y)
            for (int k = 0; k < nh; ++k) {
                /* apply_green: fftw/fft.cpp:75-82, ffte1/fft.cpp:46-52 */
                const double g = orc_green(mode, ng, x, y, k) * scale * inv_n3;
                const size_t o = (((size_t)x * ng + y) * nh + k) * 2;
                spec[o] *= g;
                spec[o + 1] *= g;
            }
    orc_fft_c2r(ng, spec, phi);
    free(spec);
    return 0;
}

/* ======================================================================
 * Periodic one-box step.
 * ==================================================================== */
/* Thread-private grids of the periodic deposit, kept between calls as the
 * reference keeps its real_buf (main.cpp:42). */
static double* g_priv = NULL;
static size_t g_priv_len = 0;

void orc_deposit_periodic(int64_t np, const double* posx, const double* posy, const double* posz,
                          double mass, int ng, double* delta) {
    /* init_buf + deposit_particles + reduce_private_to_shared + ghost fold
     * (utils.cpp:38-113, reorder_slab.cpp:43-67): each node index is taken
     * mod ng instead of landing on a ghost plane; one private grid per
     * thread as in utils.cpp:74-77. */
    const size_t nc = (size_t)ng * ng * ng;
    const int nt = omp_get_max_threads();
    const size_t need = nc * (size_t)(nt > 1 ? nt - 1 : 0);
    if (need > g_priv_len) {
        free(g_priv);
        g_priv = (double*)malloc(sizeof(double) * need);
        g_priv_len = g_priv ? need : 0;
    }
    double* priv = g_priv;
#pragma omp parallel num_threads(nt)
    {
        const int tid = omp_get_thread_num();
        double* acc = tid == 0 ? delta : priv + (size_t)(tid - 1) * nc;
        for (size_t c = 0; c < nc; ++c) acc[c] = 0.0;
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t p = 0; p < np; ++p) {
            const int ix = (int)posx[p], iy = (int)posy[p], iz = (int)posz[p];
            const double tx = posx[p] - ix, ty = posy[p] - iy, tz = posz[p] - iz;
            const size_t N = (size_t)ng;
            const size_t x0 = (size_t)ix * N * N, x1 = (size_t)(ix + 1 == ng ? 0 : ix + 1) * N * N;
            const size_t y0 = (size_t)iy * N, y1 = (size_t)(iy + 1 == ng ? 0 : iy + 1) * N;
            const size_t z0 = (size_t)iz, z1 = (size_t)(iz + 1 == ng ? 0 : iz + 1);
            const double sx0 = 1.0 - tx, sx1 = tx, sy0 = 1.0 - ty, sy1 = ty;
            const double sxy00 = sx0 * sy0, sxy10 = sx1 * sy0, sxy01 = sx0 * sy1, sxy11 = sx1 * sy1;
            const double m0 = mass * (1.0 - tz), m1 = mass * tz;
            acc[x0 + y0 + z0] += sxy00 * m0;
            acc[x1 + y0 + z0] += sxy10 * m0;
            acc[x0 + y1 + z0] += sxy01 * m0;
            acc[x1 + y1 + z0] += sxy11 * m0;
            acc[x0 + y0 + z1] += sxy00 * m1;
            acc[x1 + y0 + z1] += sxy10 * m1;
            acc[x0 + y1 + z1] += sxy01 * m1;
            acc[x1 + y1 + z1] += sxy11 * m1;
        }
    }
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < nc; ++c) {
        double s = delta[c];
        for (int t = 1; t < nt; ++t) s += priv[(size_t)(t - 1) * nc + c];
        delta[c] = s - 1.0;
    }
}

/* The three centred-difference meshes (utils.cpp:115-137) on the periodic
 * mesh, kept between calls like the reference's real_buf1..3 (main.cpp:52-54). */
static double* g_grad = NULL;
static size_t g_grad_len = 0;

static const double* periodic_gradients(const double* phi, int ng) {
    const size_t N = (size_t)ng, nc = N * N * N;
    if (3 * nc > g_grad_len) {
        free(g_grad);
        g_grad = (double*)malloc(sizeof(double) * 3 * nc);
        g_grad_len = g_grad ? 3 * nc : 0;
    }
    double* gx = g_grad;
    double* gy = g_grad + nc;
    double* gz = g_grad + 2 * nc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j) {
            const size_t im = (size_t)((i + ng - 1) % ng), ip = (size_t)((i + 1) % ng);
            const size_t jm = (size_t)((j + ng - 1) % ng), jp = (size_t)((j + 1) % ng);
            const double* row = phi + ((size_t)i * N + j) * N;
            const double* rxm = phi + (im * N + j) * N;
            const double* rxp = phi + (ip * N + j) * N;
            const double* rym = phi + ((size_t)i * N + jm) * N;
            const double* ryp = phi + ((size_t)i * N + jp) * N;
            const size_t o = ((size_t)i * N + j) * N;
            for (int k = 0; k < ng; ++k) {
                const int km = k == 0 ? ng - 1 : k - 1, kp = k + 1 == ng ? 0 : k + 1;
                gx[o + k] = (rxm[k] - rxp[k]) * 0.5;
                gy[o + k] = (rym[k] - ryp[k]) * 0.5;
                gz[o + k] = (row[km] - row[kp]) * 0.5;
            }
        }
    return g_grad;
}

static inline double trilerp_periodic(const double* f, size_t x0, size_t x1, size_t y0, size_t y1, size_t z0,
                                      size_t z1, double tx, double ty, double tz) {
    /* trilerp1 (utils_inline.h:55-70) with periodic node offsets */
    const double mx = 1.0 - tx, my = 1.0 - ty, mz = 1.0 - tz;
    const double f00 = f[x0 + y0 + z0] * mx + f[x1 + y0 + z0] * tx;
    const double f10 = f[x0 + y1 + z0] * mx + f[x1 + y1 + z0] * tx;
    const double fz0 = f00 * my + f10 * ty;
    const double g00 = f[x0 + y0 + z1] * mx + f[x1 + y0 + z1] * tx;
    const double g10 = f[x0 + y1 + z1] * mx + f[x1 + y1 + z1] * tx;
    const double fz1 = g00 * my + g10 * ty;
    return fz0 * mz + fz1 * tz;
}

static inline void force_from_grads(const double* grads, int ng, double x, double y, double z, double* g) {
    const size_t N = (size_t)ng, nc = N * N * N;
    const int ix = (int)x, iy = (int)y, iz = (int)z;
    const double tx = x - ix, ty = y - iy, tz = z - iz;
    const size_t x0 = (size_t)ix * N * N, x1 = (size_t)(ix + 1 == ng ? 0 : ix + 1) * N * N;
    const size_t y0 = (size_t)iy * N, y1 = (size_t)(iy + 1 == ng ? 0 : iy + 1) * N;
    const size_t z0 = (size_t)iz, z1 = (size_t)(iz + 1 == ng ? 0 : iz + 1);
    for (int d = 0; d < 3; ++d) g[d] = trilerp_periodic(grads + d * nc, x0, x1, y0, y1, z0, z1, tx, ty, tz);
}

void orc_interp_force(int64_t np, const double* posx, const double* posy, const double* posz,
                      const double* phi, int ng, double* gx, double* gy, double* gz) {
    const double* grads = periodic_gradients(phi, ng);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        double g[3];
        force_from_grads(grads, ng, posx[p], posy[p], posz[p], g);
        gx[p] = g[0];
        gy[p] = g[1];
        gz[p] = g[2];
    }
}

static inline double wrap_pos(double x, double box) {
    if (x < 0.0) x += box;
    if (x >= box) x -= box;
    return x;
}

void orc_interp_kick_drift(int64_t np, double* posx, double* posy, double* posz, double* vx,
                           double* vy, double* vz, const double* phi, int ng, double kick,
                           double drift) {
    /* centered_derivatives + update_particles (utils.cpp:115-170) plus the
     * drift the reference leaves out (utils.cpp:145-146 commented) */
    const double box = (double)ng;
    const double* grads = periodic_gradients(phi, ng);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        double g[3];
        force_from_grads(grads, ng, posx[p], posy[p], posz[p], g);
        vx[p] += g[0] * kick;
        vy[p] += g[1] * kick;
        vz[p] += g[2] * kick;
        if (drift != 0.0) {
            posx[p] = wrap_pos(posx[p] + vx[p] * drift, box);
            posy[p] = wrap_pos(posy[p] + vy[p] * drift, box);
            posz[p] = wrap_pos(posz[p] + vz[p] * drift, box);
        }
    }
}

int orc_run_kdk(int64_t np, double* posx, double* posy, double* posz, double* vx, double* vy,
                double* vz, int ng, int nsteps, double dt, int green_mode, double green_scale) {
    if (!is_pow2(ng) || np <= 0 || nsteps < 0) return -1;
    if (nsteps == 0) return 0; /* no step: no kick either */
    const size_t nc = (size_t)ng * ng * ng;
    double* delta = (double*)malloc(sizeof(double) * nc);
    double* phi = (double*)malloc(sizeof(double) * nc);
    if (!delta || !phi) {
        free(delta);
        free(phi);
        return -1;
    }
    const double mass = (double)nc / (double)np;
    for (int s = 0; s <= nsteps; ++s) {
        orc_deposit_periodic(np, posx, posy, posz, mass, ng, delta);
        orc_solve(ng, green_mode, green_scale, delta, phi);
        const double kick = (s == 0 || s == nsteps) ? 0.5 * dt : dt;
        const double drift = (s == nsteps) ? 0.0 : dt;
        orc_interp_kick_drift(np, posx, posy, posz, vx, vy, vz, phi, ng, kick, drift);
    }
    free(delta);
    free(phi);
    return 0;
}

/* Cosmological KDK in the expansion factor a: the reference's kick
 * p += g(a) f(a) da (update_particles, utils.cpp:141-170; f(a) of
 * utils_inline.h:51-53; g(a) carries the 1/a of apply_green(a),
 * fftw/fft.cpp:75-82) with the drift it leaves commented out
 * (utils.cpp:145-146) as x += p f(a) da / a^2 at the half step.
 * The mesh solve uses green_scale with a = 1; step s of nsteps + 1 force
 * evaluations kicks by w_s da f(a_s) / a_s (w = 1/2 at both ends, 1
 * between) and drifts by da f(a_{s+1/2}) / a_{s+1/2}^2 (not after the last). */
int orc_run_cosmo(int64_t np, double* posx, double* posy, double* posz, double* px, double* py, double* pz,
                  int ng, int nsteps, double a0, double da, int green_mode, double green_scale) {
    if (!is_pow2(ng) || np <= 0 || !(a0 > 0.0)) return -1;
    const size_t nc = (size_t)ng * ng * ng;
    double* delta = (double*)malloc(sizeof(double) * nc);
    double* phi = (double*)malloc(sizeof(double) * nc);
    if (!delta || !phi) {
        free(delta);
        free(phi);
        return -1;
    }
    const double mass = (double)nc / (double)np;
    for (int s = 0; s <= nsteps; ++s) {
        const double a = a0 + s * da, ah = a + 0.5 * da;
        orc_deposit_periodic(np, posx, posy, posz, mass, ng, delta);
        orc_solve(ng, green_mode, green_scale, delta, phi);
        const double w = (s == 0 || s == nsteps) ? 0.5 : 1.0;
        const double kick = w * da * orc_function_a(a) / a;
        const double drift = (s == nsteps) ? 0.0 : da * orc_function_a(ah) / (ah * ah);
        orc_interp_kick_drift(np, posx, posy, posz, px, py, pz, phi, ng, kick, drift);
    }
    free(delta);
    free(phi);
    return 0;
}
