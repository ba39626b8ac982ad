/*
 * pcx_oracle.c -- CPU oracle for the packed FFT / overlap-save cross-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's cpu_baseline leg as the
 * checker and as the host-core baseline.  The product library never links or calls it.
 *
 * Parity pins (each function is checked by tests/test_oracle.py against something independent):
 *   orc_dft / orc_fft          numpy.fft (pocketfft, FP64), closed-form DFT of a leaky tone
 *                              (the input of tests/test_fft.cpp:299-302), delta, Parseval,
 *                              linearity, brute force vs radix-2 cross-check.
 *   orc_unordered_q            replay of the reference's lane shuffles (orc_unordered_q_network).
 *   orc_fft_unit               golden vectors under tests/golden/ (numpy.fft + closed-form maps).
 *   orc_xcorr_naive            exact integer Barker-13 case of tests/test_fxcorr.cpp:100-139.
 *   orc_fxcorr_overlap_save    orc_xcorr_naive on multi-block inputs.
 *   orc_glibc_rand             the C library's own rand() after srand(1).
 *   orc_port_*                 orc_fft_unit / orc_xcorr_naive at FP32 tolerance.
 */
#include "pcx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ================================================================ seeded inputs */

uint64_t orc_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t* s) { /* [0, 1) with 53 random bits */
    return (double)(orc_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_gen_uniform(uint64_t seed, float* re, float* im, int64_t n) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) {
        re[i] = (float)(2.0 * u01(&s) - 1.0);
        im[i] = (float)(2.0 * u01(&s) - 1.0);
    }
}

void orc_gen_normal(uint64_t seed, double sigma, float* re, float* im, int64_t n) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) {
        double u1 = 1.0 - u01(&s); /* (0, 1] */
        double u2 = u01(&s);
        double r  = sqrt(-2.0 * log(u1));
        re[i]     = (float)(sigma * r * cos(2.0 * ORC_PI * u2));
        im[i]     = (float)(sigma * r * sin(2.0 * ORC_PI * u2));
    }
}

/* glibc random() TYPE_3: r[i] = r[i-31] + r[i-3] (mod 2^32) after a 16807 LCG seeding and 310
 * discarded outputs; each output is r >> 1.  Ring buffer of 34 words indexed by i mod 34. */
void orc_glibc_srand(orc_glibc_rand_t* st, uint32_t seed) {
    int32_t r[34];
    int64_t word = seed == 0 ? 1 : seed;
    r[0]         = (int32_t)word;
    for (int i = 1; i < 31; ++i) {
        int64_t hi = word / 127773;
        int64_t lo = word % 127773;
        word       = 16807 * lo - 2836 * hi;
        if (word < 0)
            word += 2147483647;
        r[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; ++i)
        r[i] = r[i - 31];
    memcpy(st->r, r, sizeof(r));
    st->f = 34; /* index of the next value to produce */
    st->b = 0;
    for (int i = 0; i < 310; ++i)
        (void)orc_glibc_rand(st);
}

int32_t orc_glibc_rand(orc_glibc_rand_t* st) {
    int      i   = st->f;
    uint32_t a   = (uint32_t)st->r[(i - 31) % 34];
    uint32_t b   = (uint32_t)st->r[(i - 3) % 34];
    uint32_t v   = a + b;
    st->r[i % 34] = (int32_t)v;
    st->f        = (i + 1 - 34) % 34 + 34; /* keep i in [34, 68) so (i-31) stays non-negative */
    return (int32_t)(v >> 1);
}

/* ================================================================ index maps */

int64_t orc_brev(int64_t i, int bits) {
    int64_t r = 0;
    for (int b = 0; b < bits; ++b)
        r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

int64_t orc_unordered_q(int64_t i) {
    int64_t b0 = i & 1, b1 = (i >> 1) & 1, b2 = (i >> 2) & 1, b3 = (i >> 3) & 1, b4 = (i >> 4) & 1;
    return (i & ~(int64_t)31) | (b4 << 4) | (b2 << 3) | (b0 << 2) | (b1 << 1) | b3;
}

/* Replays avx2_fft.hpp:188-213.  The unordered store writes the four 8-lane registers e0..e3
 * as they are; the bit-reversed store first applies unpack_ps, unpack_pd and unpack_128 to the
 * register pairs (e0,e1) and (e2,e3) (avx2_common.hpp:113-145).  Position p of the bit-reversed
 * store holds X[brev(p)], so unordered slot i holds X[brev(T(i))], T = where the shuffles move i. */
static void shuf_ps(const int* a, const int* b, int* lo, int* hi) {
    int l[8] = {a[0], b[0], a[1], b[1], a[4], b[4], a[5], b[5]};
    int h[8] = {a[2], b[2], a[3], b[3], a[6], b[6], a[7], b[7]};
    memcpy(lo, l, sizeof(l));
    memcpy(hi, h, sizeof(h));
}
static void shuf_pd(const int* a, const int* b, int* lo, int* hi) {
    int l[8] = {a[0], a[1], b[0], b[1], a[4], a[5], b[4], b[5]};
    int h[8] = {a[2], a[3], b[2], b[3], a[6], a[7], b[6], b[7]};
    memcpy(lo, l, sizeof(l));
    memcpy(hi, h, sizeof(h));
}
static void shuf_128(const int* a, const int* b, int* lo, int* hi) {
    int l[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
    int h[8] = {a[4], a[5], a[6], a[7], b[4], b[5], b[6], b[7]};
    memcpy(lo, l, sizeof(l));
    memcpy(hi, h, sizeof(h));
}
int64_t orc_unordered_q_network(int64_t i) {
    int e[4][8], s[4][8];
    for (int r = 0; r < 4; ++r)
        for (int l = 0; l < 8; ++l)
            e[r][l] = 8 * r + l;
    shuf_ps(e[0], e[1], s[0], s[1]);
    shuf_ps(e[2], e[3], s[2], s[3]);
    shuf_pd(s[0], s[1], e[0], e[1]);
    shuf_pd(s[2], s[3], e[2], e[3]);
    shuf_128(e[0], e[1], s[0], s[1]);
    shuf_128(e[2], e[3], s[2], s[3]);
    int T[32];
    for (int r = 0; r < 4; ++r)
        for (int l = 0; l < 8; ++l)
            T[s[r][l]] = 8 * r + l;
    return (i & ~(int64_t)31) | T[i & 31];
}

static int ilog2(int64_t n) {
    int l = 0;
    while (((int64_t)1 << l) < n)
        ++l;
    return l;
}

int64_t orc_slot_freq(int order, int64_t i, int64_t n) {
    int bits = ilog2(n);
    if (order == 1)
        return orc_brev(i, bits);
    if (order == 2)
        return orc_brev(orc_unordered_q(i), bits);
    return i;
}

/* out[i] = orc_slot_freq(order, i, n) for every i < n */
void orc_slot_map(int order, int64_t n, int64_t* out) {
    for (int64_t i = 0; i < n; ++i)
        out[i] = orc_slot_freq(order, i, n);
}

int64_t orc_pidx(int64_t i, int64_t pack) { return i + (i / pack) * pack; }

int64_t orc_real_size(int64_t n, int64_t pack) { return ((n + pack - 1) / pack) * pack * 2; }

/* ================================================================ FP64 transforms */

static void trig_table(int64_t n, double* c, double* s) {
    for (int64_t m = 0; m < n; ++m) {
        /* exact values on the axes, as wnk does (fft.hpp:77-87) */
        if (4 * m == n) {
            c[m] = 0.0;
            s[m] = 1.0;
        } else if (2 * m == n) {
            c[m] = -1.0;
            s[m] = 0.0;
        } else if (4 * m == 3 * n) {
            c[m] = 0.0;
            s[m] = -1.0;
        } else {
            double a = 2.0 * ORC_PI * (double)m / (double)n;
            c[m]     = cos(a);
            s[m]     = sin(a);
        }
    }
}

void orc_dft(const double* xr, const double* xi, double* yr, double* yi, int64_t n, int inverse) {
    double* c = (double*)malloc(sizeof(double) * n);
    double* s = (double*)malloc(sizeof(double) * n);
    trig_table(n, c, s);
    double sg = inverse ? 1.0 : -1.0;
    for (int64_t k = 0; k < n; ++k) {
        double  ar = 0, ai = 0;
        int64_t m = 0; /* (j*k) mod n, updated incrementally */
        for (int64_t j = 0; j < n; ++j) {
            double wr = c[m], wi = sg * s[m];
            ar += xr[j] * wr - xi[j] * wi;
            ai += xr[j] * wi + xi[j] * wr;
            m += k;
            if (m >= n)
                m -= n;
        }
        yr[k] = ar;
        yi[k] = ai;
    }
    free(c);
    free(s);
}

void orc_fft(const double* xr, const double* xi, double* yr, double* yi, int64_t n, int inverse) {
    int bits = ilog2(n);
    double* tr = (double*)malloc(sizeof(double) * n);
    double* ti = (double*)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = orc_brev(i, bits);
        tr[r]     = xr[i];
        ti[r]     = xi[i];
    }
    double* c = (double*)malloc(sizeof(double) * n);
    double* s = (double*)malloc(sizeof(double) * n);
    trig_table(n, c, s);
    double sg = inverse ? 1.0 : -1.0;
    for (int64_t len = 2; len <= n; len <<= 1) {
        int64_t half = len / 2, stride = n / len;
        for (int64_t g = 0; g < n; g += len) {
            for (int64_t j = 0; j < half; ++j) {
                double  wr = c[j * stride], wi = sg * s[j * stride];
                int64_t a = g + j, b = a + half;
                double  br = tr[b] * wr - ti[b] * wi;
                double  bi = tr[b] * wi + ti[b] * wr;
                tr[b]      = tr[a] - br;
                ti[b]      = ti[a] - bi;
                tr[a] += br;
                ti[a] += bi;
            }
        }
    }
    memcpy(yr, tr, sizeof(double) * n);
    memcpy(yi, ti, sizeof(double) * n);
    free(tr);
    free(ti);
    free(c);
    free(s);
}

void orc_fft_unit(int order, int inverse, int normalized, int64_t n, float* dst, int64_t dst_pack,
                  const float* src, int64_t src_pack, int64_t src_size, double* out_re, double* out_im) {
    double* ar = (double*)calloc(n, sizeof(double));
    double* ai = (double*)calloc(n, sizeof(double));
    double* yr = (double*)malloc(sizeof(double) * n);
    double* yi = (double*)malloc(sizeof(double) * n);
    if (!inverse) {
        for (int64_t i = 0; i < n && i < src_size; ++i) {
            int64_t o = orc_pidx(i, src_pack);
            ar[i]     = src[o];
            ai[i]     = src[o + src_pack];
        }
        orc_fft(ar, ai, yr, yi, n, 0);
        int64_t dp = (order == 2 && dst_pack < 8) ? 8 : dst_pack;
        for (int64_t i = 0; i < n; ++i) {
            int64_t f = orc_slot_freq(order, i, n);
            int64_t o = orc_pidx(i, dp);
            if (dst) {
                dst[o]      = (float)yr[f];
                dst[o + dp] = (float)yi[f];
            }
            if (out_re)
                out_re[i] = yr[f];
            if (out_im)
                out_im[i] = yi[f];
        }
    } else {
        int64_t sp = (order == 2 && src_pack < 8) ? 8 : src_pack;
        for (int64_t i = 0; i < n; ++i) {
            int64_t f = orc_slot_freq(order, i, n);
            int64_t o = orc_pidx(i, sp);
            ar[f]     = src[o];
            ai[f]     = src[o + sp];
        }
        orc_fft(ar, ai, yr, yi, n, 1);
        double sc = normalized ? 1.0 / (double)n : 1.0;
        for (int64_t i = 0; i < n; ++i) {
            int64_t o = orc_pidx(i, dst_pack);
            if (dst) {
                dst[o]            = (float)(yr[i] * sc);
                dst[o + dst_pack] = (float)(yi[i] * sc);
            }
            if (out_re)
                out_re[i] = yr[i] * sc;
            if (out_im)
                out_im[i] = yi[i] * sc;
        }
    }
    free(ar);
    free(ai);
    free(yr);
    free(yi);
}

void orc_xcorr_naive(const double* fr, const double* fi, int64_t L, const double* gr, const double* gi,
                     int64_t M, double* outr, double* outi) {
    for (int64_t i = 0; i < L; ++i) {
        double  sr = 0, si = 0;
        int64_t mmax = (L - i < M) ? L - i : M;
        for (int64_t m = 0; m < mmax; ++m) {
            /* f * conj(g) */
            sr += fr[i + m] * gr[m] + fi[i + m] * gi[m];
            si += fi[i + m] * gr[m] - fr[i + m] * gi[m];
        }
        outr[i] = sr;
        outi[i] = si;
    }
}

int64_t orc_fxcorr_overlap_save(const double* fr, const double* fi, int64_t L, const double* gr,
                                const double* gi, int64_t M, int64_t N, int64_t pack, double* outr,
                                double* outi) {
    int64_t overlap = M - 1;
    if (N <= overlap || M > N)
        return -1;
    int64_t step = (N - overlap) / pack * pack;
    if (step <= 0)
        return -1;
    double* kr = (double*)calloc(N, sizeof(double));
    double* ki = (double*)calloc(N, sizeof(double));
    double* Kr = (double*)malloc(sizeof(double) * N);
    double* Ki = (double*)malloc(sizeof(double) * N);
    double* ar = (double*)malloc(sizeof(double) * N);
    double* ai = (double*)malloc(sizeof(double) * N);
    double* Ar = (double*)malloc(sizeof(double) * N);
    double* Ai = (double*)malloc(sizeof(double) * N);
    memcpy(kr, gr, sizeof(double) * M);
    memcpy(ki, gi, sizeof(double) * M);
    orc_fft(kr, ki, Kr, Ki, N, 0);
    for (int64_t k = 0; k < N; ++k) { /* fxcorr.hpp:68: kernel / N */
        Kr[k] /= (double)N;
        Ki[k] /= (double)N;
    }
    int64_t offset = 0, blocks = 0;
    for (;;) {
        int64_t len;
        if (offset + N < L)
            len = N;
        else if (offset < L)
            len = L - offset;
        else
            break;
        for (int64_t j = 0; j < N; ++j) {
            ar[j] = j < len ? fr[offset + j] : 0.0;
            ai[j] = j < len ? fi[offset + j] : 0.0;
        }
        orc_fft(ar, ai, Ar, Ai, N, 0);
        for (int64_t k = 0; k < N; ++k) { /* A * conj(K) */
            double r = Ar[k] * Kr[k] + Ai[k] * Ki[k];
            double i = Ai[k] * Kr[k] - Ar[k] * Ki[k];
            Ar[k]    = r;
            Ai[k]    = i;
        }
        orc_fft(Ar, Ai, ar, ai, N, 1); /* unnormalised inverse */
        int64_t keep = (len == N && offset + N < L) ? step : (step < L - offset ? step : L - offset);
        for (int64_t j = 0; j < keep; ++j) {
            outr[offset + j] = ar[j];
            outi[offset + j] = ai[j];
        }
        offset += keep;
        ++blocks;
    }
    free(kr);
    free(ki);
    free(Kr);
    free(Ki);
    free(ar);
    free(ai);
    free(Ar);
    free(Ai);
    return blocks;
}

/* ================================================================ FP32 host-core port
 *
 * Split-complex radix-4 decimation-in-frequency FFT, in place, natural input and bit-reversed
 * output, with one radix-2 stage first when log2(n) is odd.  This is the algorithm class of the
 * reference's bit_reversed/unordered path (fft.hpp:1160-1279 DIF radix-4 nodes, then the order
 * applied on the final store, avx2_fft.hpp:114-217).  Orders are applied as permutations of the
 * bit-reversed result; the inverse runs the stages backwards with conjugate twiddles
 * (fft.hpp:1281-1424) and scales once at the end.
 */
struct orc_port_plan_s {
    int64_t n;
    int     bits;
    int     nstages;
    int64_t span[32];  /* span m of each stage */
    int     radix[32]; /* 2 or 4 */
    float*  tw[32];    /* per stage: w1r,w1i,w2r,w2i,w3r,w3i blocks of m floats */
};

orc_port_plan* orc_port_plan_create(int64_t n) {
    if (n < 2 || (n & (n - 1)))
        return NULL;
    orc_port_plan* p = (orc_port_plan*)calloc(1, sizeof(*p));
    p->n             = n;
    p->bits          = ilog2(n);
    int64_t rem      = n;
    int     st       = 0;
    if (p->bits & 1) {
        p->span[st]  = n / 2;
        p->radix[st] = 2;
        rem          = n / 2;
        ++st;
    }
    while (rem >= 4) {
        p->span[st]  = rem / 4;
        p->radix[st] = 4;
        rem /= 4;
        ++st;
    }
    p->nstages = st;
    for (int s = 0; s < st; ++s) {
        int64_t m   = p->span[s];
        int64_t len = m * p->radix[s];
        float*  t   = (float*)malloc(sizeof(float) * 6 * (m > 0 ? m : 1));
        for (int64_t j = 0; j < m; ++j) {
            for (int q = 1; q <= 3; ++q) {
                double a               = -2.0 * ORC_PI * (double)(q * j) / (double)len;
                t[(2 * q - 2) * m + j] = (float)cos(a);
                t[(2 * q - 1) * m + j] = (float)sin(a);
            }
        }
        p->tw[s] = t;
    }
    return p;
}

void orc_port_plan_destroy(orc_port_plan* p) {
    if (!p)
        return;
    for (int s = 0; s < p->nstages; ++s)
        free(p->tw[s]);
    free(p);
}

static void port_dif(const orc_port_plan* p, float* restrict re, float* restrict im) {
    int64_t n = p->n;
    for (int s = 0; s < p->nstages; ++s) {
        int64_t      m  = p->span[s];
        const float* t  = p->tw[s];
        const float *w1r = t, *w1i = t + m, *w2r = t + 2 * m, *w2i = t + 3 * m, *w3r = t + 4 * m,
                    *w3i = t + 5 * m;
        if (p->radix[s] == 2) {
            for (int64_t g = 0; g < n; g += 2 * m) {
                float *xr = re + g, *xi = im + g;
#pragma omp simd
                for (int64_t j = 0; j < m; ++j) {
                    float ar = xr[j], ai = xi[j], br = xr[j + m], bi = xi[j + m];
                    float dr = ar - br, di = ai - bi;
                    xr[j]     = ar + br;
                    xi[j]     = ai + bi;
                    xr[j + m] = dr * w1r[j] - di * w1i[j];
                    xi[j + m] = dr * w1i[j] + di * w1r[j];
                }
            }
            continue;
        }
        for (int64_t g = 0; g < n; g += 4 * m) {
            float *xr = re + g, *xi = im + g;
#pragma omp simd
            for (int64_t j = 0; j < m; ++j) {
                float a0r = xr[j], a0i = xi[j], a1r = xr[j + m], a1i = xi[j + m];
                float a2r = xr[j + 2 * m], a2i = xi[j + 2 * m], a3r = xr[j + 3 * m], a3i = xi[j + 3 * m];
                float t0r = a0r + a2r, t0i = a0i + a2i, t1r = a0r - a2r, t1i = a0i - a2i;
                float t2r = a1r + a3r, t2i = a1i + a3i, t3r = a1r - a3r, t3i = a1i - a3i;
                /* (-i)*t3 = (t3i, -t3r) */
                float y0r = t0r + t2r, y0i = t0i + t2i;
                float y2r = t0r - t2r, y2i = t0i - t2i;
                float y1r = t1r + t3i, y1i = t1i - t3r;
                float y3r = t1r - t3i, y3i = t1i + t3r;
                xr[j]         = y0r;
                xi[j]         = y0i;
                xr[j + m]     = y2r * w2r[j] - y2i * w2i[j];
                xi[j + m]     = y2r * w2i[j] + y2i * w2r[j];
                xr[j + 2 * m] = y1r * w1r[j] - y1i * w1i[j];
                xi[j + 2 * m] = y1r * w1i[j] + y1i * w1r[j];
                xr[j + 3 * m] = y3r * w3r[j] - y3i * w3i[j];
                xi[j + 3 * m] = y3r * w3i[j] + y3i * w3r[j];
            }
        }
    }
}

static void port_dit_inverse(const orc_port_plan* p, float* restrict re, float* restrict im) {
    int64_t n = p->n;
    for (int s = p->nstages - 1; s >= 0; --s) {
        int64_t      m  = p->span[s];
        const float* t  = p->tw[s];
        const float *w1r = t, *w1i = t + m, *w2r = t + 2 * m, *w2i = t + 3 * m, *w3r = t + 4 * m,
                    *w3i = t + 5 * m;
        if (p->radix[s] == 2) {
            for (int64_t g = 0; g < n; g += 2 * m) {
                float *xr = re + g, *xi = im + g;
#pragma omp simd
                for (int64_t j = 0; j < m; ++j) {
                    float ar = xr[j], ai = xi[j];
                    float br = xr[j + m] * w1r[j] + xi[j + m] * w1i[j]; /* * conj(w) */
                    float bi = xi[j + m] * w1r[j] - xr[j + m] * w1i[j];
                    xr[j]     = ar + br;
                    xi[j]     = ai + bi;
                    xr[j + m] = ar - br;
                    xi[j + m] = ai - bi;
                }
            }
            continue;
        }
        for (int64_t g = 0; g < n; g += 4 * m) {
            float *xr = re + g, *xi = im + g;
#pragma omp simd
            for (int64_t j = 0; j < m; ++j) {
                float y0r = xr[j], y0i = xi[j];
                float y2r = xr[j + m] * w2r[j] + xi[j + m] * w2i[j];
                float y2i = xi[j + m] * w2r[j] - xr[j + m] * w2i[j];
                float y1r = xr[j + 2 * m] * w1r[j] + xi[j + 2 * m] * w1i[j];
                float y1i = xi[j + 2 * m] * w1r[j] - xr[j + 2 * m] * w1i[j];
                float y3r = xr[j + 3 * m] * w3r[j] + xi[j + 3 * m] * w3i[j];
                float y3i = xi[j + 3 * m] * w3r[j] - xr[j + 3 * m] * w3i[j];
                float t0r = y0r + y2r, t0i = y0i + y2i, t1r = y0r - y2r, t1i = y0i - y2i;
                float t2r = y1r + y3r, t2i = y1i + y3i, t3r = y1r - y3r, t3i = y1i - y3i;
                /* (+i)*t3 = (-t3i, t3r) */
                xr[j]         = t0r + t2r;
                xi[j]         = t0i + t2i;
                xr[j + 2 * m] = t0r - t2r;
                xi[j + 2 * m] = t0i - t2i;
                xr[j + m]     = t1r - t3i;
                xi[j + m]     = t1i + t3r;
                xr[j + 3 * m] = t1r + t3i;
                xi[j + 3 * m] = t1i - t3r;
            }
        }
    }
}

static void port_brev_permute(float* re, float* im, int64_t n, int bits) {
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = orc_brev(i, bits);
        if (i < r) {
            float t = re[i];
            re[i]   = re[r];
            re[r]   = t;
            t       = im[i];
            im[i]   = im[r];
            im[r]   = t;
        }
    }
}

/* out[i] = in[q(i)] within each 32-block (forward) or its inverse (in[i] -> out[q(i)]). */
static void port_unordered_permute(float* re, float* im, int64_t n, int to_unordered) {
    float tr[32], ti[32];
    for (int64_t g = 0; g < n; g += 32) {
        memcpy(tr, re + g, sizeof(tr));
        memcpy(ti, im + g, sizeof(ti));
        for (int i = 0; i < 32; ++i) {
            int q = (int)orc_unordered_q(i);
            if (to_unordered) {
                re[g + i] = tr[q];
                im[g + i] = ti[q];
            } else {
                re[g + q] = tr[i];
                im[g + q] = ti[i];
            }
        }
    }
}

void orc_port_fft(const orc_port_plan* p, int order, int inverse, int normalized, float* re, float* im) {
    int64_t n = p->n;
    if (!inverse) {
        port_dif(p, re, im);
        if (order == 0)
            port_brev_permute(re, im, n, p->bits);
        else if (order == 2)
            port_unordered_permute(re, im, n, 1);
        return;
    }
    if (order == 0)
        port_brev_permute(re, im, n, p->bits);
    else if (order == 2)
        port_unordered_permute(re, im, n, 0);
    port_dit_inverse(p, re, im);
    if (normalized) {
        float sc = (float)(1.0 / (double)n);
        for (int64_t i = 0; i < n; ++i) {
            re[i] *= sc;
            im[i] *= sc;
        }
    }
}

void orc_port_fft_batch(const orc_port_plan* p, int order, int inverse, int normalized, float* re,
                        float* im, int64_t dist, int64_t batch, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t b = 0; b < batch; ++b)
        orc_port_fft(p, order, inverse, normalized, re + b * dist, im + b * dist);
}

int64_t orc_port_fxcorr(const float* fr, const float* fi, int64_t L, const float* gr, const float* gi,
                        int64_t M, int64_t N, float* outr, float* outi, int nthreads) {
    int64_t overlap = M - 1;
    if (N <= overlap || M > N)
        return -1;
    int64_t step = (N - overlap) / 8 * 8;
    if (step <= 0)
        return -1;
    orc_port_plan* p = orc_port_plan_create(N);
    if (!p)
        return -1;
    float* Kr = (float*)calloc(N, sizeof(float));
    float* Ki = (float*)calloc(N, sizeof(float));
    memcpy(Kr, gr, sizeof(float) * M);
    memcpy(Ki, gi, sizeof(float) * M);
    port_dif(p, Kr, Ki); /* bit-reversed spectrum, as the unordered-family unit keeps it */
    float sc = (float)(1.0 / (double)N);
    for (int64_t k = 0; k < N; ++k) {
        Kr[k] *= sc;
        Ki[k] *= sc;
    }
    /* block b starts at b*step while b*step + N < L; the rest are zero-extended tail blocks */
    int64_t full = 0;
    while (full * step + N < L)
        ++full;
    int64_t tail_start = full * step;
    int64_t tails      = tail_start < L ? (L - tail_start + step - 1) / step : 0;
    int64_t blocks     = full + tails;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        float* ar = (float*)malloc(sizeof(float) * N);
        float* ai = (float*)malloc(sizeof(float) * N);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < blocks; ++b) {
            int64_t off = b * step;
            int64_t len = (off + N <= L) ? N : L - off;
            if (len < 0)
                len = 0;
            memcpy(ar, fr + off, sizeof(float) * len);
            memcpy(ai, fi + off, sizeof(float) * len);
            memset(ar + len, 0, sizeof(float) * (N - len));
            memset(ai + len, 0, sizeof(float) * (N - len));
            port_dif(p, ar, ai);
            for (int64_t k = 0; k < N; ++k) {
                float r = ar[k] * Kr[k] + ai[k] * Ki[k];
                float i = ai[k] * Kr[k] - ar[k] * Ki[k];
                ar[k]   = r;
                ai[k]   = i;
            }
            port_dit_inverse(p, ar, ai);
            int64_t keep = step < L - off ? step : L - off;
            memcpy(outr + off, ar, sizeof(float) * keep);
            memcpy(outi + off, ai, sizeof(float) * keep);
        }
        free(ar);
        free(ai);
    }
    free(Kr);
    free(Ki);
    orc_port_plan_destroy(p);
    return blocks;
}
