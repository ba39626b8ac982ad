/*
 * pcx_oracle.h -- CPU oracle for the packed FFT / overlap-save cross-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls this code;
 * only tests/, the smoke check and bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * Two parts:
 *   1. An FP64 restatement of what the reference computes (values, output orders, pack layout,
 *      normalisation, zero-extension, overlap-save correlation).  Used to check GPU results.
 *   2. An FP32 split-complex radix-4 restatement of the reference's algorithm class
 *      (DIF radix-4 with bit-reversed natural output, orders applied on store), multithreaded
 *      with OpenMP.  Used as the host-core baseline that bench.py times beside the GPU.
 *
 * Reference citations are relative to the packed source tree.
 */
#ifndef PCX_ORACLE_H
#define PCX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- seeded inputs */
/* splitmix64 step (public-domain generator by S. Vigna). */
uint64_t orc_splitmix64(uint64_t* state);
/* re/im i.i.d. uniform [-1, 1): 53-bit double from splitmix64, rounded to float. */
void orc_gen_uniform(uint64_t seed, float* re, float* im, int64_t n);
/* re/im i.i.d. normal(0, sigma^2) by Box-Muller over splitmix64 uniforms, rounded to float. */
void orc_gen_normal(uint64_t seed, double sigma, float* re, float* im, int64_t n);

/* glibc random() TYPE_3 additive-feedback generator (restated from its published description),
 * used to reproduce the inputs of the reference tests, which call std::rand() unseeded (seed 1). */
typedef struct {
    int32_t r[34];
    int32_t f, b;
} orc_glibc_rand_t;
void    orc_glibc_srand(orc_glibc_rand_t* st, uint32_t seed);
int32_t orc_glibc_rand(orc_glibc_rand_t* st);

/* ---------------------------------------------------------------- index maps */
/* fft.hpp:57-68 reverse_bit_order */
int64_t orc_brev(int64_t i, int bits);
/* unordered-order 32-point lane permutation, closed form: (b4,b3,b2,b1,b0) -> (b4,b2,b0,b1,b3) */
int64_t orc_unordered_q(int64_t i);
/* same permutation derived by replaying the register shuffles of avx2_fft.hpp:188-213 */
int64_t orc_unordered_q_network(int64_t i);
/* frequency index held by output slot i for order 0 normal / 1 bit_reversed / 2 unordered */
int64_t orc_slot_freq(int order, int64_t i, int64_t n);
void    orc_slot_map(int order, int64_t n, int64_t* out);
/* element_access.hpp:8-11 pidx: float offset of the real part of complex i in pack P */
int64_t orc_pidx(int64_t i, int64_t pack);
/* floats occupied by n complex values in pack P (vector.hpp:373-378) */
int64_t orc_real_size(int64_t n, int64_t pack);

/* ---------------------------------------------------------------- FP64 transforms */
/* Brute-force DFT, sign -1 forward / +1 inverse, unnormalised; angle reduced exactly (n*k mod N). */
void orc_dft(const double* xr, const double* xi, double* yr, double* yi, int64_t n, int inverse);
/* Iterative radix-2 FFT in FP64, natural in/out order, unnormalised. */
void orc_fft(const double* xr, const double* xi, double* yr, double* yi, int64_t n, int inverse);

/* fft_unit semantics (fft.hpp:728-915) in FP64.
 * Forward: src holds src_size <= n values in pack src_pack (zero-extended to n); dst receives the
 * spectrum in `order`, pack dst_pack (unordered spectra use pack max(dst_pack, 8)).
 * Inverse: src holds a spectrum in `order` (pack src_pack, unordered -> max(src_pack, 8)); dst
 * receives natural-order samples, scaled by 1/n when `normalized`.
 * Buffers are float arrays in packed layout; the FP64 result is also written to out_re/out_im
 * (natural index of the *output slot*, may be NULL). */
void orc_fft_unit(int order, int inverse, int normalized, int64_t n, float* dst, int64_t dst_pack,
                  const float* src, int64_t src_pack, int64_t src_size, double* out_re, double* out_im);

/* Linear cross-correlation out[i] = sum_{m<M} f[i+m] * conj(g[m]), f zero beyond L (FP64, O(L*M)). */
void orc_xcorr_naive(const double* fr, const double* fi, int64_t L, const double* gr, const double* gi,
                     int64_t M, double* outr, double* outi);
/* fxcorr_unit::operator() (fxcorr.hpp:102-136) restated in FP64 with complex-element offsets:
 * block step = floor((N-(M-1))/pack)*pack, kernel spectrum FFT(g)/N, unnormalised inverse, first
 * `step` outputs of each block kept, zero-extended tail blocks.  Returns the number of blocks, or
 * -1 when N <= M-1 (no positive step). */
int64_t orc_fxcorr_overlap_save(const double* fr, const double* fi, int64_t L, const double* gr,
                                const double* gi, int64_t M, int64_t N, int64_t pack, double* outr,
                                double* outi);

/* ---------------------------------------------------------------- FP32 host-core port */
typedef struct orc_port_plan_s orc_port_plan;
orc_port_plan* orc_port_plan_create(int64_t n);
void           orc_port_plan_destroy(orc_port_plan* p);
/* In-place split-complex FFT (re[n], im[n]) producing `order`; inverse consumes `order`. */
void orc_port_fft(const orc_port_plan* p, int order, int inverse, int normalized, float* re, float* im);
/* Batch of transforms, transform b at re + b*dist / im + b*dist, OpenMP over the batch. */
void orc_port_fft_batch(const orc_port_plan* p, int order, int inverse, int normalized, float* re,
                        float* im, int64_t dist, int64_t batch, int nthreads);
/* Overlap-save cross-correlation with the FP32 port (kernel spectrum in bit-reversed order).
 * out may not alias f.  OpenMP over blocks.  Returns blocks processed or -1. */
int64_t orc_port_fxcorr(const float* fr, const float* fi, int64_t L, const float* gr, const float* gi,
                        int64_t M, int64_t N, float* outr, float* outi, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
