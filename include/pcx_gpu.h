/*
 * pcx_gpu.h -- C ABI of the MI355X (gfx950) FFT and overlap-save cross-correlation path.
 *
 * This is the device boundary of the packed library's FFT units.  Each entry point replaces the
 * body of one reference member function; the host-side C++ classes in
 * packed_amd/include/pcx_gpu/fft.hpp keep the reference's template signatures and call these.
 * Reference citations are relative to the packed source tree.
 *
 * Data layout: the packed split-complex layout of vector.hpp.  With pack size P (a power of two),
 * the real part of complex i lives at float offset i + (i / P) * P and its imaginary part P floats
 * later (element_access.hpp:8-11).  P = 1 is std::complex<float> interleaving, P = 8 the
 * pcxo::vector<float> default, P >= N fully split re/im arrays.
 * All data pointers are device (or host-pinned) memory; `stream` is a hipStream_t (NULL = the
 * legacy default stream).  Calls are asynchronous on `stream`.
 *
 * Errors: every call returns a pcx_status; pcx_last_error() returns the message of the last
 * failure on the calling thread, worded like the reference's std::invalid_argument messages.
 */
#ifndef PCX_GPU_H
#define PCX_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PCX_SUCCESS              = 0,
    PCX_ERR_INVALID_SIZE     = 1, /* fft.hpp:1806-1812 check_size */
    PCX_ERR_INVALID_SUB_SIZE = 2, /* fft.hpp:1814-1826 check_sub_size */
    PCX_ERR_SIZE_MISMATCH    = 3, /* fft.hpp:732-738, 764-777, 832-837, 868-881 */
    PCX_ERR_INPLACE_REPACK   = 4, /* fft.hpp:785-791, 889-897 */
    PCX_ERR_INVALID_PACK     = 5, /* pack sizes are powers of two (types.hpp pack_size concept) */
    PCX_ERR_INVALID_ARGUMENT = 6, /* null handle / pointer, bad enum, negative count */
    PCX_ERR_UNSUPPORTED      = 7, /* a size or mode this build does not implement */
    PCX_ERR_HIP              = 8, /* a HIP runtime call failed */
    PCX_ERR_OUT_OF_MEMORY    = 9
} pcx_status;

/* fft.hpp:540-544 */
typedef enum { PCX_ORDER_NORMAL = 0, PCX_ORDER_BIT_REVERSED = 1, PCX_ORDER_UNORDERED = 2 } pcx_order;

/* Arithmetic of the transform: plain FP32, or split-precision MFMA stages. */
typedef enum { PCX_PREC_FP32 = 0, PCX_PREC_BF16X3 = 1, PCX_PREC_FP16X2 = 2 } pcx_precision;

typedef struct pcx_plan_s*   pcx_plan_t;
typedef struct pcx_fxcorr_s* pcx_fxcorr_t;

int         pcx_version(void);
const char* pcx_last_error(void);

/* ------------------------------------------------------------------ FFT unit
 * Replaces fft_unit<float, Order, Allocator>::fft_unit(fft_size, sub_size, allocator)
 * (fft.hpp:600-605): validates fft_size (power of two > 1) and sub_size (power of two >= 8) with
 * the reference's messages and builds the device twiddle tables.  sub_size is validated for API
 * compatibility only; the GPU schedule is chosen by the plan. */
int     pcx_plan_create(int64_t fft_size, int64_t sub_size, int order, pcx_plan_t* plan);
int     pcx_plan_destroy(pcx_plan_t plan);
/* fft.hpp:620-622 size() */
int64_t pcx_plan_size(pcx_plan_t plan);
int     pcx_plan_order(pcx_plan_t plan);
/* Arithmetic of the plan's radix-16 butterflies: PCX_PREC_FP32 (default), PCX_PREC_BF16X3
 * (matrix cores, three-term BF16 split operands, FP32 accumulation) or PCX_PREC_FP16X2 (matrix
 * cores, hi + lo FP16 operands scaled by a power of two per butterfly, FP32 accumulation); the
 * split forms cover every order, N in [2^4, 2^24] (single-precision plans with their default
 * four-step shape).  No reference counterpart: the reference
 * computes in FP32 only. */
int     pcx_plan_set_precision(pcx_plan_t plan, int precision);
int     pcx_plan_precision(pcx_plan_t plan);

/* Launch-shape tuning of four-step plans (fft_size > 16384; FP64 > 8192).  The reference's
 * sub_size / NodeSizeRec / Strategy template parameters (fft.hpp:565-598) pick its cache
 * blocking; these pick the GPU's.  None changes results.
 * pcx_plan_set_shape: one of the instantiated N1 x N2 splits and tiles: column length
 *   2^n2_log2, `cols` columns per column tile, `rows` rows per row workgroup (0 = any; the first
 *   listed match wins, the default first).  Rebuilds the split's twiddle tables after a
 *   device-wide synchronisation (every enqueued use of the old tables ends first).  Split-
 *   precision plans keep the default shape.
 * pcx_plan_shape: the current split and tiles (0s for a single-pass plan).
 * pcx_plan_set_grid: workgroup caps of the column and row passes (0 = one per tile; the kernels
 *   loop over the rest).  pcx_plan_set_chunk: bytes of the intermediate of one chunk of a batch
 *   (0 = the default 2 GiB).  pcx_plan_set_one_pass: 2^15 runs as one decimation-in-frequency
 *   pass (default on) or as the four-step.  pcx_plan_set_three_pass: natural-order transforms
 *   of 2^21 .. 2^24 points (FP64: 2^21 .. 2^23) run in three passes over the intermediate
 *   (default on) or as the four-step (the permuted orders always take the four-step). */
int pcx_plan_set_shape(pcx_plan_t plan, int n2_log2, int cols, int rows);
int pcx_plan_shape(pcx_plan_t plan, int* n2_log2, int* cols, int* rows);
int pcx_plan_set_grid(pcx_plan_t plan, int64_t grid_col, int64_t grid_row);
int pcx_plan_set_chunk(pcx_plan_t plan, int64_t bytes);
int pcx_plan_set_one_pass(pcx_plan_t plan, int enable);
int pcx_plan_set_three_pass(pcx_plan_t plan, int enable);

/* Forward FFT.  Replaces fft_raw<PD,PS>(dst, src[, src_size]) and operator()(dst, src)
 * (fft.hpp:624-687, 728-820).  In place when dst == src.  src_size <= fft_size values are read
 * and zero-extended.  The spectrum is stored in the plan's order; an unordered spectrum uses pack
 * max(dst_pack, 8) (avx2_fft.hpp:209-213). */
int pcx_fft(pcx_plan_t plan, float* dst, int64_t dst_pack, const float* src, int64_t src_pack,
            int64_t src_size, void* stream);

/* Inverse FFT.  Replaces ifft_raw<Normalized,PD,PS>(dst, src) and ifft<Normalized>(dst, src)
 * (fft.hpp:689-721, 828-915).  src holds a spectrum in the plan's order (unordered: pack
 * max(src_pack, 8), avx2_fft.hpp:227); output is natural order, scaled by 1/N if normalized. */
int pcx_ifft(pcx_plan_t plan, int normalized, float* dst, int64_t dst_pack, const float* src,
             int64_t src_pack, void* stream);

/* Batched FFT over `batch` vectors: vector b at src + b*src_dist / dst + b*dst_dist floats.
 * Batched form of the two calls above (the reference runs one fft_unit call per vector; its
 * fft_unit_par, fft.hpp:2168-2572, is not followed: its output order is wrong).
 * Four-step plans keep one chunk's intermediate in the plan (allocated on the first call that
 * needs it).  Calls on another stream than the plan's previous four-step call wait for that
 * call first, so copies of one plan may run on several streams. */
int pcx_fft_batch(pcx_plan_t plan, int inverse, int normalized, float* dst, int64_t dst_pack,
                  int64_t dst_dist, const float* src, int64_t src_pack, int64_t src_dist,
                  int64_t src_size, int64_t batch, void* stream);

/* pcx_fft_batch without any allocation, for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  No counterpart
 * in the reference (its units allocate nothing per call on the CPU).  The four-step intermediate
 * (fft_size > 16384) lives in a fixed 128 MiB device array per GPU, held until the process ends,
 * instead of the plan's own buffer; batches of any size run through it in chunks of
 * 128 MiB / (8 * fft_size) transforms.  Every FP32 plan fits (2^24 points need exactly 128 MiB);
 * a larger need fails with PCX_ERR_UNSUPPORTED.  The first call on a device looks up the array's
 * address; later calls make no allocation, free or allocating library call.  All calls on one GPU
 * share the array: calls that overlap in time on one GPU (other streams, other threads) race and
 * give wrong results.  Results equal pcx_fft_batch's bit for bit. */
int pcx_fft_batch_noalloc(pcx_plan_t plan, int inverse, int normalized, float* dst, int64_t dst_pack,
                          int64_t dst_dist, const float* src, int64_t src_pack, int64_t src_dist,
                          int64_t src_size, int64_t batch, void* stream);

/* ------------------------------------------------------------------ FFT unit, double
 * Replaces fft_unit<double, Order, Allocator>::fft_unit(fft_size, sub_size, allocator)
 * (fft.hpp:600-605 with T = double; default pack 4, types.hpp:70-72).  Same checks and messages
 * as pcx_plan_create.  Orders normal and bit_reversed; fft_size <= 2^23.  The reference declares
 * the double unit but has no double kernel for the unordered order (avx2_fft.hpp:114-217 is
 * float-only), so order 2 is rejected with PCX_ERR_UNSUPPORTED. */
int pcx_plan_create_f64(int64_t fft_size, int64_t sub_size, int order, pcx_plan_t* plan);
/* Bytes per real element of a plan: 4 (pcx_plan_create) or 8 (pcx_plan_create_f64). */
int pcx_plan_element_size(pcx_plan_t plan);
/* pcx_fft / pcx_ifft / pcx_fft_batch for double data (fft.hpp:624-915 with T = double), with
 * the same arguments in double elements.  A float plan passed here, or a double plan passed to
 * the float functions, fails with PCX_ERR_INVALID_ARGUMENT. */
int pcx_fft_f64(pcx_plan_t plan, double* dst, int64_t dst_pack, const double* src, int64_t src_pack,
                int64_t src_size, void* stream);
int pcx_ifft_f64(pcx_plan_t plan, int normalized, double* dst, int64_t dst_pack, const double* src,
                 int64_t src_pack, void* stream);
int pcx_fft_batch_f64(pcx_plan_t plan, int inverse, int normalized, double* dst, int64_t dst_pack,
                      int64_t dst_dist, const double* src, int64_t src_pack, int64_t src_dist,
                      int64_t src_size, int64_t batch, void* stream);

/* FFT across a range of vectors: the transform of fft_unit_par (fft.hpp:2168-2572), whose
 * output order is wrong in the reference (its test is disabled, test_fft.cpp:789); here it is
 * natural.  For every k < length: dst_j[k] = sum_i src_i[k] * W^(i*j) (inverse: conjugate
 * twiddles, scaled by 1/fft_size if normalized), i, j < fft_size <= 8192.  src_vecs / dst_vecs
 * are device arrays of fft_size pointers to packed vectors of `length` elements; dst may equal
 * src element for element (in place) when the packs are equal; with different packs no
 * destination vector may overlap a source vector (the caller checks: the tables are device
 * memory).  Normal-order plans only. */
int pcx_fft_across(pcx_plan_t plan, int inverse, int normalized, float* const* dst_vecs, int64_t dst_pack,
                   const float* const* src_vecs, int64_t src_pack, int64_t length, void* stream);
/* The same for a double plan (fft_unit_par<double>, fft.hpp:2168-2172 with T = double):
 * fft_size <= 2048. */
int pcx_fft_across_f64(pcx_plan_t plan, int inverse, int normalized, double* const* dst_vecs, int64_t dst_pack,
                       const double* const* src_vecs, int64_t src_pack, int64_t length, void* stream);

/* ------------------------------------------------------------------ cross-correlation
 * Replaces fxcorr_unit<float>(g, fft_size) (fxcorr.hpp:62-69): spectrum of g (device, pack
 * g_pack, g_size taps) zero-extended to fft_size, scaled by 1/fft_size.  fft_size is any power
 * of two from 2 to 2^24: up to 16384 every block runs in one fused kernel (forward FFT,
 * conj-multiply, inverse FFT on chip); above, the blocks go through HBM and the batched
 * four-step FFT. */
int pcx_fxcorr_create(const float* g, int64_t g_pack, int64_t g_size, int64_t fft_size, void* stream,
                      pcx_fxcorr_t* xc);
int pcx_fxcorr_destroy(pcx_fxcorr_t xc);
/* Replaces fxcorr_unit::operator()(vector&) (fxcorr.hpp:102-136) when dst == src, and the
 * out-of-place overload (fxcorr.hpp:138-140, an empty stub in the reference) otherwise:
 * out[i] = sum_{m < g_size} f[i+m] * conj(g[m]), i < len, f zero beyond len, by overlap-save with
 * block step floor((fft_size - g_size + 1) / src_pack) * src_pack (at least 1). */
int pcx_fxcorr_exec(pcx_fxcorr_t xc, float* dst, int64_t dst_pack, const float* src, int64_t src_pack,
                    int64_t len, void* stream);
int64_t pcx_fxcorr_step(pcx_fxcorr_t xc, int64_t src_pack);
/* Multi-channel form: `channels` signals of `len` samples, channel c at src + c*src_dist and
 * dst + c*dst_dist floats, all correlated with the same kernel spectrum in one launch. */
int pcx_fxcorr_exec_batch(pcx_fxcorr_t xc, float* dst, int64_t dst_pack, int64_t dst_dist, const float* src,
                          int64_t src_pack, int64_t src_dist, int64_t len, int64_t channels, void* stream);
/* PCX_PREC_FP32 (default), PCX_PREC_BF16X3 or PCX_PREC_FP16X2 (matrix-core radix-16 butterflies
 * in both transforms, fft_size 16 .. 16384; no reference counterpart). */
int pcx_fxcorr_set_precision(pcx_fxcorr_t xc, int precision);
/* The reference's TmpFactory_ slot (fxcorr.hpp:10-32, 47, 82-91: scratch supplied by the caller,
 * one fft_size vector).  pcx_fxcorr_exec_batch_ws is pcx_fxcorr_exec_batch with the unit's
 * scratch taken from `workspace` (device memory of workspace_bytes >= pcx_fxcorr_workspace_min,
 * which is 8 * fft_size bytes): in place, it holds the block order's ticket counter and read
 * flags (4 bytes per block, up to workspace_bytes / 4 - 8 blocks per group of launches) when the
 * overlap fft_size - step is at most step, else the saved overlap of up to workspace_bytes /
 * (8 * (fft_size - step)) blocks at a time; above fft_size 16384 it holds the blocks in flight
 * (the four-step FFT keeps its intermediate in its plan).  The unit allocates no buffer of its
 * own.  Calls that share one workspace must not overlap. */
int64_t pcx_fxcorr_workspace_min(pcx_fxcorr_t xc);
int     pcx_fxcorr_exec_batch_ws(pcx_fxcorr_t xc, float* dst, int64_t dst_pack, int64_t dst_dist, const float* src,
                                 int64_t src_pack, int64_t src_dist, int64_t len, int64_t channels, void* workspace,
                                 int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ uneven batches
 * A work queue of transforms of uneven power-of-two sizes and mixed directions, natural order,
 * each vector in split layout (pack = its size: re[N] then im[N]) at its own offset.  The
 * reference runs one fft_unit per size and one call per vector; this is its batched form.
 * Every transform of 2^1 .. 2^20 points runs in one persistent launch whose workgroups pull
 * items from a device counter, largest first: groups of single-pass transforms up to 2^14, and
 * whole four-step transforms of 2^15 .. 2^20 (each through an intermediate owned by its
 * workgroup: the queue holds 2 * 2^L floats per CU, 2^L the largest such size in the list).
 * Sizes above 2^20 follow as grouped four-step launches.  pcx_queue_capture records the whole
 * pass into a hipGraph for fixed (dst, src); pcx_queue_autotune times the launch shapes on the
 * caller's data and keeps the fastest (optionally persisted in a cache file). */
typedef struct {
    int64_t log2n;      /* transform size 2^log2n, 1 <= log2n <= 24 */
    int32_t inverse;    /* 0 forward, 1 inverse */
    int32_t reserved;   /* 0 */
    int64_t src_offset; /* floats from the source base */
    int64_t dst_offset; /* floats from the destination base */
} pcx_job;
typedef struct pcx_queue_s* pcx_queue_t;

int pcx_queue_create(const pcx_job* jobs, int64_t njobs, int normalized, pcx_queue_t* queue);
int pcx_queue_destroy(pcx_queue_t queue);
int pcx_queue_exec(pcx_queue_t queue, float* dst, const float* src, void* stream);
/* Times the launch shapes on (dst, src) and keeps the fastest.  dst is overwritten many times
 * and must not alias src.  A graph captured before is dropped (pcx_queue_launch then fails until
 * pcx_queue_capture runs again); the call synchronises the device. */
int pcx_queue_autotune(pcx_queue_t queue, float* dst, const float* src, void* stream, const char* cache_path);
int pcx_queue_capture(pcx_queue_t queue, float* dst, const float* src, void* stream);
int pcx_queue_launch(pcx_queue_t queue, void* stream);
/* algorithmic work of one pass: sum 5*N*log2(N) flops and 16*N bytes */
int pcx_queue_work(pcx_queue_t queue, double* flops, double* bytes);

/* ------------------------------------------------------------------ vector expressions
 * Device-side evaluation of the reference's vector expressions (vector_arithm.hpp:450-1130:
 * binary + - * / of two vectors of any pack sizes, + - * / with a real or complex scalar on
 * either side, conj) and of pack conversions (vector.hpp:55-75 copy between pack sizes;
 * avx2_common.hpp:243-324).  An expression tree is flattened into a program for a machine with
 * four complex registers per element; one fused kernel reads every operand once and writes the
 * destination once, in its pack size, for every element i < n.
 *   dst[i] = r0 after running prog[0 .. nprog) with the sources' element i.
 * Arithmetic follows the reference's formulas (simd_common.hpp:296-410): complex division is
 * (a * conj(b)) / (br^2 + bi^2); a real scalar adds to / subtracts from the real part only. */
typedef enum {
    PCX_VOP_LOAD = 0,  /* r[d] = src[a] element i */
    PCX_VOP_ADD  = 1,  /* r[d] = r[a] + r[b] */
    PCX_VOP_SUB  = 2,  /* r[d] = r[a] - r[b] */
    PCX_VOP_MUL  = 3,  /* r[d] = r[a] * r[b] */
    PCX_VOP_DIV  = 4,  /* r[d] = r[a] / r[b] */
    PCX_VOP_CONJ = 5,  /* r[d] = conj(r[a]) */
    PCX_VOP_ADDS = 6,  /* r[d] = r[a] + s          (s complex: s_re + i s_im) */
    PCX_VOP_SUBS = 7,  /* r[d] = r[a] - s */
    PCX_VOP_RSUBS = 8, /* r[d] = s - r[a] */
    PCX_VOP_MULS = 9,  /* r[d] = r[a] * s */
    PCX_VOP_DIVS = 10, /* r[d] = r[a] / s */
    PCX_VOP_RDIVS = 11,/* r[d] = s / r[a] */
    PCX_VOP_ADDR = 12, /* the same six with a real scalar s_re (s_im ignored) */
    PCX_VOP_SUBR = 13,
    PCX_VOP_RSUBR = 14,
    PCX_VOP_MULR = 15,
    PCX_VOP_DIVR = 16,
    PCX_VOP_RDIVR = 17
} pcx_vop;
enum { PCX_VEXPR_MAX_PROG = 64, PCX_VEXPR_MAX_SRC = 16, PCX_VEXPR_REGS = 4 };
typedef struct {
    int32_t op;   /* pcx_vop */
    int32_t d;    /* destination register, < PCX_VEXPR_REGS */
    int32_t a;    /* register operand, or the source index for PCX_VOP_LOAD */
    int32_t b;    /* second register operand (binary ops) */
    float   s_re; /* scalar operand */
    float   s_im;
} pcx_vinstr;
typedef struct {
    const float* data; /* device pointer, packed layout */
    int64_t      pack; /* pack size, a power of two */
} pcx_vsource;
int pcx_vexpr_eval(const pcx_vsource* srcs, int32_t nsrc, const pcx_vinstr* prog, int32_t nprog, float* dst,
                   int64_t dst_pack, int64_t n, void* stream);
/* Pack conversion: dst (pack dst_pack) = src (pack src_pack), n elements; in place only when the
 * packs are equal (vector.hpp copy between pack sizes). */
int pcx_repack(float* dst, int64_t dst_pack, const float* src, int64_t src_pack, int64_t n, void* stream);
/* The same for double vectors and scalars (tests/test_op.cpp runs every operator for float and
 * double): programs of pcx_vinstr_f64 over pcx_vsource_f64, double arithmetic throughout. */
typedef struct {
    int32_t op, d, a, b; /* as in pcx_vinstr */
    double  s_re, s_im;
} pcx_vinstr_f64;
typedef struct {
    const double* data;
    int64_t       pack;
} pcx_vsource_f64;
int pcx_vexpr_eval_f64(const pcx_vsource_f64* srcs, int32_t nsrc, const pcx_vinstr_f64* prog, int32_t nprog,
                       double* dst, int64_t dst_pack, int64_t n, void* stream);
int pcx_repack_f64(double* dst, int64_t dst_pack, const double* src, int64_t src_pack, int64_t n, void* stream);

/* ------------------------------------------------------------------ memory and utilities
 * The reference's Allocator template parameter (allocators.hpp:11-48) decides where buffers
 * live; these back the C++ shim's device allocator. */
int pcx_malloc(void** ptr, size_t bytes);
int pcx_free(void* ptr);
int pcx_memcpy_htod(void* dst, const void* src, size_t bytes);
int pcx_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int pcx_memcpy_dtod(void* dst, const void* src, size_t bytes);
int pcx_memset(void* dst, int value, size_t bytes);
int pcx_synchronize(void* stream);
/* Seeded synthetic input written straight into HBM: splitmix64 outputs 2i and 2i+1 give
 * uniform [-1,1) re/im (dist 0) or Box-Muller normal(0, sigma^2) re/im (dist 1). */
int pcx_generate(float* dst, int64_t pack, int64_t n, uint64_t seed, int dist, double sigma, void* stream);
/* The same values for double data (not rounded to float). */
int pcx_generate_f64(double* dst, int64_t pack, int64_t n, uint64_t seed, int dist, double sigma, void* stream);

#ifdef __cplusplus
}
#endif
#endif
