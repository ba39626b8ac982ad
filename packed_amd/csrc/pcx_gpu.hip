// pcx_gpu.hip -- C ABI of the gfx950 FFT / overlap-save cross-correlation path (include/pcx_gpu.h).
//
// Plans own device twiddle tables and, for N > 16384, a buffer for the four-step intermediate.  Everything is launched asynchronously on the caller's stream.
#include "pcx_gpu.h"

#include "dispatch.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>
#include <algorithm>

using namespace pcxk;
using namespace pcxd;

namespace {

thread_local std::string g_err;

int fail(int code, std::string msg) {
    g_err = std::move(msg);
    return code;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return fail(PCX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));              \
    } while (0)

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
int  ilog2(int64_t v) {
    int l = 0;
    while (((int64_t)1 << l) < v)
        ++l;
    return l;
}

std::string s(int64_t v) { return std::to_string(v); }

// fft.hpp:1806-1812
int check_size(int64_t n) {
    if (n > 1 && is_pow2(n))
        return PCX_SUCCESS;
    return fail(PCX_ERR_INVALID_SIZE, "fft_size (which is  " + s(n) + ") is not an integer power of two");
}
// fft.hpp:1814-1826
int check_sub_size(int64_t sub) {
    if (sub >= 8 && is_pow2(sub))
        return PCX_SUCCESS;
    if (sub < 8)
        return fail(PCX_ERR_INVALID_SUB_SIZE,
                    "fft_sub_size (which is  " + s(sub) + ") is smaller than minimum (which is 8)");
    return fail(PCX_ERR_INVALID_SUB_SIZE, "fft_sub_size (which is  " + s(sub) + ") is not an integer power of two");
}
int check_pack(int64_t p, const char* what) {
    if (is_pow2(p) && p <= ((int64_t)1 << 40))
        return PCX_SUCCESS;
    return fail(PCX_ERR_INVALID_PACK, std::string(what) + " pack size (which is " + s(p) + ") is not a power of two");
}

// Plan-time work (twiddle tables, kernel spectra) runs on one non-blocking stream per device, and
// set-up waits for that stream only: other streams of the process are never serialised by it.
constexpr int kMaxDevices = 64;
std::mutex    g_setup_mu;
hipStream_t   g_setup[kMaxDevices] = {};

hipError_t setup_stream(hipStream_t* out) {
    int        dev = 0;
    hipError_t e   = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_setup_mu);
    if (!g_setup[dev]) {
        e = hipStreamCreateWithFlags(&g_setup[dev], hipStreamNonBlocking);
        if (e != hipSuccess)
            return e;
    }
    *out = g_setup[dev];
    return hipSuccess;
}

template<class V>
hipError_t gen_twiddles(V* t, int64_t M, int64_t stride, int64_t count, hipStream_t st) {
    using T    = decltype(V{}.x);
    int64_t nb = (count + 255) / 256;
    hipLaunchKernelGGL(k_twiddles<T>, dim3((unsigned)nb), dim3(256), 0, st, t, M, stride, count);
    return hipGetLastError();
}


// ---------------------------------------------------------------- shape lookup
template<class T>
const FourStepT<T>* find_in(const ShapeList<T>& l, int lg, int n2, int g, int c) {
    for (int i = 0; i < l.count; ++i) {
        const FourStepT<T>& f = l.data[i];
        if (f.lg == lg && (n2 == 0 || f.n2 == n2) && (g == 0 || f.g == g) && (c == 0 || f.c == c))
            return &f;
    }
    return nullptr;
}
// The first listed shape that matches (0 = any): the defaults first, then the alternatives.
const FourStep* find_shape(int lg, int n2, int g, int c) {
    if (const FourStep* f = find_in(shapes_default(), lg, n2, g, c))
        return f;
    return find_in(shapes_alt(), lg, n2, g, c);
}
// The default shape of a plan: the first listed one of its size, except for the permuted orders
// of 2^21 .. 2^24.  Their row pass stores contiguous blocks, so rows of 2^14 points, one per
// workgroup, with the widest column tiles that leaves win (per 2 GiB, forward / inverse, against
// the first listed shape, whose wide row groups serve the natural order's strided store):
// 2^21 2.09 / 2.12 ms (2.33 / 2.25), 2^22 2.05 / 2.19 (2.58 / 2.51), 2^23 2.29 / 2.46
// (3.14 / 3.21), 2^24 2.75 / 2.99 (2.93 / 3.29).  At 2^20, 4096-point rows (four per workgroup)
// and 64-column tiles: 1.97 / 1.93 ms against 2.12 / 2.06 (one 16384-point row per workgroup
// made the inverse slower there, 2.66 ms).
const FourStep* default_shape(int lg, int order) {
    if (order != PCX_ORDER_NORMAL && lg >= 21 && lg <= 24)
        if (const FourStep* f = find_shape(lg, lg - 14, 1, 16384 >> (lg - 14)))
            return f;
    if (order != PCX_ORDER_NORMAL && lg == 20)
        if (const FourStep* f = find_shape(20, 8, 4, 64))
            return f;
    return find_shape(lg, 0, 0, 0);
}
const FourStepT<double>* find_shape64(int lg, int n2, int g, int c) {
    if (const FourStepT<double>* f = find_in(shapes64_default(), lg, n2, g, c))
        return f;
    return find_in(shapes64_alt(), lg, n2, g, c);
}
// The same for FP64: bit_reversed 2^21 .. 2^23 take rows of 2^13 points, one per workgroup, and
// the widest column tiles (per 2 GiB, forward / inverse: 2.02 / 1.92, 2.06 / 2.14, 2.53 / 2.61 ms
// against 2.41 / 2.47, 2.53 / 2.64, 4.23 / 4.44 with the first listed shapes).
const FourStepT<double>* default_shape64(int lg, int order) {
    if (order != PCX_ORDER_NORMAL && lg >= 21 && lg <= 23)
        if (const FourStepT<double>* f = find_shape64(lg, lg - 13, 1, 8192 >> (lg - 13)))
            return f;
    return find_shape64(lg, 0, 0, 0);
}
// The split kernels for the plan's shape: same size and the same N1 x N2 split (the plan's
// scratch layout and row / column counts follow its shape); only the default shapes have them.
const FourStepMfma* find_mfma(int lg, const FourStep* sh) {
    for (const auto& f : shapes_mfma())
        if (f.lg == lg && sh == find_shape(lg, 0, 0, 0) && f.n1 == sh->n1 && f.n2 == sh->n2)
            return &f;
    return nullptr;
}
// The same above 2^20: the split shapes of 2^21 .. 2^24 follow the default shapes.
// (Same split N1 x N2 as the plan's shape: the plan's tables follow the split; rows per
// workgroup and columns per tile are the split kernels' own.)
const FourStep* find_mfma_big(int lg, const FourStep* sh, int precision) {
    if (lg < 21 || lg > 24 || precision < PCX_PREC_BF16X3 || precision > PCX_PREC_FP16X2 || !sh)
        return nullptr;
    const FourStep* f = &shapes_mfma_big()[precision - 1][lg - 21];
    return f->n1 == sh->n1 && f->n2 == sh->n2 ? f : nullptr;
}

} // namespace

// ---------------------------------------------------------------- plan
struct pcx_plan_s {
    int64_t n     = 0;
    int     lg    = 0;
    int     order = 0;
    bool    two_pass = false;
    int     n1_lg = 0, n2_lg = 0; // N = N1 * N2, N1 contiguous (row) length, N2 column length
    float2* tw_row = nullptr;     // W_N (single pass) or W_N1
    float2* tw_col = nullptr;     // W_N2
    float2* twl    = nullptr;     // W_N^e, e < 2^tw_ls
    float2* twh    = nullptr;     // W_N^(e << tw_ls)
    int     tw_ls  = 0;
    // four-step intermediate of one chunk (allocated on the first large call); calls on different
    // streams are ordered on it: tmp_ev is recorded after the last use, on tmp_stream
    float*      tmp        = nullptr;
    int64_t     tmp_cap    = 0; // transforms the intermediate holds
    hipEvent_t  tmp_ev     = nullptr;
    hipStream_t tmp_stream = nullptr;
    bool        tmp_used   = false;
    int64_t chunk        = 0; // transforms per four-step chunk
    const FourStep* shape = nullptr; // four-step kernels and tile sizes
    int     precision    = PCX_PREC_FP32;
    int64_t grid_col     = 2048; // grid caps of the two passes (persistent loop; 0: one workgroup per tile)
    int64_t grid_row     = 2048;
    int     nt_flags     = 0; // bit 0/1: streaming load/store of src/dst, bit 2/3: of the intermediate
    float2* tw_dif       = nullptr; // W_M for the one-pass k_dif kernel
    bool    use_dif      = false;   // out-of-place transforms of 2^15 run as one k_dif pass
    // three-pass natural-order transforms of 2^21 .. 2^24 (FP64: 2^21 .. 2^23): stage tables
    // W_N2, W_Nb, W_Na and the two-level W_N1 pass table (lo, hi), in the plan's precision
    bool    use_three    = false;
    void*   tw3[5]       = {};
    int     tw_ls1       = 0;
    // FP64 plans (pcx_plan_create_f64): the same roles in double precision
    bool     f64      = false;
    double2* tw64_row = nullptr;
    double2* tw64_col = nullptr;
    double2* twl64    = nullptr;
    double2* twh64    = nullptr;
    const FourStepT<double>* shape64 = nullptr;
    std::mutex mu;
};

extern "C" int pcx_version(void) { return 2; }
extern "C" const char* pcx_last_error(void) { return g_err.c_str(); }

namespace {

// Default bytes of one four-step chunk's intermediate (2 GiB: the whole 2^28-point headline batch
// in one chunk; a 288 GB part holds it several times over).
constexpr int64_t kChunkBytes = (int64_t)2048 << 20;

// Twiddle tables of a four-step FP32 plan for its shape's N1 x N2 split (W_N1, W_N2) and the
// pass tables W_N = lo[e mod 2^ls] * hi[e >> ls].
// Tables of the three-pass form: stage tables W_N2, W_Nb, W_Na and the two-level W_N1 table of
// pass 2 (the W_N pass table is the four-step one).  T = float or double.
template<class T>
hipError_t three_pass_tables(pcx_plan_t p, hipStream_t st) {
    using V                   = vec2<T>;
    const ThreePassT<T>* tp   = three_of<T>(p->lg);
    if (!tp)
        return hipSuccess;
    const int     l1 = p->lg - tp->n2_lg;
    const int64_t n1 = (int64_t)1 << l1;
    p->tw_ls1        = (l1 + 1) / 2;
    const int64_t len[5]   = {(int64_t)1 << tp->n2_lg, (int64_t)1 << kThreeNbLog2, (int64_t)1 << tp->na_lg,
                              (int64_t)1 << p->tw_ls1, n1 >> p->tw_ls1};
    const int64_t m[5]     = {len[0], len[1], len[2], n1, n1};
    const int64_t stride[5] = {1, 1, 1, 1, len[3]};
    hipError_t    he       = hipSuccess;
    for (int k = 0; k < 5 && he == hipSuccess; ++k) {
        he = hipMalloc(&p->tw3[k], sizeof(V) * len[k]);
        if (he == hipSuccess)
            he = gen_twiddles((V*)p->tw3[k], m[k], stride[k], len[k], st);
    }
    if (he == hipSuccess)
        p->use_three = true;
    return he;
}

hipError_t fourstep_tables(pcx_plan_t p, hipStream_t st) {
    const int64_t n1 = (int64_t)1 << p->n1_lg, n2 = (int64_t)1 << p->n2_lg;
    const int64_t nl = (int64_t)1 << p->tw_ls, nh = p->n >> p->tw_ls;
    hipError_t    he = hipMalloc(&p->tw_row, sizeof(float2) * n1);
    if (he == hipSuccess)
        he = hipMalloc(&p->tw_col, sizeof(float2) * n2);
    if (he == hipSuccess && !p->twl)
        he = hipMalloc(&p->twl, sizeof(float2) * nl);
    if (he == hipSuccess && !p->twh)
        he = hipMalloc(&p->twh, sizeof(float2) * nh);
    if (he == hipSuccess)
        he = gen_twiddles(p->tw_row, n1, 1, n1, st);
    if (he == hipSuccess)
        he = gen_twiddles(p->tw_col, n2, 1, n2, st);
    if (he == hipSuccess)
        he = gen_twiddles(p->twl, p->n, 1, nl, st);
    if (he == hipSuccess)
        he = gen_twiddles(p->twh, p->n, nl, nh, st);
    return he;
}

} // namespace

extern "C" int pcx_plan_create(int64_t fft_size, int64_t sub_size, int order, pcx_plan_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan output pointer is null");
    *out = nullptr;
    if (int e = check_size(fft_size))
        return e;
    if (int e = check_sub_size(sub_size))
        return e;
    if (order < 0 || order > 2)
        return fail(PCX_ERR_INVALID_ARGUMENT, "order (which is " + s(order) + ") is not a fft_order");
    int lg = ilog2(fft_size);
    if (lg > kMaxLog2)
        return fail(PCX_ERR_UNSUPPORTED, "fft_size (which is " + s(fft_size) + ") exceeds the supported maximum 2^24");
    if (order == PCX_ORDER_UNORDERED && lg < 5)
        return fail(PCX_ERR_UNSUPPORTED,
                    "unordered fft_size (which is " + s(fft_size) + ") is below the 32-point finisher size");
    hipStream_t st = nullptr;
    HIPCHK(setup_stream(&st));
    auto* p  = new pcx_plan_s;
    p->n     = fft_size;
    p->lg    = lg;
    p->order = order;
    hipError_t he = hipSuccess;
    if (lg <= kMaxSinglePass) {
        he = hipMalloc(&p->tw_row, sizeof(float2) * fft_size);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw_row, fft_size, 1, fft_size, st);
    } else {
        p->two_pass = true;
        p->shape    = default_shape(lg, order);
        if (!p->shape) {
            delete p;
            return fail(PCX_ERR_UNSUPPORTED, "no four-step kernels for fft_size " + s(fft_size));
        }
        p->n2_lg = p->shape->n2;
        p->n1_lg = p->shape->n1;
        p->tw_ls = (lg + 1) / 2;
        he       = fourstep_tables(p, st);
        // transforms per four-step chunk: the intermediate holds one chunk
        p->chunk = std::max<int64_t>(1, kChunkBytes / (fft_size * 8));
        if (lg == kDifLog2) {
            p->use_dif     = true;
            const int64_t m = (int64_t)1 << dif_core_log2(order);
            if (he == hipSuccess)
                he = hipMalloc(&p->tw_dif, sizeof(float2) * m);
            if (he == hipSuccess)
                he = gen_twiddles(p->tw_dif, m, 1, m, st);
        }
        // three passes are the default: 2^21 / 2^22 / 2^23 / 2^24 in 2.74 / 2.83 / 2.93 /
        // 2.95 ms per 2 GiB forward against 3.03 / 3.67 / 4.45 / 6.00 with the four-step
        if (he == hipSuccess)
            he = three_pass_tables<float>(p, st);
        if (he == hipSuccess)
            he = hipEventCreateWithFlags(&p->tmp_ev, hipEventDisableTiming);
    }
    if (he == hipSuccess)
        he = hipStreamSynchronize(st);
    if (he != hipSuccess) {
        pcx_plan_destroy(p);
        return fail(PCX_ERR_HIP, std::string("plan setup: ") + hipGetErrorString(he));
    }
    *out = p;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_destroy(pcx_plan_t p) {
    if (!p)
        return PCX_SUCCESS;
    // the intermediate may still be in use by enqueued work
    if (p->tmp_ev && p->tmp_used)
        (void)hipEventSynchronize(p->tmp_ev);
    (void)hipFree(p->tw_row);
    (void)hipFree(p->tw_col);
    (void)hipFree(p->twl);
    (void)hipFree(p->twh);
    (void)hipFree(p->tmp);
    (void)hipFree(p->tw_dif);
    for (void* t : p->tw3)
        (void)hipFree(t);
    (void)hipFree(p->tw64_row);
    (void)hipFree(p->tw64_col);
    (void)hipFree(p->twl64);
    (void)hipFree(p->twh64);
    if (p->tmp_ev)
        (void)hipEventDestroy(p->tmp_ev);
    delete p;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_create_f64(int64_t fft_size, int64_t sub_size, int order, pcx_plan_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan output pointer is null");
    *out = nullptr;
    if (int e = check_size(fft_size))
        return e;
    if (int e = check_sub_size(sub_size))
        return e;
    if (order < 0 || order > 2)
        return fail(PCX_ERR_INVALID_ARGUMENT, "order (which is " + s(order) + ") is not a fft_order");
    if (order == PCX_ORDER_UNORDERED)
        return fail(PCX_ERR_UNSUPPORTED, "the unordered order is defined for float only (the reference's "
                                         "unsorted kernels are float-only)");
    int lg = ilog2(fft_size);
    if (lg > kMaxLog2_64)
        return fail(PCX_ERR_UNSUPPORTED,
                    "double fft_size (which is " + s(fft_size) + ") exceeds the supported maximum 2^23");
    hipStream_t st = nullptr;
    HIPCHK(setup_stream(&st));
    auto* p  = new pcx_plan_s;
    p->n     = fft_size;
    p->lg    = lg;
    p->order = order;
    p->f64   = true;
    hipError_t he = hipSuccess;
    if (lg <= kMaxSinglePass64) {
        he = hipMalloc(&p->tw64_row, sizeof(double2) * fft_size);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw64_row, fft_size, 1, fft_size, st);
    } else {
        p->two_pass = true;
        p->shape64  = default_shape64(lg, order);
        p->n1_lg    = p->shape64->n1;
        p->n2_lg    = p->shape64->n2;
        p->tw_ls    = (lg + 1) / 2;
        int64_t n1 = (int64_t)1 << p->n1_lg, n2 = (int64_t)1 << p->n2_lg;
        int64_t nl = (int64_t)1 << p->tw_ls, nh = fft_size >> p->tw_ls;
        he = hipMalloc(&p->tw64_row, sizeof(double2) * n1);
        if (he == hipSuccess)
            he = hipMalloc(&p->tw64_col, sizeof(double2) * n2);
        if (he == hipSuccess)
            he = hipMalloc(&p->twl64, sizeof(double2) * nl);
        if (he == hipSuccess)
            he = hipMalloc(&p->twh64, sizeof(double2) * nh);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw64_row, n1, 1, n1, st);
        if (he == hipSuccess)
            he = gen_twiddles(p->tw64_col, n2, 1, n2, st);
        if (he == hipSuccess)
            he = gen_twiddles(p->twl64, fft_size, 1, nl, st);
        if (he == hipSuccess)
            he = gen_twiddles(p->twh64, fft_size, nl, nh, st);
        p->chunk = std::max<int64_t>(1, kChunkBytes / (fft_size * 16));
        if (he == hipSuccess)
            he = three_pass_tables<double>(p, st);
        if (he == hipSuccess)
            he = hipEventCreateWithFlags(&p->tmp_ev, hipEventDisableTiming);
    }
    if (he == hipSuccess)
        he = hipStreamSynchronize(st);
    if (he != hipSuccess) {
        pcx_plan_destroy(p);
        return fail(PCX_ERR_HIP, std::string("plan setup: ") + hipGetErrorString(he));
    }
    *out = p;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_element_size(pcx_plan_t p) { return p ? (p->f64 ? 8 : 4) : 0; }

extern "C" int64_t pcx_plan_size(pcx_plan_t p) { return p ? p->n : 0; }

namespace {
int plan_apply_shape(pcx_plan_t p, const FourStep* sh);
}

extern "C" int pcx_plan_set_precision(pcx_plan_t p, int precision) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->f64)
        return fail(PCX_ERR_UNSUPPORTED, "a double plan has no precision modes");
    if (precision == PCX_PREC_FP32) {
        p->precision = precision;
        return PCX_SUCCESS;
    }
    if (precision != PCX_PREC_BF16X3 && precision != PCX_PREC_FP16X2)
        return fail(PCX_ERR_UNSUPPORTED, "precision " + s(precision) + " is not implemented");
    // the split kernels of four-step sizes follow the first listed shape's split: a plan on
    // another one (the permuted orders' default at 2^20 .. 2^24) moves to it
    if (p->two_pass && p->lg <= 24) {
        const FourStep* d = find_shape(p->lg, 0, 0, 0);
        const bool has = p->lg <= 20 ? find_mfma(p->lg, p->shape) != nullptr
                                     : find_mfma_big(p->lg, p->shape, precision) != nullptr;
        const bool d_has = d && (p->lg <= 20 ? find_mfma(p->lg, d) != nullptr
                                             : find_mfma_big(p->lg, d, precision) != nullptr);
        if (!has && d_has)
            if (int e = plan_apply_shape(p, d))
                return e;
    }
    const bool covered = p->lg >= 4 && (p->lg <= 20 ? (!p->two_pass || find_mfma(p->lg, p->shape))
                                                     : find_mfma_big(p->lg, p->shape, precision) != nullptr);
    if (!covered)
        return fail(PCX_ERR_UNSUPPORTED, "split-precision transforms cover 2^4 .. 2^24 (default four-step shapes)");
    p->precision = precision;
    return PCX_SUCCESS;
}
extern "C" int pcx_plan_precision(pcx_plan_t p) { return p ? p->precision : -1; }
extern "C" int     pcx_plan_order(pcx_plan_t p) { return p ? p->order : -1; }

namespace {

// Switches a four-step plan to another instantiated shape, rebuilding the W_N1 / W_N2 tables
// when the split changes (the pass tables W_N do not depend on the split).
int plan_apply_shape(pcx_plan_t p, const FourStep* sh) {
    if (!p->two_pass || p->f64 || !sh || sh->lg != p->lg)
        return fail(PCX_ERR_INVALID_ARGUMENT, "shape does not match the plan");
    if (sh->n1 != p->n1_lg || sh->n2 != p->n2_lg) {
        // the W_N1 / W_N2 tables follow the split; every enqueued use of the old ones must end
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(p->tw_row));
        HIPCHK(hipFree(p->tw_col));
        p->tw_row = p->tw_col = nullptr;
        p->n1_lg = sh->n1, p->n2_lg = sh->n2;
        hipStream_t st = nullptr;
        HIPCHK(setup_stream(&st));
        HIPCHK(fourstep_tables(p, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    p->shape = sh, p->n1_lg = sh->n1, p->n2_lg = sh->n2;
    return PCX_SUCCESS;
}

int plan_apply_shape64(pcx_plan_t p, const FourStepT<double>* sh) {
    if (!p->two_pass || !p->f64 || !sh || sh->lg != p->lg)
        return fail(PCX_ERR_INVALID_ARGUMENT, "shape does not match the plan");
    if (sh->n1 != p->n1_lg || sh->n2 != p->n2_lg) {
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(p->tw64_row));
        HIPCHK(hipFree(p->tw64_col));
        p->tw64_row = p->tw64_col = nullptr;
        const int64_t n1 = (int64_t)1 << sh->n1, n2 = (int64_t)1 << sh->n2;
        hipStream_t   st = nullptr;
        HIPCHK(setup_stream(&st));
        HIPCHK(hipMalloc(&p->tw64_row, sizeof(double2) * n1));
        HIPCHK(hipMalloc(&p->tw64_col, sizeof(double2) * n2));
        HIPCHK(gen_twiddles(p->tw64_row, n1, 1, n1, st));
        HIPCHK(gen_twiddles(p->tw64_col, n2, 1, n2, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    p->shape64 = sh, p->n1_lg = sh->n1, p->n2_lg = sh->n2;
    return PCX_SUCCESS;
}

} // namespace

namespace {

// A four-step intermediate the caller provides (the allocation-free entry point): `bytes` of
// device memory at `buf`, used instead of the plan's own.
struct ExtTmp {
    void*   buf;
    int64_t bytes;
};

// src_offs / dst_offs (optional, device arrays of `batch` element offsets) replace b * dist.
// T is the element type: float, or double for a plan made by pcx_plan_create_f64.
template<class T>
int run_batch(pcx_plan_t p, int inverse, int normalized, T* dst, int64_t dst_pack, int64_t dst_dist,
              const T* src, int64_t src_pack, int64_t src_dist, int64_t src_size, int64_t batch,
              hipStream_t st, const int64_t* src_offs = nullptr, const int64_t* dst_offs = nullptr,
              const ExtTmp* ext = nullptr) {
    constexpr bool F64 = std::is_same_v<T, double>;
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->f64 != F64)
        return fail(PCX_ERR_INVALID_ARGUMENT, p->f64 ? "the plan is for double data (use the _f64 functions)"
                                                     : "the plan is for float data (made by pcx_plan_create)");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (batch < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "batch (which is " + s(batch) + ") is negative");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    if (int e = check_pack(src_pack, "source"))
        return e;
    const int64_t n = p->n;
    if (src_size > n)
        return fail(PCX_ERR_SIZE_MISMATCH,
                    "source size (which is " + s(src_size) + ") is bigger than fft size (which is " + s(n) + ")");
    if (src_size < 0)
        return fail(PCX_ERR_SIZE_MISMATCH, "source size (which is " + s(src_size) + ") is negative");
    // spectra in unordered order keep pack max(P, 8) (avx2_fft.hpp:209-213, 227)
    if (p->order == PCX_ORDER_UNORDERED) {
        if (inverse && src_pack < 8)
            src_pack = 8;
        if (!inverse && dst_pack < 8)
            dst_pack = 8;
    }
    // fft.hpp:783, 889: a pack change wider than one SIMD register (8 floats, 4 doubles) cannot run
    // in place
    constexpr int64_t kReg = F64 ? 4 : 8;
    if ((const T*)dst == src && dst_pack != src_pack && (dst_pack > kReg || src_pack > kReg))
        return fail(PCX_ERR_INPLACE_REPACK, "cannot perform in-place repack from source pack size (which is " +
                                                s(src_pack) + ") to destination pack size (which is " + s(dst_pack) + ")");
    if (batch == 0)
        return PCX_SUCCESS;
    const int   src_lp = ilog2(src_pack), dst_lp = ilog2(dst_pack);
    const T     scale  = (inverse && normalized) ? (T)(1.0 / (double)n) : T(1);
    if (inverse)
        src_size = n;
    const vec2<T> *tw_row, *tw_col, *twl, *twh;
    if constexpr (F64)
        tw_row = p->tw64_row, tw_col = p->tw64_col, twl = p->twl64, twh = p->twh64;
    else
        tw_row = p->tw_row, tw_col = p->tw_col, twl = p->twl, twh = p->twh;

    if (!p->two_pass) {
        RowArgsT<T> a{};
        a.src = src, a.dst = dst, a.src_dist = src_dist, a.dst_dist = dst_dist;
        a.src_size = src_size, a.nrows = batch, a.src_lp = src_lp, a.dst_lp = dst_lp, a.rows_lg = 0;
        a.order = p->order, a.conj_in = inverse, a.conj_out = inverse, a.scale = scale;
        a.twA = tw_row, a.twl = nullptr, a.twh = nullptr, a.tw_ls = 0;
        a.src_offs = src_offs, a.dst_offs = dst_offs;
        if constexpr (F64)
            single64_tab()[p->lg - 1][single_variant(inverse, p->order)](a, batch, 0, st);
        else if (p->precision != PCX_PREC_FP32 && p->order == PCX_ORDER_NORMAL)
            single_mfma_tab()[p->precision - 1][p->lg - 4](a, batch, 0, st);
        else if (p->precision != PCX_PREC_FP32)
            single_mfma_perm_tab()[p->precision - 1][p->lg - 4][single_variant(inverse, p->order)](a, batch, 0, st);
        else
            single_tab()[p->lg - 1][single_variant(inverse, p->order)](a, batch, 0, st);
        HIPCHK(hipGetLastError());
        return PCX_SUCCESS;
    }

    // ---- one pass, radix-R decimation in frequency + M-point Stockham (no intermediate)
    if constexpr (!F64)
    if (p->use_dif && p->precision == PCX_PREC_FP32 && (const float*)dst != src &&
        (!inverse || p->order == PCX_ORDER_NORMAL)) {
        RowArgs a{};
        a.src = src, a.dst = dst, a.src_dist = src_dist, a.dst_dist = dst_dist;
        a.src_size = src_size, a.nrows = batch, a.src_lp = src_lp, a.dst_lp = dst_lp;
        a.conj_in = inverse, a.conj_out = inverse, a.scale = scale;
        a.twA = p->tw_dif, a.twl = p->twl, a.twh = p->twh, a.tw_ls = p->tw_ls, a.tw_lg = p->lg;
        a.nt_store = (p->nt_flags >> 1) & 1;
        a.src_offs = src_offs, a.dst_offs = dst_offs;
        dif_tab()[p->order](a, batch, st);
        HIPCHK(hipGetLastError());
        return PCX_SUCCESS;
    }

    // ---- four-step: N = N1 * N2 through an intermediate, `chunk` transforms at a time
    std::lock_guard<std::mutex> lock(p->mu);
    int64_t chunk = batch < p->chunk ? batch : p->chunk;
    T*      tmp   = nullptr;
    if (ext) {
        const int64_t cap = ext->bytes / (int64_t)(2 * sizeof(T) * n);
        if (cap < 1)
            return fail(PCX_ERR_UNSUPPORTED, "fft_size (which is " + s(n) + ") needs " + s(2 * sizeof(T) * n) +
                                                 " bytes of intermediate; the fixed workspace holds " + s(ext->bytes));
        chunk = std::min(chunk, cap);
        tmp   = reinterpret_cast<T*>(ext->buf);
    } else {
        // Calls on different streams share the plan's intermediate: a call on another stream than
        // the last user waits for that use (copies of a C++ fft_unit share one plan).  Not while
        // the stream is being captured into a graph (the event belongs to no capture).
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(st, &cs));
        const bool capturing = cs != hipStreamCaptureStatusNone;
        if (p->tmp_cap < chunk) {
            if (capturing)
                return fail(PCX_ERR_INVALID_ARGUMENT, "the plan's intermediate must be allocated before graph capture "
                                                      "(run one call of this batch size first)");
            if (p->tmp) {
                HIPCHK(hipEventSynchronize(p->tmp_ev)); // the last use, on whichever stream, has ended
                HIPCHK(hipFree(p->tmp));
                p->tmp     = nullptr;
                p->tmp_cap = 0;
            }
            hipError_t he = hipMalloc(&p->tmp, sizeof(T) * 2 * n * chunk);
            if (he != hipSuccess)
                return fail(PCX_ERR_OUT_OF_MEMORY,
                            "intermediate allocation of " + s(2 * sizeof(T) * n * chunk) + " bytes failed");
            p->tmp_cap = chunk;
        } else if (p->tmp_used && p->tmp_stream != st && !capturing) {
            HIPCHK(hipStreamWaitEvent(st, p->tmp_ev, 0));
        }
        tmp = reinterpret_cast<T*>(p->tmp);
    }
    const FourStepT<T>* shape;
    if constexpr (F64)
        shape = p->shape64;
    else
        shape = p->shape;
    int           C    = shape->c;
    ColLaunchT<T> colf = shape->cols[inverse ? 0 : 1];
    RowLaunchT<T> rowf = shape->rows[(inverse ? 3 : 0) + p->order];
    if constexpr (!F64)
    if (p->precision != PCX_PREC_FP32 && p->lg > 20) {
        const FourStep* fb = find_mfma_big(p->lg, p->shape, p->precision);
        if (!fb)
            return fail(PCX_ERR_UNSUPPORTED, "no split-precision kernels for this four-step shape");
        colf = fb->cols[inverse ? 0 : 1];
        rowf = fb->rows[(inverse ? 3 : 0) + p->order];
        C    = fb->c;
    } else if (p->precision != PCX_PREC_FP32) {
        const FourStepMfma* fm = find_mfma(p->lg, p->shape);
        if (!fm)
            return fail(PCX_ERR_UNSUPPORTED, "no split-precision kernels for this four-step shape");
        colf = fm->cols[p->precision - 1][inverse ? 0 : 1];
        rowf = p->order == PCX_ORDER_NORMAL
                   ? fm->rows[p->precision - 1][inverse ? 1 : 0]
                   : shapes_mfma_perm()[fm - shapes_mfma().data()][p->precision - 1][(inverse ? 2 : 0) + p->order - 1];
        C    = fm->c; // the split kernels keep their own column tile
    }
    const int64_t n1 = (int64_t)1 << p->n1_lg, n2 = (int64_t)1 << p->n2_lg;
    const ThreePassT<T>* tp3 = p->use_three && p->order == PCX_ORDER_NORMAL && p->precision == PCX_PREC_FP32
                                   ? three_of<T>(p->lg)
                                   : nullptr;
    for (int64_t b0 = 0; b0 < batch; b0 += chunk) {
        int64_t      cb = batch - b0 < chunk ? batch - b0 : chunk;
        const T*       cs  = src_offs ? src : src + b0 * src_dist;
        T*             cd  = dst_offs ? dst : dst + b0 * dst_dist;
        const int64_t* cso = src_offs ? src_offs + b0 : nullptr;
        const int64_t* cdo = dst_offs ? dst_offs + b0 : nullptr;
        if (tp3) {
            // natural order, forward or inverse (conj in pass 1, conj and 1/N in pass 3)
            using V          = vec2<T>;
            const int     l1 = p->lg - tp3->n2_lg; // log2 N1; the intermediate's pack
            const int64_t m1 = (int64_t)1 << l1;
            ColArgsT<T>   c1{}, c2{};
            RowArgsT<T>   r3{};
            c1.src = cs, c1.src_dist = src_dist, c1.src_lp = src_lp, c1.src_size = src_size, c1.src_offs = cso;
            c1.dst = tmp, c1.dst_dist = 2 * n, c1.dst_lp = l1;
            c1.ntiles = cb * (m1 / tp3->c1), c1.n1_lg = l1, c1.conj_in = inverse, c1.conj_out = 0, c1.scale = T(1);
            c1.twA = (const V*)p->tw3[0], c1.twl = twl, c1.twh = twh, c1.tw_ls = p->tw_ls, c1.tw_lg = p->lg;
            c1.nt_load = p->nt_flags & 1, c1.nt_store = (p->nt_flags >> 2) & 1;
            tp3->col1(c1, p->grid_col, st);
            c2.src = tmp, c2.src_dist = 2 * m1, c2.src_lp = l1, c2.src_size = m1;
            c2.dst = tmp, c2.dst_dist = 2 * m1, c2.dst_lp = l1;
            c2.ntiles = (cb << tp3->n2_lg) * (((int64_t)1 << tp3->na_lg) / tp3->c2), c2.n1_lg = tp3->na_lg;
            c2.conj_in = 0, c2.conj_out = 0, c2.scale = T(1);
            c2.twA = (const V*)p->tw3[1], c2.twl = (const V*)p->tw3[3], c2.twh = (const V*)p->tw3[4];
            c2.tw_ls = p->tw_ls1, c2.tw_lg = l1;
            tp3->col2(c2, p->grid_col, st);
            r3.src = tmp, r3.src_dist = 2 * n, r3.src_lp = l1, r3.src_size = n;
            r3.dst = cd, r3.dst_dist = dst_dist, r3.dst_lp = dst_lp, r3.dst_offs = cdo;
            r3.rows_lg = p->lg - tp3->na_lg, r3.rsplit_lg = tp3->n2_lg, r3.nrows = cb << (p->lg - tp3->na_lg);
            r3.order = PCX_ORDER_NORMAL, r3.conj_in = 0, r3.conj_out = inverse, r3.scale = scale;
            r3.twA = (const V*)p->tw3[2];
            r3.nt_load = (p->nt_flags >> 3) & 1, r3.nt_store = (p->nt_flags >> 1) & 1;
            tp3->row3(r3, r3.nrows, p->grid_row, st);
            HIPCHK(hipGetLastError());
            continue;
        }
        ColArgsT<T>  ca{};
        RowArgsT<T>  ra{};
        ca.twA = tw_col, ca.n1_lg = p->n1_lg, ca.ntiles = cb * (n1 / C);
        ra.twA = tw_row, ra.rows_lg = p->n2_lg, ra.nrows = cb * n2, ra.order = p->order;
        ca.tw_lg = p->lg, ra.tw_lg = p->lg;
        if (!inverse) {
            // pass A: columns of src -> intermediate, times W_N^(n1*k2)
            ca.src = cs, ca.src_dist = src_dist, ca.src_lp = src_lp, ca.src_size = src_size, ca.src_offs = cso;
            ca.dst = tmp, ca.dst_dist = 2 * n, ca.dst_lp = p->lg;
            ca.conj_in = 0, ca.conj_out = 0, ca.scale = T(1);
            ca.twl = twl, ca.twh = twh, ca.tw_ls = p->tw_ls;
            ca.nt_load = p->nt_flags & 1, ca.nt_store = (p->nt_flags >> 2) & 1;
            colf(ca, p->grid_col, st);
            // pass B: rows of the intermediate -> dst in the plan's order
            ra.src = tmp, ra.src_dist = 2 * n, ra.src_lp = p->lg, ra.src_size = n;
            ra.dst = cd, ra.dst_dist = dst_dist, ra.dst_lp = dst_lp, ra.dst_offs = cdo;
            ra.conj_in = 0, ra.conj_out = 0, ra.scale = T(1);
            ra.twl = nullptr, ra.twh = nullptr;
            ra.nt_load = (p->nt_flags >> 3) & 1, ra.nt_store = (p->nt_flags >> 1) & 1;
            rowf(ra, ra.nrows, p->grid_row, st);
        } else {
            // pass 1: rows (over k1) of the spectrum -> intermediate, times W_N^(n1*k2)
            ra.src = cs, ra.src_dist = src_dist, ra.src_lp = src_lp, ra.src_size = n, ra.src_offs = cso;
            ra.dst = tmp, ra.dst_dist = 2 * n, ra.dst_lp = p->lg;
            ra.conj_in = 1, ra.conj_out = 0, ra.scale = T(1);
            ra.twl = twl, ra.twh = twh, ra.tw_ls = p->tw_ls;
            ra.nt_load = p->nt_flags & 1, ra.nt_store = (p->nt_flags >> 2) & 1;
            rowf(ra, ra.nrows, p->grid_row, st);
            // pass 2: columns (over k2) of the intermediate -> dst, natural order
            ca.src = tmp, ca.src_dist = 2 * n, ca.src_lp = p->lg, ca.src_size = n;
            ca.dst = cd, ca.dst_dist = dst_dist, ca.dst_lp = dst_lp, ca.dst_offs = cdo;
            ca.conj_in = 0, ca.conj_out = 1, ca.scale = scale;
            ca.twl = nullptr, ca.twh = nullptr;
            ca.nt_load = (p->nt_flags >> 3) & 1, ca.nt_store = (p->nt_flags >> 1) & 1;
            colf(ca, p->grid_col, st);
        }
        HIPCHK(hipGetLastError());
    }
    if (!ext) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(st, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            HIPCHK(hipEventRecord(p->tmp_ev, st));
            p->tmp_stream = st;
            p->tmp_used   = true;
        }
    }
    return PCX_SUCCESS;
}

int64_t footprint(int64_t n, int64_t pack) { return ((n + pack - 1) / pack) * pack * 2; }

} // namespace

// ---------------------------------------------------------------- tuning of four-step plans
extern "C" int pcx_plan_set_shape(pcx_plan_t p, int n2_log2, int cols, int rows) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (!p->two_pass)
        return fail(PCX_ERR_UNSUPPORTED, "a single-pass plan (fft_size " + s(p->n) + ") has no four-step shape");
    if (p->precision != PCX_PREC_FP32)
        return fail(PCX_ERR_UNSUPPORTED, "split-precision plans keep the default four-step shape");
    if (n2_log2 < 0 || cols < 0 || rows < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "shape selectors are negative");
    std::lock_guard<std::mutex> lock(p->mu);
    if (p->f64) {
        const FourStepT<double>* sh = find_shape64(p->lg, n2_log2, rows, cols);
        if (!sh)
            return fail(PCX_ERR_UNSUPPORTED, "no double four-step shape of fft_size " + s(p->n) + " matches");
        return plan_apply_shape64(p, sh);
    }
    const FourStep* sh = find_shape(p->lg, n2_log2, rows, cols);
    if (!sh)
        return fail(PCX_ERR_UNSUPPORTED, "no four-step shape of fft_size " + s(p->n) + " matches");
    return plan_apply_shape(p, sh);
}

extern "C" int pcx_plan_shape(pcx_plan_t p, int* n2_log2, int* cols, int* rows) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    int a = 0, c = 0, g = 0;
    if (p->two_pass) {
        a = p->n2_lg;
        c = p->f64 ? p->shape64->c : p->shape->c;
        g = p->f64 ? p->shape64->g : p->shape->g;
    }
    if (n2_log2)
        *n2_log2 = a;
    if (cols)
        *cols = c;
    if (rows)
        *rows = g;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_set_grid(pcx_plan_t p, int64_t grid_col, int64_t grid_row) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (grid_col < 0 || grid_row < 0 || grid_col > ((int64_t)1 << 31) || grid_row > ((int64_t)1 << 31))
        return fail(PCX_ERR_INVALID_ARGUMENT, "grid caps are not in [0, 2^31]");
    std::lock_guard<std::mutex> lock(p->mu);
    p->grid_col = grid_col, p->grid_row = grid_row;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_set_chunk(pcx_plan_t p, int64_t bytes) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (bytes < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "chunk size (which is " + s(bytes) + ") is negative");
    std::lock_guard<std::mutex> lock(p->mu);
    const int64_t per = p->n * (p->f64 ? 16 : 8);
    p->chunk          = std::max<int64_t>(1, (bytes ? bytes : kChunkBytes) / per);
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_set_three_pass(pcx_plan_t p, int enable) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (!p->tw3[0])
        return enable ? fail(PCX_ERR_UNSUPPORTED, "no three-pass form for fft_size " + s(p->n)) : PCX_SUCCESS;
    p->use_three = enable != 0;
    return PCX_SUCCESS;
}

extern "C" int pcx_plan_set_one_pass(pcx_plan_t p, int enable) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (!p->tw_dif)
        return enable ? fail(PCX_ERR_UNSUPPORTED, "no one-pass kernel for fft_size " + s(p->n)) : PCX_SUCCESS;
    p->use_dif = enable != 0;
    return PCX_SUCCESS;
}

extern "C" int pcx_fft(pcx_plan_t p, float* dst, int64_t dst_pack, const float* src, int64_t src_pack,
                       int64_t src_size, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 0, 0, dst, dst_pack, 0, src, src_pack, 0, src_size, 1, (hipStream_t)stream);
}

extern "C" int pcx_ifft(pcx_plan_t p, int normalized, float* dst, int64_t dst_pack, const float* src,
                        int64_t src_pack, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 1, normalized, dst, dst_pack, 0, src, src_pack, 0, p->n, 1, (hipStream_t)stream);
}

namespace {
template<class T>
int fft_batch(pcx_plan_t p, int inverse, int normalized, T* dst, int64_t dst_pack, int64_t dst_dist, const T* src,
              int64_t src_pack, int64_t src_dist, int64_t src_size, int64_t batch, void* stream,
              const ExtTmp* ext = nullptr) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (batch > 1) {
        int64_t dp = dst_pack, sp = src_pack;
        if (p->order == PCX_ORDER_UNORDERED) {
            if (!inverse && dp < 8)
                dp = 8;
            if (inverse && sp < 8)
                sp = 8;
        }
        // a forward source holds src_size values (zero-extended to the fft size)
        const int64_t src_n = inverse ? p->n : std::min(std::max<int64_t>(src_size, 0), p->n);
        if (is_pow2(dp) && is_pow2(sp) &&
            (dst_dist < footprint(p->n, dp) || src_dist < footprint(src_n, sp)))
            return fail(PCX_ERR_INVALID_ARGUMENT, "batch distance is smaller than one vector");
        if ((const T*)dst == src && dst_dist != src_dist)
            return fail(PCX_ERR_INVALID_ARGUMENT, "in-place batch needs equal source and destination distances");
    }
    return run_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist,
                     inverse ? p->n : src_size, batch, (hipStream_t)stream, nullptr, nullptr, ext);
}
} // namespace

extern "C" int pcx_fft_batch(pcx_plan_t p, int inverse, int normalized, float* dst, int64_t dst_pack,
                             int64_t dst_dist, const float* src, int64_t src_pack, int64_t src_dist,
                             int64_t src_size, int64_t batch, void* stream) {
    return fft_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist, src_size, batch,
                     stream);
}

// ---------------------------------------------------------------- allocation-free batched FFT
// The four-step intermediate of pcx_fft_batch_noalloc: a fixed 128 MiB array per GPU (one
// 2^24-point FP32 transform, the largest plan), held until the process ends.  Host code uses the
// address hipGetSymbolAddress returns for the current device, never the array's name (a host-side
// stand-in).  The array needs no clearing: pass A writes every element pass B reads.
constexpr int64_t kNoAllocBytes = (int64_t)128 << 20;
static __device__ __align__(256) float scratch[kNoAllocBytes / (int64_t)sizeof(float)];

namespace {
std::mutex g_scratch_mu;
void*      g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int noalloc_scratch(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(PCX_ERR_UNSUPPORTED, "device " + s(dev) + " is beyond the supported " + s(kMaxDevices));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr)
            return fail(PCX_ERR_HIP, "hipGetSymbolAddress returned no address for the fixed intermediate");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return PCX_SUCCESS;
}
} // namespace

extern "C" int pcx_fft_batch_noalloc(pcx_plan_t p, int inverse, int normalized, float* dst, int64_t dst_pack,
                                     int64_t dst_dist, const float* src, int64_t src_pack, int64_t src_dist,
                                     int64_t src_size, int64_t batch, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->two_pass && 8 * p->n > kNoAllocBytes)
        return fail(PCX_ERR_UNSUPPORTED, "fft_size (which is " + s(p->n) + ") needs " + s(8 * p->n) +
                                             " bytes of intermediate; the fixed intermediate holds " + s(kNoAllocBytes));
    void* buf = nullptr;
    if (int e = noalloc_scratch(&buf))
        return e;
    const ExtTmp ext{buf, kNoAllocBytes};
    return fft_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist, src_size, batch, stream,
                     &ext);
}

// ---------------------------------------------------------------- FP64 transforms
extern "C" int pcx_fft_f64(pcx_plan_t p, double* dst, int64_t dst_pack, const double* src, int64_t src_pack,
                           int64_t src_size, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 0, 0, dst, dst_pack, 0, src, src_pack, 0, src_size, 1, (hipStream_t)stream);
}

extern "C" int pcx_ifft_f64(pcx_plan_t p, int normalized, double* dst, int64_t dst_pack, const double* src,
                            int64_t src_pack, void* stream) {
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    return run_batch(p, 1, normalized, dst, dst_pack, 0, src, src_pack, 0, p->n, 1, (hipStream_t)stream);
}

extern "C" int pcx_fft_batch_f64(pcx_plan_t p, int inverse, int normalized, double* dst, int64_t dst_pack,
                                 int64_t dst_dist, const double* src, int64_t src_pack, int64_t src_dist,
                                 int64_t src_size, int64_t batch, void* stream) {
    return fft_batch(p, inverse, normalized, dst, dst_pack, dst_dist, src, src_pack, src_dist, src_size, batch,
                     stream);
}

// ---------------------------------------------------------------- cross-correlation
struct pcx_fxcorr_s {
    int64_t    n = 0, m = 0;
    int        lg        = 0;
    int        precision = PCX_PREC_FP32;
    float2*    tw        = nullptr; // W_N of the fused kernel (N <= 16384)
    float*     kspec     = nullptr; // FFT(g)/N (fused kernel) or FFT(g) (block path), split re[N] im[N]
    pcx_plan_t plan      = nullptr; // block path (N > 16384): the batched four-step FFT of size N
    // the unit's own buffers: the in-place overlap of every block (fused kernel) or the blocks in
    // flight (block path); calls on another stream than the last user wait for that use (ev)
    float*      own      = nullptr;
    int64_t     own_cap  = 0; // floats
    hipEvent_t  ev       = nullptr;
    hipStream_t ev_stream = nullptr;
    bool        ev_used  = false;
    std::mutex  mu;
};

namespace {

// Block path: at most this many bytes of blocks in flight in the unit's own buffer.
constexpr int64_t kXcBlockBytes = (int64_t)512 << 20;

// The unit's own buffer with at least `floats` floats, ordered after its last use on any stream.
int xc_own(pcx_fxcorr_t x, int64_t floats, hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (x->own_cap < floats) {
        if (capturing)
            return fail(PCX_ERR_INVALID_ARGUMENT, "the correlation's buffers must be allocated before graph capture "
                                                  "(run one call of this size first)");
        if (x->own) {
            if (x->ev_used)
                HIPCHK(hipEventSynchronize(x->ev));
            HIPCHK(hipFree(x->own));
            x->own     = nullptr;
            x->own_cap = 0;
        }
        if (hipMalloc(&x->own, sizeof(float) * floats) != hipSuccess)
            return fail(PCX_ERR_OUT_OF_MEMORY, "correlation buffer allocation of " + s(4 * floats) + " bytes failed");
        x->own_cap = floats;
    } else if (x->ev_used && x->ev_stream != st && !capturing) {
        HIPCHK(hipStreamWaitEvent(st, x->ev, 0));
    }
    return PCX_SUCCESS;
}
int xc_own_done(pcx_fxcorr_t x, hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    if (cs == hipStreamCaptureStatusNone) {
        HIPCHK(hipEventRecord(x->ev, st));
        x->ev_stream = st;
        x->ev_used   = true;
    }
    return PCX_SUCCESS;
}

int64_t grid_of(int64_t work) { return std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, (int64_t)1 << 20)); }

} // namespace

extern "C" int pcx_fxcorr_create(const float* g, int64_t g_pack, int64_t g_size, int64_t fft_size, void* stream,
                                 pcx_fxcorr_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "output pointer is null");
    *out = nullptr;
    if (!g)
        return fail(PCX_ERR_INVALID_ARGUMENT, "kernel pointer is null");
    if (int e = check_size(fft_size))
        return e;
    if (int e = check_pack(g_pack, "kernel"))
        return e;
    if (g_size < 1 || g_size > fft_size)
        return fail(PCX_ERR_SIZE_MISMATCH,
                    "source size (which is " + s(g_size) + ") is bigger than fft size (which is " + s(fft_size) + ")");
    int lg = ilog2(fft_size);
    if (lg > kMaxLog2)
        return fail(PCX_ERR_UNSUPPORTED, "fxcorr fft_size (which is " + s(fft_size) + ") exceeds 2^24");
    hipStream_t st    = (hipStream_t)stream;
    hipStream_t setup = nullptr;
    HIPCHK(setup_stream(&setup));
    auto* x = new pcx_fxcorr_s;
    x->n = fft_size, x->m = g_size, x->lg = lg;
    hipError_t he = hipMalloc(&x->kspec, sizeof(float) * 2 * fft_size);
    if (he == hipSuccess)
        he = hipEventCreateWithFlags(&x->ev, hipEventDisableTiming);
    if (he == hipSuccess && lg <= kMaxSinglePass) {
        he = hipMalloc(&x->tw, sizeof(float2) * fft_size);
        if (he == hipSuccess)
            he = gen_twiddles(x->tw, fft_size, 1, fft_size, setup);
        if (he == hipSuccess)
            he = hipStreamSynchronize(setup);
    }
    if (he != hipSuccess) {
        pcx_fxcorr_destroy(x);
        return fail(PCX_ERR_HIP, std::string("fxcorr setup: ") + hipGetErrorString(he));
    }
    if (lg <= kMaxSinglePass) {
        // kernel spectrum: FFT of g zero-extended to N, scaled by 1/N (fxcorr.hpp:67-68)
        RowArgs a{};
        a.src = g, a.dst = x->kspec, a.src_dist = 0, a.dst_dist = 0, a.src_size = g_size, a.nrows = 1;
        a.src_lp = ilog2(g_pack), a.dst_lp = lg, a.rows_lg = 0;
        a.order = 0, a.conj_in = 0, a.conj_out = 0, a.scale = (float)(1.0 / (double)fft_size);
        a.twA = x->tw;
        single_tab()[lg - 1][0](a, 1, 0, st);
        he = hipGetLastError();
    } else {
        // block path: FFT(g) through the four-step plan; the 1/N is applied with the product
        if (int e = pcx_plan_create(fft_size, 2048, PCX_ORDER_NORMAL, &x->plan)) {
            pcx_fxcorr_destroy(x);
            return e;
        }
        if (int e = run_batch(x->plan, 0, 0, x->kspec, fft_size, 0, g, g_pack, 0, g_size, 1, st)) {
            pcx_fxcorr_destroy(x);
            return e;
        }
    }
    if (he == hipSuccess)
        he = hipStreamSynchronize(st);
    if (he != hipSuccess) {
        pcx_fxcorr_destroy(x);
        return fail(PCX_ERR_HIP, std::string("fxcorr kernel spectrum: ") + hipGetErrorString(he));
    }
    *out = x;
    return PCX_SUCCESS;
}

extern "C" int pcx_fxcorr_destroy(pcx_fxcorr_t x) {
    if (!x)
        return PCX_SUCCESS;
    if (x->ev && x->ev_used)
        (void)hipEventSynchronize(x->ev);
    (void)hipFree(x->tw);
    (void)hipFree(x->kspec);
    (void)hipFree(x->own);
    if (x->ev)
        (void)hipEventDestroy(x->ev);
    pcx_plan_destroy(x->plan);
    delete x;
    return PCX_SUCCESS;
}

extern "C" int64_t pcx_fxcorr_step(pcx_fxcorr_t x, int64_t src_pack) {
    if (!x || !is_pow2(src_pack))
        return -1;
    int64_t overlap = x->m - 1;
    if (x->n <= overlap)
        return -1;
    // fxcorr.hpp:110: step = ((N - overlap) / P) * P
    return (x->n - overlap) / src_pack * src_pack;
}

extern "C" int64_t pcx_fxcorr_workspace_min(pcx_fxcorr_t x) { return x ? 8 * x->n : -1; }

extern "C" int pcx_fxcorr_set_precision(pcx_fxcorr_t x, int precision) {
    if (!x)
        return fail(PCX_ERR_INVALID_ARGUMENT, "fxcorr handle is null");
    if (precision != PCX_PREC_FP32 && precision != PCX_PREC_BF16X3 && precision != PCX_PREC_FP16X2)
        return fail(PCX_ERR_UNSUPPORTED, "precision " + s(precision) + " is not implemented");
    if (precision != PCX_PREC_FP32 && (x->lg < 4 || x->lg > kMaxSinglePass))
        return fail(PCX_ERR_UNSUPPORTED, "split-precision correlation covers fft_size 16 .. 16384");
    x->precision = precision;
    return PCX_SUCCESS;
}

namespace {

int fxcorr_run(pcx_fxcorr_t x, float* dst, int64_t dst_pack, int64_t dst_dist, const float* src, int64_t src_pack,
               int64_t src_dist, int64_t len, int64_t channels, const ExtTmp* ws, hipStream_t st) {
    if (!x)
        return fail(PCX_ERR_INVALID_ARGUMENT, "fxcorr handle is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    if (int e = check_pack(src_pack, "source"))
        return e;
    if (len < 0 || channels < 0)
        return fail(PCX_ERR_SIZE_MISMATCH, "signal length (which is " + s(len) + ") or channel count is negative");
    int64_t step = pcx_fxcorr_step(x, src_pack);
    if (step <= 0)
        return fail(PCX_ERR_SIZE_MISMATCH, "fft size (which is " + s(x->n) + ") leaves no output per block for a kernel of " +
                                               s(x->m) + " taps at pack " + s(src_pack));
    const bool inplace = (const float*)dst == src;
    if (channels > 1 && (src_dist < footprint(len, src_pack) || dst_dist < footprint(len, dst_pack)))
        return fail(PCX_ERR_INVALID_ARGUMENT, "channel distance is smaller than one signal");
    if (inplace && (dst_pack != src_pack || dst_dist != src_dist))
        return fail(PCX_ERR_INPLACE_REPACK, "cannot perform in-place repack from source pack size (which is " +
                                                s(src_pack) + ") to destination pack size (which is " + s(dst_pack) + ")");
    if (ws && (!ws->buf || ws->bytes < 8 * x->n))
        return fail(PCX_ERR_INVALID_ARGUMENT, "workspace (which is " + s(ws ? ws->bytes : 0) +
                                                  " bytes) is smaller than one fft_size vector (which is " + s(8 * x->n) +
                                                  " bytes)");
    if (len == 0 || channels == 0)
        return PCX_SUCCESS;
    const int64_t nblocks = (len + step - 1) / step;
    const int64_t total   = nblocks * channels;
    const int64_t n       = x->n;
    std::lock_guard<std::mutex> lock(x->mu);

    if (x->lg > kMaxSinglePass) {
        // ---- block path: gather nb blocks, forward FFT, times conj(FFT(g)) / N, unnormalised
        // inverse, keep `step` outputs per block.  Blocks run in increasing order, and a block's
        // outputs lie below every later block's samples, so in place needs no saved overlap.
        float*  buf = nullptr;
        int64_t nb  = 0;
        if (ws) {
            buf = reinterpret_cast<float*>(ws->buf);
            nb  = std::min(total, ws->bytes / (8 * n));
        } else {
            nb = std::min(total, std::max<int64_t>(1, kXcBlockBytes / (8 * n)));
            if (int e = xc_own(x, 2 * n * nb, st))
                return e;
            buf = x->own;
        }
        XcBlockArgs a{};
        a.src = src, a.dst = dst, a.buf = buf, a.kre = x->kspec, a.kim = x->kspec + n;
        a.len = len, a.step = step, a.nblocks = nblocks, a.src_dist = src_dist, a.dst_dist = dst_dist;
        a.src_lp = ilog2(src_pack), a.dst_lp = ilog2(dst_pack), a.lg = x->lg;
        a.scale = (float)(1.0 / (double)n);
        for (a.bg0 = 0; a.bg0 < total; a.bg0 += nb) {
            a.nb = std::min(nb, total - a.bg0);
            hipLaunchKernelGGL(k_xc_gather, dim3((unsigned)grid_of(a.nb * n)), dim3(256), 0, st, a);
            HIPCHK(hipGetLastError());
            if (int e = run_batch(x->plan, 0, 0, buf, n, 2 * n, (const float*)buf, n, 2 * n, n, a.nb, st))
                return e;
            hipLaunchKernelGGL(k_xc_mul, dim3((unsigned)grid_of(a.nb * n)), dim3(256), 0, st, a);
            HIPCHK(hipGetLastError());
            if (int e = run_batch(x->plan, 1, 0, buf, n, 2 * n, (const float*)buf, n, 2 * n, n, a.nb, st))
                return e;
            hipLaunchKernelGGL(k_xc_scatter, dim3((unsigned)grid_of(a.nb * step)), dim3(256), 0, st, a);
            HIPCHK(hipGetLastError());
        }
        return ws ? PCX_SUCCESS : xc_own_done(x, st);
    }

    // ---- fused kernel: one workgroup per block
    XcArgs a{};
    a.src = src, a.dst = dst, a.kre = x->kspec, a.kim = x->kspec + n, a.twA = x->tw;
    a.len = len, a.step = step, a.nblocks = nblocks, a.hlen = n - step;
    a.src_lp = ilog2(src_pack), a.dst_lp = ilog2(dst_pack);
    a.src_dist = src_dist, a.dst_dist = dst_dist, a.channels = channels;
    const XcLaunch launch = x->precision != PCX_PREC_FP32 ? xc_mfma_tab()[x->precision - 1][x->lg - 4] : xc_tab()[x->lg - 1];
    if (!(inplace && nblocks > 1 && a.hlen > 0)) {
        a.blk0 = 0, a.blk1 = total;
        launch(a, st);
        HIPCHK(hipGetLastError());
        return PCX_SUCCESS;
    }
    // In place: a block overwrites samples that its predecessor reads.  When they lie within the
    // next block only (N - step <= step), workgroups run in ticket order and each stores after its
    // predecessor has read (XcArgs::order: one word per workgroup and launch); a workspace of W
    // bytes takes W / 4 - 8 blocks per group of launches.
    const int64_t owords = ws ? ws->bytes / 4 - 8 : total;
    if (a.hlen <= step && owords >= 1) {
        const int64_t group = std::min(total, owords);
        uint32_t*     obuf  = nullptr;
        if (ws) {
            obuf = reinterpret_cast<uint32_t*>(ws->buf);
        } else {
            if (int e = xc_own(x, total + 8, st))
                return e;
            obuf = reinterpret_cast<uint32_t*>(x->own);
        }
        a.order = obuf;
        for (int64_t g0 = 0; g0 < total; g0 += group) {
            a.blk0 = g0, a.blk1 = std::min(total, g0 + group);
            launch(a, st);
            HIPCHK(hipGetLastError());
        }
        return ws ? PCX_SUCCESS : xc_own_done(x, st);
    }
    // Otherwise the overlap is saved first.  With a workspace of W bytes the blocks run in groups
    // of W / (8 * (N - step)), one group after another (a group only overwrites samples of its own
    // blocks).
    const int64_t group = ws ? std::min(total, ws->bytes / (8 * a.hlen)) : total;
    float*        hbuf  = nullptr;
    if (ws) {
        hbuf = reinterpret_cast<float*>(ws->buf);
    } else {
        if (int e = xc_own(x, 2 * total * a.hlen, st))
            return e;
        hbuf = x->own;
    }
    for (int64_t g0 = 0; g0 < total; g0 += group) {
        const int64_t cnt = std::min(group, total - g0);
        a.hre = hbuf, a.him = hbuf + cnt * a.hlen;
        a.blk0 = g0, a.blk1 = g0 + cnt, a.hblk0 = g0;
        hipLaunchKernelGGL(k_fxcorr_halo, dim3((unsigned)grid_of(cnt * a.hlen)), dim3(256), 0, st, src, a.src_lp,
                           src_dist, len, step, a.hlen, nblocks, g0, cnt, hbuf, hbuf + cnt * a.hlen);
        HIPCHK(hipGetLastError());
        launch(a, st);
        HIPCHK(hipGetLastError());
    }
    return ws ? PCX_SUCCESS : xc_own_done(x, st);
}

} // namespace

extern "C" int pcx_fxcorr_exec_batch(pcx_fxcorr_t x, float* dst, int64_t dst_pack, int64_t dst_dist, const float* src,
                                     int64_t src_pack, int64_t src_dist, int64_t len, int64_t channels, void* stream) {
    return fxcorr_run(x, dst, dst_pack, dst_dist, src, src_pack, src_dist, len, channels, nullptr, (hipStream_t)stream);
}

extern "C" int pcx_fxcorr_exec(pcx_fxcorr_t x, float* dst, int64_t dst_pack, const float* src, int64_t src_pack,
                               int64_t len, void* stream) {
    return pcx_fxcorr_exec_batch(x, dst, dst_pack, 0, src, src_pack, 0, len, 1, stream);
}

extern "C" int pcx_fxcorr_exec_batch_ws(pcx_fxcorr_t x, float* dst, int64_t dst_pack, int64_t dst_dist,
                                        const float* src, int64_t src_pack, int64_t src_dist, int64_t len,
                                        int64_t channels, void* workspace, int64_t workspace_bytes, void* stream) {
    const ExtTmp ws{workspace, workspace_bytes};
    return fxcorr_run(x, dst, dst_pack, dst_dist, src, src_pack, src_dist, len, channels, &ws, (hipStream_t)stream);
}

// ---------------------------------------------------------------- memory and utilities
extern "C" int pcx_malloc(void** ptr, size_t bytes) {
    if (!ptr)
        return fail(PCX_ERR_INVALID_ARGUMENT, "pointer is null");
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess)
        return fail(PCX_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    return PCX_SUCCESS;
}
extern "C" int pcx_free(void* ptr) {
    HIPCHK(hipFree(ptr));
    return PCX_SUCCESS;
}
extern "C" int pcx_memcpy_htod(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PCX_SUCCESS;
}
extern "C" int pcx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PCX_SUCCESS;
}
extern "C" int pcx_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return PCX_SUCCESS;
}
extern "C" int pcx_memset(void* dst, int value, size_t bytes) {
    HIPCHK(hipMemset(dst, value, bytes));
    return PCX_SUCCESS;
}
extern "C" int pcx_synchronize(void* stream) {
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return PCX_SUCCESS;
}
extern "C" int pcx_generate(float* dst, int64_t pack, int64_t n, uint64_t seed, int dist, double sigma, void* stream) {
    if (!dst)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (int e = check_pack(pack, "destination"))
        return e;
    if (n <= 0)
        return PCX_SUCCESS;
    int64_t nb = std::min<int64_t>((n + 255) / 256, (int64_t)1 << 20);
    hipLaunchKernelGGL(k_generate<float>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, dst, ilog2(pack), n, seed,
                       dist, sigma);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}
extern "C" int pcx_generate_f64(double* dst, int64_t pack, int64_t n, uint64_t seed, int dist, double sigma,
                                void* stream) {
    if (!dst)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if (int e = check_pack(pack, "destination"))
        return e;
    if (n <= 0)
        return PCX_SUCCESS;
    int64_t nb = std::min<int64_t>((n + 255) / 256, (int64_t)1 << 20);
    hipLaunchKernelGGL(k_generate<double>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, dst, ilog2(pack), n,
                       seed, dist, sigma);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}

// ---------------------------------------------------------------- uneven-batch work queue
namespace {

struct QGroup { // transforms of one size above 2^20 and one direction: grouped four-step launches
    int      lg   = 0;
    int      inv  = 0;
    int64_t  count = 0;
    int64_t* src_offs = nullptr; // device
    int64_t* dst_offs = nullptr;
    pcx_plan_t plan = nullptr;
};

} // namespace

struct pcx_queue_s {
    int                 normalized = 1;
    int64_t             njobs      = 0;
    double              flops = 0, bytes = 0;
    // persistent part: every transform of 2^1 .. 2^20 points
    int32_t             nitems     = 0;
    QItem*              items      = nullptr; // device
    int64_t*            q_src_offs = nullptr;
    int64_t*            q_dst_offs = nullptr;
    uint32_t*           counter    = nullptr;
    float2*             tw         = nullptr; // W_A for A = 2 .. 2^14
    float2*             ptw        = nullptr; // pass tables of 2^15 .. 2^20
    int32_t             ptw_off[kQueueMaxLog2 + 1] = {};
    float*              tmp        = nullptr; // four-step intermediates, one per workgroup
    int64_t             tmp_floats = 0;
    int                 grid       = 0;       // workgroups of the persistent launch
    int                 grid_cap   = 0;       // workgroups the intermediates were sized for
    // grouped part: sizes above 2^20, and the classes the autotuner moved out of the persistent
    // launch (grouped[log2n][inverse])
    std::vector<QGroup> groups;
    bool                grouped[kMaxLog2 + 1][2] = {};
    std::vector<pcx_job> jobs; // the job list, to rebuild the items when the split changes
    // captured graph
    hipGraph_t          graph = nullptr;
    hipGraphExec_t      exec  = nullptr;
};

namespace {

int queue_run(pcx_queue_t q, float* dst, const float* src, hipStream_t st) {
    if (q->nitems > 0) {
        launch_reset_counter(q->counter, st);
        QueueArgs a{};
        a.src = src, a.dst = dst, a.tmp = q->tmp, a.tmp_floats = q->tmp_floats;
        a.src_offs = q->q_src_offs, a.dst_offs = q->q_dst_offs;
        a.items = q->items, a.nitems = q->nitems, a.normalized = q->normalized, a.counter = q->counter;
        a.tw = q->tw, a.ptw = q->ptw;
        std::memcpy(a.ptw_off, q->ptw_off, sizeof(a.ptw_off));
        launch_queue(a, std::min<int>(q->grid, q->nitems), st);
        HIPCHK(hipGetLastError());
    }
    for (auto& g : q->groups) {
        int64_t n = (int64_t)1 << g.lg;
        if (int e = run_batch(g.plan, g.inv, q->normalized, dst, n, 2 * n, src, n, 2 * n, n, g.count, st, g.src_offs,
                              g.dst_offs))
            return e;
    }
    return PCX_SUCCESS;
}

// Frees the device side of the item list and the groups (the job list and settings stay).
void queue_release(pcx_queue_t q) {
    if (q->exec)
        (void)hipGraphExecDestroy(q->exec);
    if (q->graph)
        (void)hipGraphDestroy(q->graph);
    q->exec = nullptr, q->graph = nullptr;
    (void)hipFree(q->items);
    (void)hipFree(q->q_src_offs);
    (void)hipFree(q->q_dst_offs);
    (void)hipFree(q->counter);
    (void)hipFree(q->tw);
    (void)hipFree(q->ptw);
    (void)hipFree(q->tmp);
    q->items = nullptr, q->q_src_offs = q->q_dst_offs = nullptr, q->counter = nullptr;
    q->tw = q->ptw = nullptr, q->tmp = nullptr, q->tmp_floats = 0, q->nitems = 0;
    for (auto& g : q->groups) {
        (void)hipFree(g.src_offs);
        (void)hipFree(g.dst_offs);
        pcx_plan_destroy(g.plan);
    }
    q->groups.clear();
}

void queue_free(pcx_queue_t q) {
    queue_release(q);
    delete q;
}

// Builds the persistent items (every class up to 2^20 not marked grouped, largest first) and the
// grouped launches (the rest, largest first) from the job list.
int queue_build(pcx_queue_t q) {
    const pcx_job* jobs  = q->jobs.data();
    const int64_t  njobs = (int64_t)q->jobs.size();
    std::vector<std::vector<int64_t>> bucket((kMaxLog2 + 1) * 2);
    for (int64_t i = 0; i < njobs; ++i)
        bucket[jobs[i].log2n * 2 + jobs[i].inverse].push_back(i);
    std::vector<QItem>   items;
    std::vector<int64_t> qs, qd;
    int                  max_four = 0;
    // persistent items, largest size first (longest-processing-time order)
    for (int lg = kQueueMaxLog2; lg >= 1; --lg) {
        const int per = lg >= kQueueMinFour ? 1 : kQueueThreads / (lg >= 4 ? (1 << lg) / 16 : 1);
        for (int inv = 0; inv < 2; ++inv) {
            if (q->grouped[lg][inv])
                continue;
            auto& b = bucket[lg * 2 + inv];
            for (size_t k = 0; k < b.size(); k += per) {
                int cnt = (int)std::min<size_t>(per, b.size() - k);
                items.push_back({lg >= kQueueMinFour ? QK_FOUR : QK_SINGLE, lg, inv, (int32_t)qs.size(), cnt});
                for (int c = 0; c < cnt; ++c) {
                    qs.push_back(jobs[b[k + c]].src_offset);
                    qd.push_back(jobs[b[k + c]].dst_offset);
                }
                if (lg >= kQueueMinFour)
                    max_four = std::max(max_four, lg);
            }
        }
    }
    q->nitems = (int32_t)items.size();
    hipStream_t setup = nullptr;
    hipError_t  he    = setup_stream(&setup);
    if (!items.empty() && he == hipSuccess) {
        he = hipMalloc(&q->items, sizeof(QItem) * items.size());
        if (he == hipSuccess)
            he = hipMalloc(&q->q_src_offs, sizeof(int64_t) * qs.size());
        if (he == hipSuccess)
            he = hipMalloc(&q->q_dst_offs, sizeof(int64_t) * qd.size());
        if (he == hipSuccess)
            he = hipMalloc(&q->counter, sizeof(uint32_t));
        if (he == hipSuccess)
            he = hipMalloc(&q->tw, sizeof(float2) * ((1 << (kQueueMaxSingle + 1)) - 2));
        if (he == hipSuccess)
            he = hipMemcpy(q->items, items.data(), sizeof(QItem) * items.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy(q->q_src_offs, qs.data(), sizeof(int64_t) * qs.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy(q->q_dst_offs, qd.data(), sizeof(int64_t) * qd.size(), hipMemcpyHostToDevice);
        for (int lg = 1; lg <= kQueueMaxSingle && he == hipSuccess; ++lg)
            he = gen_twiddles(q->tw + ((1 << lg) - 2), (int64_t)1 << lg, 1, (int64_t)1 << lg, setup);
        if (max_four > 0 && he == hipSuccess) {
            // pass tables W_N^e = lo[e mod 2^ls] * hi[e >> ls] of every four-step size, and one
            // intermediate of the largest size per workgroup
            int64_t total = 0;
            for (int lg = kQueueMinFour; lg <= max_four; ++lg) {
                q->ptw_off[lg] = (int32_t)total;
                total += ((int64_t)1 << ((lg + 1) / 2)) + ((int64_t)1 << (lg - (lg + 1) / 2));
            }
            he = hipMalloc(&q->ptw, sizeof(float2) * total);
            for (int lg = kQueueMinFour; lg <= max_four && he == hipSuccess; ++lg) {
                const int     ls = (lg + 1) / 2;
                const int64_t nl = (int64_t)1 << ls, nh = ((int64_t)1 << lg) >> ls;
                he = gen_twiddles(q->ptw + q->ptw_off[lg], (int64_t)1 << lg, 1, nl, setup);
                if (he == hipSuccess)
                    he = gen_twiddles(q->ptw + q->ptw_off[lg] + nl, (int64_t)1 << lg, nl, nh, setup);
            }
            q->tmp_floats = (int64_t)2 << max_four;
            if (he == hipSuccess)
                he = hipMalloc(&q->tmp, sizeof(float) * q->tmp_floats * q->grid_cap);
        }
    }
    // grouped launches, largest first
    for (int lg = kMaxLog2; lg >= 1 && he == hipSuccess; --lg) {
        for (int inv = 0; inv < 2; ++inv) {
            auto& b = bucket[lg * 2 + inv];
            if (b.empty() || (lg <= kQueueMaxLog2 && !q->grouped[lg][inv]))
                continue;
            QGroup g;
            g.lg = lg, g.inv = inv, g.count = (int64_t)b.size();
            std::vector<int64_t> so(b.size()), dof(b.size());
            for (size_t k = 0; k < b.size(); ++k)
                so[k] = jobs[b[k]].src_offset, dof[k] = jobs[b[k]].dst_offset;
            he = hipMalloc(&g.src_offs, sizeof(int64_t) * b.size());
            if (he == hipSuccess)
                he = hipMalloc(&g.dst_offs, sizeof(int64_t) * b.size());
            if (he == hipSuccess)
                he = hipMemcpy(g.src_offs, so.data(), sizeof(int64_t) * b.size(), hipMemcpyHostToDevice);
            if (he == hipSuccess)
                he = hipMemcpy(g.dst_offs, dof.data(), sizeof(int64_t) * b.size(), hipMemcpyHostToDevice);
            q->groups.push_back(g);
            if (he == hipSuccess)
                if (int e = pcx_plan_create((int64_t)1 << lg, 2048, PCX_ORDER_NORMAL, &q->groups.back().plan))
                    return e;
        }
    }
    if (he == hipSuccess)
        he = hipStreamSynchronize(setup);
    if (he != hipSuccess)
        return fail(PCX_ERR_HIP, std::string("queue setup: ") + hipGetErrorString(he));
    return PCX_SUCCESS;
}

} // namespace

extern "C" int pcx_queue_create(const pcx_job* jobs, int64_t njobs, int normalized, pcx_queue_t* out) {
    if (!out)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue output pointer is null");
    *out = nullptr;
    if (njobs < 0 || (njobs > 0 && !jobs))
        return fail(PCX_ERR_INVALID_ARGUMENT, "job list is null or negative");
    for (int64_t i = 0; i < njobs; ++i) {
        if (jobs[i].log2n < 1 || jobs[i].log2n > kMaxLog2)
            return fail(PCX_ERR_INVALID_SIZE, "fft_size (which is  2^" + s(jobs[i].log2n) + ") is not supported");
        if (jobs[i].inverse != 0 && jobs[i].inverse != 1)
            return fail(PCX_ERR_INVALID_ARGUMENT, "job " + s(i) + ": inverse must be 0 or 1");
        if (jobs[i].src_offset < 0 || jobs[i].dst_offset < 0)
            return fail(PCX_ERR_INVALID_ARGUMENT, "job " + s(i) + ": negative offset");
    }
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    auto* q       = new pcx_queue_s;
    q->normalized = normalized ? 1 : 0;
    q->njobs      = njobs;
    // one workgroup per CU (the largest item takes most of a CU's LDS); the autotuner may pick fewer
    q->grid = q->grid_cap = std::max(1, ncu);
    for (int64_t i = 0; i < njobs; ++i) {
        double n = (double)((int64_t)1 << jobs[i].log2n);
        q->flops += 5.0 * n * (double)jobs[i].log2n;
        q->bytes += 16.0 * n;
    }
    q->jobs.assign(jobs, jobs + njobs);
    if (int e = queue_build(q)) {
        queue_free(q);
        return e;
    }
    *out = q;
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_destroy(pcx_queue_t q) {
    if (q)
        queue_free(q);
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_exec(pcx_queue_t q, float* dst, const float* src, void* stream) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    return queue_run(q, dst, src, (hipStream_t)stream);
}

extern "C" int pcx_queue_capture(pcx_queue_t q, float* dst, const float* src, void* stream) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    hipStream_t st = (hipStream_t)stream;
    if (!st)
        return fail(PCX_ERR_INVALID_ARGUMENT, "graph capture needs a non-default stream");
    // make sure every group's intermediate exists before capture (allocation is not capturable)
    for (auto& g : q->groups) {
        int64_t n = (int64_t)1 << g.lg;
        if (g.plan->two_pass && g.plan->tmp_cap < std::min(g.count, g.plan->chunk)) {
            if (int e = run_batch(g.plan, g.inv, q->normalized, dst, n, 2 * n, src, n, 2 * n, n, g.count, st,
                                  g.src_offs, g.dst_offs))
                return e;
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    if (q->exec) {
        HIPCHK(hipGraphExecDestroy(q->exec));
        q->exec = nullptr;
    }
    if (q->graph) {
        HIPCHK(hipGraphDestroy(q->graph));
        q->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int        e  = queue_run(q, dst, src, st);
    hipGraph_t gr = nullptr;
    hipError_t he = hipStreamEndCapture(st, &gr);
    if (e)
        return e;
    if (he != hipSuccess)
        return fail(PCX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    q->graph = gr;
    HIPCHK(hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0));
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_launch(pcx_queue_t q, void* stream) {
    if (!q || !q->exec)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue has no captured graph");
    HIPCHK(hipGraphLaunch(q->exec, (hipStream_t)stream));
    return PCX_SUCCESS;
}

extern "C" int pcx_queue_work(pcx_queue_t q, double* flops, double* bytes) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (flops)
        *flops = q->flops;
    if (bytes)
        *bytes = q->bytes;
    return PCX_SUCCESS;
}

namespace {

// Times fn (which enqueues work on st) with events: best of `reps` after one warm-up.
template<class F>
float time_ms(hipStream_t st, F&& fn, int reps = 3) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess)
        return 1e30f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return 1e30f;
    }
    float best = 1e30f;
    if (fn() != PCX_SUCCESS)
        best = -1.f;
    for (int r = 0; r < reps && best >= 0.f; ++r) {
        float ms = 0.f;
        (void)hipEventRecord(a, st);
        if (fn() != PCX_SUCCESS) {
            best = -1.f;
            break;
        }
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        best = ms < best ? ms : best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best < 0.f ? 1e30f : best;
}

} // namespace

// Autotuner: workgroups of the persistent launch, then per grouped size the four-step shape and
// grid caps.  Winners are written to cache_path (one "key value..." line each) and read back
// from it when present.
extern "C" int pcx_queue_autotune(pcx_queue_t q, float* dst, const float* src, void* stream, const char* cache_path) {
    if (!q)
        return fail(PCX_ERR_INVALID_ARGUMENT, "queue is null");
    if (!dst || !src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "data pointer is null");
    if ((const float*)dst == src)
        return fail(PCX_ERR_INVALID_ARGUMENT, "autotuning overwrites dst many times: it must not alias src");
    hipStream_t st = (hipStream_t)stream;
    // a captured graph holds the kernels, grids and twiddle tables of the shapes it was captured
    // with; the tuning below may change all of them, so the graph is dropped (capture again after)
    HIPCHK(hipDeviceSynchronize());
    if (q->exec) {
        HIPCHK(hipGraphExecDestroy(q->exec));
        q->exec = nullptr;
    }
    if (q->graph) {
        HIPCHK(hipGraphDestroy(q->graph));
        q->graph = nullptr;
    }
    if (cache_path) { // reuse earlier decisions
        if (FILE* f = std::fopen(cache_path, "r")) {
            char key[32];
            int  a = 0, b = 0, c = 0, d = 0, e = 0;
            long long gc = 0, gr = 0;
            // the split between the persistent launch and grouped launches comes first
            bool mask[kMaxLog2 + 1][2] = {};
            long pos = std::ftell(f);
            while (std::fscanf(f, "%31s", key) == 1) {
                if (!std::strcmp(key, "grouped") && std::fscanf(f, "%d %d", &a, &b) == 2 && a >= 1 && a <= kQueueMaxLog2 &&
                    (b == 0 || b == 1))
                    mask[a][b] = true;
            }
            std::fseek(f, pos, SEEK_SET);
            if (std::memcmp(mask, q->grouped, sizeof(mask)) != 0) {
                std::memcpy(q->grouped, mask, sizeof(mask));
                queue_release(q);
                if (int err = queue_build(q)) {
                    std::fclose(f);
                    return err;
                }
            }
            while (std::fscanf(f, "%31s", key) == 1) {
                if (!std::strcmp(key, "grouped") && std::fscanf(f, "%d %d", &a, &b) == 2) {
                    continue;
                } else if (!std::strcmp(key, "queue_grid") && std::fscanf(f, "%d", &a) == 1) {
                    q->grid = std::max(1, std::min(a, q->grid_cap));
                } else if (!std::strcmp(key, "group") && std::fscanf(f, "%d %d %d %d %d %lld %lld", &a, &b, &c, &d, &e, &gc, &gr) == 7) {
                    for (auto& g : q->groups)
                        if (g.lg == a && g.inv == b && g.plan->two_pass) {
                            if (const FourStep* sh = find_shape(a, c, d, e)) {
                                if (int err = plan_apply_shape(g.plan, sh)) {
                                    std::fclose(f);
                                    return err;
                                }
                                g.plan->grid_col = gc, g.plan->grid_row = gr;
                            }
                        }
                } else {
                    break;
                }
            }
            std::fclose(f);
            return PCX_SUCCESS;
        }
    }
    if (q->nitems > 0) {
        float best = 1e30f;
        int   bg   = q->grid;
        for (int grid : {q->grid_cap / 2, (3 * q->grid_cap) / 4, q->grid_cap}) {
            if (grid < 1)
                continue;
            q->grid  = grid;
            float ms = time_ms(st, [&] {
                std::vector<QGroup> keep;
                keep.swap(q->groups); // time the persistent part alone
                int e = queue_run(q, dst, src, st);
                keep.swap(q->groups);
                return e;
            });
            if (ms < best)
                best = ms, bg = grid;
        }
        q->grid = bg;
    }
    // which size classes run in the persistent launch and which as grouped launches after it:
    // each four-step class (largest first), then the single-pass classes together, moves out when
    // that makes the whole pass faster
    {
        bool present[kMaxLog2 + 1] = {};
        for (const auto& j : q->jobs)
            present[j.log2n] = true;
        auto whole = [&] { return queue_run(q, dst, src, st); };
        auto rebuild = [&]() -> int {
            queue_release(q);
            return queue_build(q);
        };
        float best = time_ms(st, whole);
        auto  trial = [&](int lo, int hi) -> int {
            bool any = false;
            for (int lg = lo; lg <= hi; ++lg)
                any = any || (present[lg] && !q->grouped[lg][0]);
            if (!any)
                return PCX_SUCCESS;
            bool keep[kMaxLog2 + 1][2];
            std::memcpy(keep, q->grouped, sizeof(keep));
            for (int lg = lo; lg <= hi; ++lg)
                q->grouped[lg][0] = q->grouped[lg][1] = present[lg];
            if (int e = rebuild())
                return e;
            const float ms = time_ms(st, whole);
            if (ms < best) {
                best = ms;
                return PCX_SUCCESS;
            }
            std::memcpy(q->grouped, keep, sizeof(keep));
            return rebuild();
        };
        for (int lg = kQueueMaxLog2; lg >= kQueueMinFour; --lg)
            if (int e = trial(lg, lg))
                return e;
        if (int e = trial(1, kQueueMaxSingle))
            return e;
    }
    for (auto& g : q->groups) {
        if (!g.plan->two_pass)
            continue;
        int64_t         n    = (int64_t)1 << g.lg;
        const FourStep* best = g.plan->shape;
        int64_t         bgc = g.plan->grid_col, bgr = g.plan->grid_row;
        float           bms = 1e30f;
        std::vector<const FourStep*> cands;
        for (const ShapeList<float>& l : {shapes_default(), shapes_alt()})
            for (int i = 0; i < l.count; ++i)
                if (l.data[i].lg == g.lg)
                    cands.push_back(&l.data[i]);
        for (const FourStep* shp : cands) {
            const FourStep& sh = *shp;
            if (int err = plan_apply_shape(g.plan, &sh))
                return err;
            for (int64_t gc : {(int64_t)0, (int64_t)1024, (int64_t)2048}) {
                g.plan->grid_col = g.plan->grid_row = gc;
                float ms = time_ms(st, [&] {
                    return run_batch(g.plan, g.inv, q->normalized, dst, n, 2 * n, src, n, 2 * n, n, g.count, st,
                                     g.src_offs, g.dst_offs);
                });
                if (ms < bms)
                    bms = ms, best = &sh, bgc = gc, bgr = gc;
            }
        }
        if (int err = plan_apply_shape(g.plan, best))
            return err;
        g.plan->grid_col = bgc, g.plan->grid_row = bgr;
    }
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "w")) {
            for (int lg = 1; lg <= kQueueMaxLog2; ++lg)
                for (int inv = 0; inv < 2; ++inv)
                    if (q->grouped[lg][inv])
                        std::fprintf(f, "grouped %d %d\n", lg, inv);
            std::fprintf(f, "queue_grid %d\n", q->grid);
            for (auto& g : q->groups)
                if (g.plan->two_pass)
                    std::fprintf(f, "group %d %d %d %d %d %lld %lld\n", g.lg, g.inv, g.plan->shape->n2, g.plan->shape->g,
                                 g.plan->shape->c, (long long)g.plan->grid_col, (long long)g.plan->grid_row);
            std::fclose(f);
        }
    }
    return PCX_SUCCESS;
}

// ---------------------------------------------------------------- vector expressions
namespace {
// One validation and launch path for both element types (S, I: the C ABI's source and
// instruction structs of that type).
template<class T, class S, class I>
int vexpr_eval(const S* srcs, int32_t nsrc, const I* prog, int32_t nprog, T* dst, int64_t dst_pack, int64_t n,
               void* stream) {
    if (nsrc < 0 || nsrc > PCX_VEXPR_MAX_SRC || (nsrc > 0 && !srcs))
        return fail(PCX_ERR_INVALID_ARGUMENT, "source count (which is " + s(nsrc) + ") is not in [0, 16]");
    if (nprog < 1 || nprog > PCX_VEXPR_MAX_PROG || !prog)
        return fail(PCX_ERR_INVALID_ARGUMENT, "program length (which is " + s(nprog) + ") is not in [1, 64]");
    if (!dst)
        return fail(PCX_ERR_INVALID_ARGUMENT, "destination pointer is null");
    if (n < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "size (which is " + s(n) + ") is negative");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    VExprArgsT<T> a{};
    for (int k = 0; k < nsrc; ++k) {
        if (!srcs[k].data)
            return fail(PCX_ERR_INVALID_ARGUMENT, "source " + s(k) + " is null");
        if (int e = check_pack(srcs[k].pack, "source"))
            return e;
        if ((const T*)dst == srcs[k].data && srcs[k].pack != dst_pack)
            return fail(PCX_ERR_INPLACE_REPACK, "cannot perform in-place repack from source pack size (which is " +
                                                    s(srcs[k].pack) + ") to destination pack size (which is " +
                                                    s(dst_pack) + ")");
        a.src[k].data = srcs[k].data;
        a.src[k].lp   = ilog2(srcs[k].pack);
    }
    for (int k = 0; k < nprog; ++k) {
        const I& in = prog[k];
        bool     ok = in.op >= 0 && in.op < VOP_COUNT && in.d >= 0 && in.d < PCX_VEXPR_REGS;
        if (in.op == VOP_LOAD)
            ok = ok && in.a >= 0 && in.a < nsrc;
        else
            ok = ok && in.a >= 0 && in.a < PCX_VEXPR_REGS && (in.b >= 0 && in.b < PCX_VEXPR_REGS);
        if (!ok)
            return fail(PCX_ERR_INVALID_ARGUMENT, "instruction " + s(k) + " is malformed");
        a.prog[k] = VInstrT<T>{in.op, in.d, in.a, in.b, in.s_re, in.s_im};
    }
    a.nprog  = nprog;
    a.dst    = dst;
    a.dst_lp = ilog2(dst_pack);
    a.n      = n;
    if (n == 0)
        return PCX_SUCCESS;
    int64_t nwg = (n + 255) / 256;
    if (nwg > 16384)
        nwg = 16384;
    launch_vexpr(a, nwg, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}
} // namespace

extern "C" int pcx_vexpr_eval(const pcx_vsource* srcs, int32_t nsrc, const pcx_vinstr* prog, int32_t nprog, float* dst,
                              int64_t dst_pack, int64_t n, void* stream) {
    return vexpr_eval(srcs, nsrc, prog, nprog, dst, dst_pack, n, stream);
}

extern "C" int pcx_repack(float* dst, int64_t dst_pack, const float* src, int64_t src_pack, int64_t n, void* stream) {
    pcx_vsource sv{src, src_pack};
    pcx_vinstr  ld{PCX_VOP_LOAD, 0, 0, 0, 0.f, 0.f};
    return pcx_vexpr_eval(&sv, 1, &ld, 1, dst, dst_pack, n, stream);
}

extern "C" int pcx_vexpr_eval_f64(const pcx_vsource_f64* srcs, int32_t nsrc, const pcx_vinstr_f64* prog, int32_t nprog,
                                  double* dst, int64_t dst_pack, int64_t n, void* stream) {
    return vexpr_eval(srcs, nsrc, prog, nprog, dst, dst_pack, n, stream);
}

extern "C" int pcx_repack_f64(double* dst, int64_t dst_pack, const double* src, int64_t src_pack, int64_t n,
                              void* stream) {
    pcx_vsource_f64 sv{src, src_pack};
    pcx_vinstr_f64  ld{PCX_VOP_LOAD, 0, 0, 0, 0.0, 0.0};
    return pcx_vexpr_eval_f64(&sv, 1, &ld, 1, dst, dst_pack, n, stream);
}

// ---------------------------------------------------------------- FFT across vectors
namespace {
template<class T, class Tab>
int fft_across(pcx_plan_t p, int inverse, int normalized, T* const* dst_vecs, int64_t dst_pack, const T* const* src_vecs,
               int64_t src_pack, int64_t length, hipStream_t st, const Tab& tab) {
    constexpr bool F64 = std::is_same_v<T, double>;
    if (!p)
        return fail(PCX_ERR_INVALID_ARGUMENT, "plan is null");
    if (p->f64 != F64)
        return fail(PCX_ERR_INVALID_ARGUMENT, p->f64 ? "the plan is for double data (use pcx_fft_across_f64)"
                                                     : "the plan is for float data (use pcx_fft_across)");
    if (!dst_vecs || !src_vecs)
        return fail(PCX_ERR_INVALID_ARGUMENT, "vector table is null");
    if (length < 0)
        return fail(PCX_ERR_INVALID_ARGUMENT, "length (which is " + s(length) + ") is negative");
    if (int e = check_pack(dst_pack, "destination"))
        return e;
    if (int e = check_pack(src_pack, "source"))
        return e;
    if (p->order != PCX_ORDER_NORMAL)
        return fail(PCX_ERR_UNSUPPORTED, "transforms across vectors are built for the normal order");
    if (p->lg > (int)tab.size())
        return fail(PCX_ERR_UNSUPPORTED, "transforms across vectors cover at most " + s((int64_t)1 << tab.size()) +
                                             " vectors (fft_size " + s(p->n) + ")");
    if (length == 0)
        return PCX_SUCCESS;
    AcrossArgsT<T> a{};
    a.src = src_vecs, a.dst = dst_vecs, a.len = length;
    a.src_lp = ilog2(src_pack), a.dst_lp = ilog2(dst_pack);
    a.conj_in = inverse, a.conj_out = inverse;
    a.scale   = (inverse && normalized) ? (T)(1.0 / (double)p->n) : T(1);
    if constexpr (F64)
        a.twA = p->tw64_row;
    else
        a.twA = p->tw_row;
    tab[p->lg - 1](a, 2048, st);
    HIPCHK(hipGetLastError());
    return PCX_SUCCESS;
}
} // namespace

extern "C" int pcx_fft_across(pcx_plan_t p, int inverse, int normalized, float* const* dst_vecs, int64_t dst_pack,
                              const float* const* src_vecs, int64_t src_pack, int64_t length, void* stream) {
    return fft_across(p, inverse, normalized, dst_vecs, dst_pack, src_vecs, src_pack, length, (hipStream_t)stream,
                      across_tab());
}

extern "C" int pcx_fft_across_f64(pcx_plan_t p, int inverse, int normalized, double* const* dst_vecs,
                                  int64_t dst_pack, const double* const* src_vecs, int64_t src_pack, int64_t length,
                                  void* stream) {
    return fft_across(p, inverse, normalized, dst_vecs, dst_pack, src_vecs, src_pack, length, (hipStream_t)stream,
                      across64_tab());
}
