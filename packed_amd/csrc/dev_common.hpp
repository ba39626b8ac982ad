// dev_common.hpp -- device-side complex arithmetic, in-register DFTs and output-order maps for
// the gfx950 FFT kernels.  Split-complex throughout: a complex value is (re, im) in two VGPRs.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace pcxk {

// A complex value of element type T (float or double).
template<class T>
struct cx {
    T x, y;
};
using cf = cx<float>;
using cd = cx<double>;

__device__ __forceinline__ float  fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }

template<class T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template<class T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
// a * w
template<class T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> w) {
    return {fmat(a.x, w.x, -a.y * w.y), fmat(a.x, w.y, a.y * w.x)};
}
// a * conj(w)
template<class T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> w) {
    return {fmat(a.x, w.x, a.y * w.y), fmat(a.y, w.x, -a.x * w.y)};
}
template<class T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.x, -a.y}; }

// Two-component vector of T as stored in twiddle tables (float2 / double2).
template<class T>
struct vec2_of;
template<>
struct vec2_of<float> {
    using type = float2;
};
template<>
struct vec2_of<double> {
    using type = double2;
};
template<class T>
using vec2 = typename vec2_of<T>::type;

// Compile-time loop.
template<int I, int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int brevc(int v, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b)
        r |= ((v >> b) & 1) << (bits - 1 - b);
    return r;
}

// cos(pi*m/16) for m = 0..8, to 20 digits.
constexpr double kCos16[9] = {1.0,
                              0.98078528040323044913,
                              0.92387953251128675613,
                              0.83146961230254523708,
                              0.70710678118654752440,
                              0.55557023301960222474,
                              0.38268343236508977173,
                              0.19509032201612826785,
                              0.0};
// cos(2*pi*k/32) = cos(pi*k/16) for k in [0, 32)
constexpr double cos32(int k) {
    k &= 31;
    if (k <= 8)
        return kCos16[k];
    if (k <= 16)
        return -kCos16[16 - k];
    if (k <= 24)
        return -kCos16[k - 16];
    return kCos16[32 - k];
}
constexpr double sin32(int k) { return cos32(k - 8); } // sin(a) = cos(a - pi/2)

// z * exp(-2*pi*i*K/N) for N | 32, with the trivial factors done without multiplies.
template<int N, int K, class T>
__device__ __forceinline__ cx<T> twc(cx<T> z) {
    constexpr int k = ((K % N) + N) % N;
    if constexpr (k == 0) {
        return z;
    } else if constexpr (4 * k == N) {
        return {z.y, -z.x};
    } else if constexpr (2 * k == N) {
        return {-z.x, -z.y};
    } else if constexpr (4 * k == 3 * N) {
        return {-z.y, z.x};
    } else if constexpr (8 * k == N) {
        constexpr T r = (T)0.70710678118654752440;
        return {(z.x + z.y) * r, (z.y - z.x) * r};
    } else if constexpr (8 * k == 3 * N) {
        constexpr T r = (T)0.70710678118654752440;
        return {(z.y - z.x) * r, -(z.x + z.y) * r};
    } else {
        constexpr int k32 = k * (32 / N);
        constexpr T   c   = (T)cos32(k32);
        constexpr T   s   = (T)(-sin32(k32));
        return cmul(z, cx<T>{c, s});
    }
}

// In-register forward DFT of size R (power of two, R <= 32): natural order in, natural order out.
// Radix-2 decimation in frequency with compile-time twiddles, then a register renaming.
template<int R, class T>
__device__ __forceinline__ void dft(cx<T>* v) {
    constexpr int L = ilog2c(R);
    sfor<0, L>([&](auto ss) {
        constexpr int s = decltype(ss)::value;
        constexpr int h = R >> (s + 1);
        sfor<0, R / 2>([&](auto bb) {
            constexpr int b = decltype(bb)::value;
            constexpr int g = (b / h) * 2 * h;
            constexpr int j = b % h;
            cx<T> a = v[g + j], c = v[g + j + h];
            v[g + j]     = cadd(a, c);
            v[g + j + h] = twc<2 * h, j>(csub(a, c));
        });
    });
    cx<T> t[R];
    sfor<0, R>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        t[k]            = v[brevc(k, L)];
    });
    sfor<0, R>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        v[k]            = t[k];
    });
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t v, int bits) {
    return bits == 0 ? 0u : (__builtin_bitreverse32(v) >> (32 - bits));
}
// unordered 32-point lane permutation q: (b4,b3,b2,b1,b0) -> (b4,b2,b0,b1,b3)
__device__ __forceinline__ uint32_t uq(uint32_t i) {
    return (i & ~31u) | (i & 16u) | ((i & 4u) << 1) | ((i & 1u) << 2) | (i & 2u) | ((i >> 3) & 1u);
}

// Output orders (fft.hpp:540-544): slot i holds frequency slot_freq(i).
enum : int { ORD_NORMAL = 0, ORD_BITREV = 1, ORD_UNORD = 2 };
__device__ __forceinline__ uint32_t slot_freq(int order, uint32_t i, int bits) {
    if (order == ORD_BITREV)
        return brev_bits(i, bits);
    if (order == ORD_UNORD)
        return brev_bits(uq(i), bits);
    return i;
}

// Packed layout (element_access.hpp:8-11): real part of complex i at i + (i / P) * P, imaginary
// part P elements later.  lp = log2(P).
__device__ __forceinline__ int64_t pidx(int64_t i, int lp) { return i + ((i >> lp) << lp); }

} // namespace pcxk
