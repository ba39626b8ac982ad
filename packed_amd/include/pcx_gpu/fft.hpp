// pcx_gpu/fft.hpp -- C++20 host interface of the gfx950 FFT path, mirroring the reference's
// fft_unit (fft.hpp:565-915) and fxcorr_unit (fxcorr.hpp:44-147) template surface over the C ABI
// in pcx_gpu.h.  Header-only; link libpcx_gpu.so.  No HIP headers are needed: streams are passed
// as `void*` (a hipStream_t).
//
// Data lives in device memory in the packed split-complex layout of vector.hpp: complex i of a
// pack-P vector has its real part at element offset i + (i / P) * P, its imaginary part P
// elements later.  pcx_gpu::vector<T, P> is that container with a device allocator; raw-pointer
// members (fft_raw / ifft_raw) take device pointers.  T is float or double (fft_unit<double>
// covers the normal and bit_reversed orders).
//
// Errors throw what the reference throws: std::invalid_argument with the reference's message for
// a bad size, sub-size, size mismatch or in-place repack; std::runtime_error for a HIP failure and
// std::bad_alloc when device memory runs out.
#pragma once

#include "pcx_gpu.h"

#include <complex>
#include <concepts>
#include <cstddef>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pcx_gpu {

using uZ = std::size_t;

// fft.hpp:540-544
enum class fft_order { normal, bit_reversed, unordered };

inline void throw_on(int rc) {
    if (rc == PCX_SUCCESS)
        return;
    std::string msg = pcx_last_error();
    switch (rc) {
    case PCX_ERR_OUT_OF_MEMORY: throw std::bad_alloc();
    case PCX_ERR_HIP: throw std::runtime_error(msg);
    default: throw std::invalid_argument(msg);
    }
}

template<uZ P>
concept pack_size = P > 0 && (P & (P - 1)) == 0;

template<typename T>
concept gpu_real = std::same_as<T, float> || std::same_as<T, double>;

// types.hpp:70-72
template<gpu_real T>
constexpr uZ default_pack_size = 32 / sizeof(T);

// The C ABI by element type.
namespace detail {
inline int plan_create(float*, int64_t n, int64_t sub, int order, pcx_plan_t* p) {
    return pcx_plan_create(n, sub, order, p);
}
inline int plan_create(double*, int64_t n, int64_t sub, int order, pcx_plan_t* p) {
    return pcx_plan_create_f64(n, sub, order, p);
}
inline int fft(pcx_plan_t p, float* d, int64_t dp, const float* s, int64_t sp, int64_t n, void* st) {
    return pcx_fft(p, d, dp, s, sp, n, st);
}
inline int fft(pcx_plan_t p, double* d, int64_t dp, const double* s, int64_t sp, int64_t n, void* st) {
    return pcx_fft_f64(p, d, dp, s, sp, n, st);
}
inline int ifft(pcx_plan_t p, int nz, float* d, int64_t dp, const float* s, int64_t sp, void* st) {
    return pcx_ifft(p, nz, d, dp, s, sp, st);
}
inline int ifft(pcx_plan_t p, int nz, double* d, int64_t dp, const double* s, int64_t sp, void* st) {
    return pcx_ifft_f64(p, nz, d, dp, s, sp, st);
}
inline int batch(pcx_plan_t p, int inv, int nz, float* d, int64_t dp, int64_t dd, const float* s, int64_t sp,
                 int64_t sd, int64_t n, int64_t b, void* st) {
    return pcx_fft_batch(p, inv, nz, d, dp, dd, s, sp, sd, n, b, st);
}
inline int batch(pcx_plan_t p, int inv, int nz, double* d, int64_t dp, int64_t dd, const double* s, int64_t sp,
                 int64_t sd, int64_t n, int64_t b, void* st) {
    return pcx_fft_batch_f64(p, inv, nz, d, dp, dd, s, sp, sd, n, b, st);
}
} // namespace detail

// Allocator of device memory (the reference's Allocator parameter decides where buffers live,
// allocators.hpp:11-48).
template<typename T>
struct device_allocator {
    using value_type = T;
    device_allocator() = default;
    template<typename U>
    device_allocator(const device_allocator<U>&) noexcept {}
    T* allocate(uZ n) {
        void* p = nullptr;
        throw_on(pcx_malloc(&p, n * sizeof(T)));
        return static_cast<T*>(p);
    }
    void deallocate(T* p, uZ) noexcept { (void)pcx_free(p); }
    template<typename U>
    bool operator==(const device_allocator<U>&) const noexcept {
        return true;
    }
};

// pcxo::vector<T, PackSize, Allocator> (vector.hpp:26-382) resident in device memory.
// Lazy vector expressions (vector_arithm.hpp in this directory) are marked by this trait.
template<typename E>
struct is_vexpr : std::false_type {};
template<typename E, typename V>
void store(const E& expr, V& dest);

template<typename T, uZ PackSize = default_pack_size<T>, typename Allocator = device_allocator<T>>
    requires gpu_real<T> && pack_size<PackSize>
class vector {
public:
    using real_type      = T;
    using allocator_type = Allocator;
    static constexpr uZ pack_size = PackSize;

    vector() = default;
    explicit vector(uZ size, allocator_type alloc = allocator_type{})
    : m_alloc(alloc)
    , m_size(size) {
        if (size > 0) {
            m_ptr = m_alloc.allocate(real_size(size));
            throw_on(pcx_memset(m_ptr, 0, real_size(size) * sizeof(T))); // zero-filled, vector.hpp:58-63
        }
    }
    vector(const std::vector<std::complex<T>>& host, allocator_type alloc = allocator_type{})
    : vector(host.size(), alloc) {
        assign(host.data(), host.size());
    }
    vector(const vector& other)
    : vector(other.m_size, other.m_alloc) {
        if (m_size)
            throw_on(pcx_memcpy_dtod(m_ptr, other.m_ptr, real_size(m_size) * sizeof(T)));
    }
    vector(vector&& other) noexcept
    : m_alloc(other.m_alloc)
    , m_size(std::exchange(other.m_size, 0))
    , m_ptr(std::exchange(other.m_ptr, nullptr)) {}
    vector& operator=(vector other) noexcept {
        std::swap(m_size, other.m_size);
        std::swap(m_ptr, other.m_ptr);
        return *this;
    }
    // vector = expression (vector_arithm.hpp:405-447 store): evaluated on the GPU in one pass
    template<typename E>
        requires is_vexpr<E>::value
    vector& operator=(const E& expr) {
        store(expr, *this);
        return *this;
    }
    ~vector() {
        if (m_ptr)
            m_alloc.deallocate(m_ptr, real_size(m_size));
    }

    [[nodiscard]] uZ size() const { return m_size; }
    [[nodiscard]] T* data() { return m_ptr; }
    [[nodiscard]] const T* data() const { return m_ptr; }

    // vector.hpp:373-381
    static constexpr uZ num_packs(uZ n) { return (n % PackSize > 0) ? n / PackSize + 1 : n / PackSize; }
    static constexpr uZ real_size(uZ n) { return num_packs(n) * PackSize * 2; }
    static constexpr uZ packed_idx(uZ i) { return i + i / PackSize * PackSize; }

    // host <-> device in the packed layout
    void assign(const std::complex<T>* host, uZ n) {
        if (n > m_size)
            throw std::out_of_range("assign: " + std::to_string(n) + " values into a vector of " + std::to_string(m_size));
        std::vector<T> buf(real_size(m_size), T{});
        for (uZ i = 0; i < n; ++i) {
            buf[packed_idx(i)]            = host[i].real();
            buf[packed_idx(i) + PackSize] = host[i].imag();
        }
        throw_on(pcx_memcpy_htod(m_ptr, buf.data(), buf.size() * sizeof(T)));
    }
    [[nodiscard]] std::vector<std::complex<T>> to_host(uZ pack_override = 0) const {
        uZ             p = pack_override ? pack_override : PackSize;
        std::vector<T> buf(real_size(m_size));
        if (m_size)
            throw_on(pcx_memcpy_dtoh(buf.data(), m_ptr, buf.size() * sizeof(T)));
        std::vector<std::complex<T>> out(m_size);
        for (uZ i = 0; i < m_size; ++i) {
            uZ o   = i + i / p * p;
            out[i] = {buf[o], buf[o + p]};
        }
        return out;
    }

private:
    [[no_unique_address]] allocator_type m_alloc{};
    uZ                                   m_size = 0;
    T*                                   m_ptr  = nullptr;
};

// The layout contract of cx_vector_traits (types.hpp:113-148): element type, pack size, size
// and data.
// A batch storage (pcx_gpu/mdstorage.hpp) is not a vector: its size() is the outer extent.
template<typename S>
concept md_batch = requires { typename S::md_storage_tag; };

template<typename V>
concept device_complex_vector = !md_batch<V> && requires(V v, const V cv) {
    typename V::real_type;
    { cv.size() } -> std::convertible_to<uZ>;
    { v.data() } -> std::convertible_to<typename V::real_type*>;
    V::pack_size;
};

// fft_unit<T, Order, Allocator> (fft.hpp:565-915) on the GPU, T = float or double.  The twiddle
// tables live in the plan; copies share the (immutable) plan.  Calls are asynchronous on
// stream(); the raw forms take device pointers.
template<typename T, fft_order Order = fft_order::normal, typename Allocator = device_allocator<T>>
    requires gpu_real<T>
class fft_unit {
    static_assert(std::same_as<T, float> || Order != fft_order::unordered,
                  "the unordered order is defined for float only (the reference's unsorted kernels are "
                  "float-only)");

public:
    using real_type      = T;
    using allocator_type = Allocator;

    explicit fft_unit(uZ fft_size, uZ sub_size = 2048, allocator_type = allocator_type{})
    : m_plan(make_plan(fft_size, sub_size)) {}

    fft_unit(const fft_unit& other)     = default;
    fft_unit(fft_unit&& other) noexcept = default;
    ~fft_unit()                         = default;
    fft_unit& operator=(const fft_unit& other)     = delete;
    fft_unit& operator=(fft_unit&& other) noexcept = delete;

    [[nodiscard]] uZ size() const { return static_cast<uZ>(pcx_plan_size(m_plan.get())); }
    void             set_stream(void* stream) { m_stream = stream; }
    [[nodiscard]] void* stream() const { return m_stream; }

    // fft.hpp:624-687
    template<uZ PackSize = 1>
        requires pack_size<PackSize>
    void fft_raw(T* data) {
        throw_on(detail::fft(m_plan.get(), data, PackSize, data, PackSize, size(), m_stream));
    }
    template<uZ PackSizeDest = 1, uZ PackSizeSrc = 1>
        requires pack_size<PackSizeDest> && pack_size<PackSizeSrc>
    void fft_raw(T* dest, const T* source) {
        throw_on(detail::fft(m_plan.get(), dest, PackSizeDest, source, PackSizeSrc, size(), m_stream));
    }
    template<uZ PackSizeDest = 1, uZ PackSizeSrc = 1>
        requires pack_size<PackSizeDest> && pack_size<PackSizeSrc>
    void fft_raw(T* dest, const T* source, uZ source_size) {
        throw_on(detail::fft(m_plan.get(), dest, PackSizeDest, source, PackSizeSrc, source_size, m_stream));
    }
    // fft.hpp:689-721
    template<bool Normalized = true, uZ PackSize = 1>
        requires pack_size<PackSize>
    void ifft_raw(T* data) {
        throw_on(detail::ifft(m_plan.get(), Normalized, data, PackSize, data, PackSize, m_stream));
    }
    template<bool Normalized = true, uZ PackSizeDest = 1, uZ PackSizeSrc = 1>
        requires pack_size<PackSizeDest> && pack_size<PackSizeSrc>
    void ifft_raw(T* dest, const T* source) {
        throw_on(detail::ifft(m_plan.get(), Normalized, dest, PackSizeDest, source, PackSizeSrc, m_stream));
    }

    // fft.hpp:728-750
    template<device_complex_vector Vect_>
        requires std::same_as<typename Vect_::real_type, T>
    void operator()(Vect_& vector) {
        if (vector.size() != size())
            throw std::invalid_argument("input size (which is " + std::to_string(vector.size()) +
                                        ") is not equal to fft size (which is " + std::to_string(size()) + ")");
        fft_raw<Vect_::pack_size>(vector.data());
    }
    // fft.hpp:759-820 (source may be shorter: zero-extended)
    template<device_complex_vector DestVect_, device_complex_vector SrcVect_>
        requires std::same_as<typename DestVect_::real_type, T> && std::same_as<typename SrcVect_::real_type, T>
    void operator()(DestVect_& dest, const SrcVect_& source) {
        if (dest.size() != size())
            throw std::invalid_argument("destination size (which is " + std::to_string(dest.size()) +
                                        ") is not equal to fft size (which is " + std::to_string(size()) + ")");
        if (source.size() > size())
            throw std::invalid_argument("source size (which is " + std::to_string(source.size()) +
                                        ") is bigger than fft size (which is " + std::to_string(size()) + ")");
        fft_raw<DestVect_::pack_size, SrcVect_::pack_size>(dest.data(), source.data(), source.size());
    }
    // fft.hpp:828-853
    template<bool Normalized = true, device_complex_vector Vect_>
        requires std::same_as<typename Vect_::real_type, T>
    void ifft(Vect_& vector) {
        if (vector.size() != size())
            throw std::invalid_argument("input size (which is " + std::to_string(vector.size()) +
                                        " is not equal to fft size (which is " + std::to_string(size()) + ")");
        ifft_raw<Normalized, Vect_::pack_size>(vector.data());
    }
    // fft.hpp:863-915
    template<bool Normalized = true, device_complex_vector DestVect_, device_complex_vector SrcVect_>
        requires std::same_as<typename DestVect_::real_type, T> && std::same_as<typename SrcVect_::real_type, T>
    void ifft(DestVect_& dest, const SrcVect_& source) {
        if (dest.size() != size())
            throw std::invalid_argument("destination size (which is " + std::to_string(dest.size()) +
                                        ") is not equal to fft size (which is " + std::to_string(size()) + ")");
        if (source.size() != size())
            throw std::invalid_argument("source size (which is " + std::to_string(source.size()) +
                                        ") is not equal to fft size (which is " + std::to_string(size()) + ")");
        ifft_raw<Normalized, DestVect_::pack_size, SrcVect_::pack_size>(dest.data(), source.data());
    }

    // Every slice of a batch storage (pcx_gpu/mdstorage.hpp, the layout of a two-axis
    // md::dynamic_storage) in one launch: the slices are the rows of the batch, row_stride()
    // reals apart.  The reference transforms md slices one call at a time.
    template<md_batch S>
        requires std::same_as<typename S::real_type, T>
    void operator()(S& storage) {
        check_inner(storage.inner_extent(), "input");
        batch<S::pack_size, S::pack_size>(storage.data(), storage.row_stride(), storage.data(), storage.row_stride(),
                                          storage.size());
    }
    template<md_batch D, md_batch S>
        requires std::same_as<typename D::real_type, T> && std::same_as<typename S::real_type, T>
    void operator()(D& dest, const S& source) {
        check_inner(dest.inner_extent(), "destination");
        check_batch(dest.size(), source.size());
        if (source.inner_extent() > size())
            throw std::invalid_argument("source size (which is " + std::to_string(source.inner_extent()) +
                                        ") is bigger than fft size (which is " + std::to_string(size()) + ")");
        throw_on(detail::batch(m_plan.get(), 0, 0, dest.data(), D::pack_size, dest.row_stride(), source.data(),
                               S::pack_size, source.row_stride(), source.inner_extent(), dest.size(), m_stream));
    }
    template<bool Normalized = true, md_batch S>
        requires std::same_as<typename S::real_type, T>
    void ifft(S& storage) {
        check_inner(storage.inner_extent(), "input");
        batch<S::pack_size, S::pack_size>(storage.data(), storage.row_stride(), storage.data(), storage.row_stride(),
                                          storage.size(), true, Normalized);
    }
    template<bool Normalized = true, md_batch D, md_batch S>
        requires std::same_as<typename D::real_type, T> && std::same_as<typename S::real_type, T>
    void ifft(D& dest, const S& source) {
        check_inner(dest.inner_extent(), "destination");
        check_inner(source.inner_extent(), "source");
        check_batch(dest.size(), source.size());
        batch<D::pack_size, S::pack_size>(dest.data(), dest.row_stride(), source.data(), source.row_stride(),
                                          dest.size(), true, Normalized);
    }

    // Batched form: `batch` vectors at a distance of dist elements (device pointers).
    template<uZ PackSizeDest, uZ PackSizeSrc>
    void batch(T* dest, uZ dest_dist, const T* source, uZ source_dist, uZ count, bool inverse = false,
               bool normalized = true) {
        throw_on(detail::batch(m_plan.get(), inverse, normalized, dest, PackSizeDest, dest_dist, source, PackSizeSrc,
                               source_dist, size(), count, m_stream));
    }

private:
    void check_inner(uZ n, const char* what) const {
        if (n != size())
            throw std::invalid_argument(std::string(what) + " size (which is " + std::to_string(n) +
                                        ") is not equal to fft size (which is " + std::to_string(size()) + ")");
    }
    static void check_batch(uZ a, uZ b) {
        if (a != b)
            throw std::invalid_argument("destination slice count (which is " + std::to_string(a) +
                                        ") is not equal to source slice count (which is " + std::to_string(b) + ")");
    }
    struct plan_deleter {
        void operator()(pcx_plan_s* p) const noexcept { (void)pcx_plan_destroy(p); }
    };
    static std::shared_ptr<pcx_plan_s> make_plan(uZ n, uZ sub) {
        pcx_plan_t p = nullptr;
        throw_on(detail::plan_create(static_cast<T*>(nullptr), static_cast<int64_t>(n), static_cast<int64_t>(sub),
                                     static_cast<int>(Order), &p));
        return {p, plan_deleter{}};
    }
    std::shared_ptr<pcx_plan_s> m_plan;
    void*                       m_stream = nullptr;
};

// fxcorr_unit<float, Allocator, FFT_, TmpFactory_> (fxcorr.hpp:44-147) on the GPU.  FFT_ and
// TmpFactory_ keep the reference's policy slots: the unit takes its FFT size from FFT_::size(), and
// fft_unit_par<float, Order> (fft.hpp:2168-2572): FFT across a range of vectors, for every
// element index k: dest[j][k] = sum_i source[i][k] W^(i j).  The reference's output order is wrong
// (its test is disabled, test_fft.cpp:789); this one is natural.  The ranges hold device vectors
// (pcx_gpu::vector or anything with data(), size() and pack_size); size() of a range is the
// number of vectors and must equal fft size, as the reference requires.
template<typename T, fft_order Order = fft_order::normal, typename Allocator = device_allocator<T>>
    requires std::same_as<T, float>
class fft_unit_par {
public:
    using real_type      = T;
    using allocator_type = Allocator;

    explicit fft_unit_par(uZ size, allocator_type = allocator_type{})
    : m_size(size) {
        pcx_plan_t p = nullptr;
        throw_on(pcx_plan_create(static_cast<int64_t>(size), 2048, static_cast<int>(Order), &p));
        m_plan = std::shared_ptr<pcx_plan_s>(p, [](pcx_plan_s* q) { (void)pcx_plan_destroy(q); });
    }
    [[nodiscard]] uZ size() const { return m_size; }
    void             set_stream(void* stream) { m_stream = stream; }

    template<typename DestR_, typename SrcR_>
    void operator()(DestR_& dest, const SrcR_& source) {
        if (dest.size() != m_size)
            throw std::invalid_argument("destination size (which is " + std::to_string(dest.size()) +
                                        " is not equal to fft size (which is " + std::to_string(m_size) + ")");
        if (source.size() != m_size)
            throw std::invalid_argument("source size (which is " + std::to_string(source.size()) +
                                        " is not equal to fft size (which is " + std::to_string(m_size) + ")");
        using DV = std::remove_cvref_t<decltype(*std::begin(dest))>;
        using SV = std::remove_cvref_t<decltype(*std::begin(source))>;
        const uZ len = std::begin(source)->size();
        std::vector<const float*> sp;
        std::vector<float*>       dp;
        for (const auto& v : source) {
            if (v.size() != len)
                throw std::invalid_argument("vectors of the source range differ in size");
            sp.push_back(v.data());
        }
        for (auto& v : dest) {
            if (v.size() != len)
                throw std::invalid_argument("vectors of the destination range differ in size");
            dp.push_back(v.data());
        }
        const uZ bytes = m_size * sizeof(void*);
        if (!m_tables) {
            void* t = nullptr;
            throw_on(pcx_malloc(&t, 2 * bytes));
            m_tables = std::shared_ptr<void>(t, [](void* q) { (void)pcx_free(q); });
        }
        auto* tab = static_cast<char*>(m_tables.get());
        throw_on(pcx_synchronize(m_stream)); // the tables of a previous call may still be in use
        throw_on(pcx_memcpy_htod(tab, sp.data(), bytes));
        throw_on(pcx_memcpy_htod(tab + bytes, dp.data(), bytes));
        throw_on(pcx_fft_across(m_plan.get(), 0, 0, reinterpret_cast<float* const*>(tab + bytes), (int64_t)DV::pack_size,
                                reinterpret_cast<const float* const*>(tab), (int64_t)SV::pack_size,
                                static_cast<int64_t>(len), m_stream));
    }

private:
    uZ                          m_size;
    std::shared_ptr<pcx_plan_s> m_plan;
    std::shared_ptr<void>       m_tables;
    void*                       m_stream = nullptr;
};

// needs no caller scratch (the fused kernel keeps each block on chip), so TmpFactory_ is unused.
struct no_tmp_factory {};

template<typename T,
         typename Allocator   = device_allocator<T>,
         typename FFT_        = fft_unit<T, fft_order::unordered, Allocator>,
         typename TmpFactory_ = no_tmp_factory>
    requires std::same_as<T, float>
class fxcorr_unit {
public:
    using real_type      = T;
    using allocator_type = Allocator;

    fxcorr_unit() = delete;
    // fxcorr.hpp:62-69
    template<device_complex_vector G>
    fxcorr_unit(const G& g, uZ fft_size, allocator_type alloc = allocator_type{})
    : fxcorr_unit(g, std::make_shared<FFT_>(fft_size, 2048, alloc), alloc) {}
    // fxcorr.hpp:71-80
    template<device_complex_vector G>
    fxcorr_unit(const G& g, std::shared_ptr<FFT_> fft_unit, allocator_type = allocator_type{})
    : m_fft(std::move(fft_unit))
    , m_overlap(g.size() - 1) {
        pcx_fxcorr_t h = nullptr;
        throw_on(pcx_fxcorr_create(g.data(), G::pack_size, static_cast<int64_t>(g.size()),
                                   static_cast<int64_t>(m_fft->size()), m_stream, &h));
        m_xc = std::shared_ptr<pcx_fxcorr_s>(h, [](pcx_fxcorr_s* x) { (void)pcx_fxcorr_destroy(x); });
    }
    template<device_complex_vector G>
    fxcorr_unit(const G& g, std::shared_ptr<FFT_> fft_unit, TmpFactory_, allocator_type alloc = allocator_type{})
    : fxcorr_unit(g, std::move(fft_unit), alloc) {}

    fxcorr_unit(const fxcorr_unit&)            = default;
    fxcorr_unit(fxcorr_unit&&) noexcept        = default;
    fxcorr_unit& operator=(const fxcorr_unit&) = delete;
    fxcorr_unit& operator=(fxcorr_unit&&)      = delete;
    ~fxcorr_unit()                             = default;

    void set_stream(void* stream) { m_stream = stream; }

    // fxcorr.hpp:102-136: in place, out[i] = sum_m f[i+m] * conj(g[m]) for i < f.size()
    template<device_complex_vector Vect_>
    void operator()(Vect_& vector) {
        throw_on(pcx_fxcorr_exec(m_xc.get(), vector.data(), Vect_::pack_size, vector.data(), Vect_::pack_size,
                                 static_cast<int64_t>(vector.size()), m_stream));
    }
    // fxcorr.hpp:138-140 (an empty stub in the reference): out of place
    template<device_complex_vector DestVect_, device_complex_vector SrcVect_>
    void operator()(DestVect_& dest, const SrcVect_& source) {
        if (dest.size() < source.size())
            throw std::invalid_argument("destination size (which is " + std::to_string(dest.size()) +
                                        ") is smaller than source size (which is " + std::to_string(source.size()) + ")");
        throw_on(pcx_fxcorr_exec(m_xc.get(), dest.data(), DestVect_::pack_size, const_cast<T*>(source.data()),
                                 SrcVect_::pack_size, static_cast<int64_t>(source.size()), m_stream));
    }
    // Every slice of a batch storage (pcx_gpu/mdstorage.hpp) is a channel: one launch, one shared
    // kernel spectrum (the reference correlates one vector per call).
    template<md_batch S>
        requires std::same_as<typename S::real_type, T>
    void operator()(S& storage) {
        throw_on(pcx_fxcorr_exec_batch(m_xc.get(), storage.data(), S::pack_size, storage.row_stride(), storage.data(),
                                       S::pack_size, storage.row_stride(),
                                       static_cast<int64_t>(storage.inner_extent()),
                                       static_cast<int64_t>(storage.size()), m_stream));
    }
    template<md_batch D, md_batch S>
        requires std::same_as<typename D::real_type, T> && std::same_as<typename S::real_type, T>
    void operator()(D& dest, const S& source) {
        if (dest.size() != source.size() || dest.inner_extent() < source.inner_extent())
            throw std::invalid_argument("destination storage (which is " + std::to_string(dest.size()) + " x " +
                                        std::to_string(dest.inner_extent()) + ") does not hold source storage (which is " +
                                        std::to_string(source.size()) + " x " + std::to_string(source.inner_extent()) +
                                        ")");
        throw_on(pcx_fxcorr_exec_batch(m_xc.get(), dest.data(), D::pack_size, dest.row_stride(), source.data(),
                                       S::pack_size, source.row_stride(),
                                       static_cast<int64_t>(source.inner_extent()), static_cast<int64_t>(source.size()),
                                       m_stream));
    }
    [[nodiscard]] uZ step(uZ pack = 8) const { return static_cast<uZ>(pcx_fxcorr_step(m_xc.get(), pack)); }

private:
    std::shared_ptr<FFT_>         m_fft;
    std::shared_ptr<pcx_fxcorr_s> m_xc;
    uZ                            m_overlap;
    void*                         m_stream = nullptr;
};

} // namespace pcx_gpu

#include "vector_arithm.hpp"
