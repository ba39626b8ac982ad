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

#include <array>
#include <complex>
#include <concepts>
#include <cstddef>
#include <functional>
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

// vector_util.hpp:15: the SubSize of a unit whose sub-size is chosen at run time
inline constexpr uZ dynamic_size = static_cast<uZ>(-1);

// Node-size strategies of the reference's CPU recursion (fft.hpp:546-563).  The GPU schedule is
// chosen by the plan, so these only keep the template interface: any of them is accepted.
struct strategy42 {
    static constexpr uZ                 node_size = 4;
    static constexpr std::array<uZ, 2> align_node_size{2, 2};
    static constexpr std::array<uZ, 2> align_node_count{1, 1};
};
struct strategy48 {
    static constexpr uZ                 node_size = 4;
    static constexpr std::array<uZ, 2> align_node_size{8, 2};
    static constexpr std::array<uZ, 2> align_node_count{1, 1};
};
struct strategy8 {
    static constexpr uZ                 node_size = 8;
    static constexpr std::array<uZ, 2> align_node_size{4, 2};
    static constexpr std::array<uZ, 3> align_node_count{1, 1, 1};
};

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

// fft_unit<T, Order, Allocator, SubSize, NodeSizeRec, Strategy, StrategyRec> (fft.hpp:565-915)
// on the GPU, T = float or double.  The template parameters and constraints are the
// reference's; SubSize, NodeSizeRec and the strategies pick the reference's CPU cache blocking and
// are accepted and otherwise unused (the plan picks the GPU's schedule; pcx_plan_set_shape tunes
// it).  Allocator is accepted for the interface: device buffers come from the plan and from
// device_allocator vectors.  The twiddle tables and the four-step intermediate live in the plan,
// which copies share; a call on another stream than the plan's previous four-step call waits for
// that call, so copies may run on different streams.  Calls are asynchronous on stream(); the
// raw forms take device pointers.
template<typename T,
         fft_order Order      = fft_order::normal,
         typename Allocator   = device_allocator<T>,
         uZ SubSize           = dynamic_size,
         uZ NodeSizeRec       = 4,
         typename Strategy    = strategy42,
         typename StrategyRec = strategy42>
    requires gpu_real<T> && ((pack_size<SubSize> && SubSize >= default_pack_size<T>) || (SubSize == dynamic_size))
class fft_unit {
    static_assert(std::same_as<T, float> || Order != fft_order::unordered,
                  "the unordered order is defined for float only (the reference's unsorted kernels are "
                  "float-only)");

public:
    using real_type      = T;
    using allocator_type = Allocator;

    // fft.hpp:594-598
    explicit fft_unit(uZ fft_size, allocator_type = allocator_type{})
        requires(SubSize != dynamic_size)
    : m_plan(make_plan(fft_size, 2048)) {}
    // fft.hpp:600-605
    explicit fft_unit(uZ fft_size, uZ sub_size = 2048, allocator_type = allocator_type{})
        requires(SubSize == dynamic_size)
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

// fft_unit_par<T, Order, Allocator, NodeSizeStrategy> (fft.hpp:2168-2572), T = float (up to 8192
// vectors) or double (up to 2048): FFT across a range of vectors, for every
// element index k: dest[j][k] = sum_i source[i][k] W^(i j).  The reference's output order is wrong
// (its test is disabled, test_fft.cpp:789); this one is natural.  The ranges hold device vectors
// (pcx_gpu::vector or anything with data(), size() and pack_size); size() of a range is the
// number of vectors and must equal fft size, as the reference requires.
template<typename T,
         fft_order Order     = fft_order::normal,
         typename Allocator  = device_allocator<T>,
         uZ NodeSizeStrategy = 4>
    requires gpu_real<T>
class fft_unit_par {
public:
    using real_type      = T;
    using allocator_type = Allocator;
    using size_type      = uZ;

    explicit fft_unit_par(uZ size, allocator_type = allocator_type{})
    : m_size(size) {
        pcx_plan_t p = nullptr;
        throw_on(detail::plan_create(static_cast<T*>(nullptr), static_cast<int64_t>(size), 2048,
                                     static_cast<int>(Order), &p));
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
        // ranges of vectors or of pointers to vectors (fft.hpp:2214-2215, 2221-2228)
        const auto at = [](auto& v) -> decltype(auto) {
            if constexpr (std::is_pointer_v<std::remove_cvref_t<decltype(v)>>)
                return *v;
            else
                return (v);
        };
        using DV     = std::remove_cvref_t<decltype(at(*std::begin(dest)))>;
        using SV     = std::remove_cvref_t<decltype(at(*std::begin(source)))>;
        const uZ len = at(*std::begin(source)).size();
        std::vector<const T*> sp;
        std::vector<T*>       dp;
        for (const auto& e : source) {
            const auto& v = at(e);
            if (v.size() != len)
                throw std::invalid_argument("vectors of the source range differ in size");
            sp.push_back(v.data());
        }
        for (auto& e : dest) {
            auto& v = at(e);
            if (v.size() != len)
                throw std::invalid_argument("vectors of the destination range differ in size");
            dp.push_back(v.data());
        }
        if constexpr (DV::pack_size != SV::pack_size) {
            // k_across reads element k of every source while other workgroups write element k of
            // the destinations at other packed offsets: no destination may overlap a source
            // (fft.hpp:783 refuses an in-place repack)
            // vector.hpp:373-378: ceil(n / P) * P * 2 floats
            const auto footprint = [&](uZ p) { return (len + p - 1) / p * p * 2; };
            const auto send      = [&](const T* p) { return p + footprint(SV::pack_size); };
            const auto dend      = [&](const T* p) { return p + footprint(DV::pack_size); };
            for (const T* d : dp)
                for (const T* q : sp)
                    if (std::less<const T*>{}(d, send(q)) && std::less<const T*>{}(q, dend(d)))
                        throw std::invalid_argument("cannot perform in-place repack from source pack size (which is " +
                                                    std::to_string(SV::pack_size) +
                                                    ") to destination pack size (which is " +
                                                    std::to_string(DV::pack_size) + ")");
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
        if constexpr (std::same_as<T, double>)
            throw_on(pcx_fft_across_f64(m_plan.get(), 0, 0, reinterpret_cast<double* const*>(tab + bytes),
                                        (int64_t)DV::pack_size, reinterpret_cast<const double* const*>(tab),
                                        (int64_t)SV::pack_size, static_cast<int64_t>(len), m_stream));
        else
            throw_on(pcx_fft_across(m_plan.get(), 0, 0, reinterpret_cast<float* const*>(tab + bytes),
                                    (int64_t)DV::pack_size, reinterpret_cast<const float* const*>(tab),
                                    (int64_t)SV::pack_size, static_cast<int64_t>(len), m_stream));
    }

private:
    uZ                          m_size;
    std::shared_ptr<pcx_plan_s> m_plan;
    std::shared_ptr<void>       m_tables;
    void*                       m_stream = nullptr;
};

// The reference's default TmpFactory_ (fxcorr.hpp:10-32): owns one fft-size vector and hands out
// a pointer to it.  Here the vector is in device memory.
namespace detail_ {
template<typename T, typename Allocator = device_allocator<T>>
class pseudo_vector_factory {
public:
    explicit pseudo_vector_factory(uZ size, const Allocator& allocator = Allocator{})
    : m_vector(size, allocator) {}
    pseudo_vector_factory(const pseudo_vector_factory&)            = delete;
    pseudo_vector_factory(pseudo_vector_factory&&) noexcept        = delete;
    pseudo_vector_factory& operator=(const pseudo_vector_factory&) = delete;
    pseudo_vector_factory& operator=(pseudo_vector_factory&&)      = delete;
    ~pseudo_vector_factory()                                       = default;
    auto* operator()() { return &m_vector; }

private:
    vector<T, default_pack_size<T>, Allocator> m_vector;
};
} // namespace detail_

// The default TmpFactory_ of this unit: none.  The unit's own buffers then hold the saved overlap
// of every block, so one launch covers the whole signal.
struct no_tmp_factory {};

template<typename F>
struct is_gpu_fft_unit : std::false_type {};
template<typename T, fft_order O, typename A, uZ S, uZ N, typename St, typename StR>
struct is_gpu_fft_unit<fft_unit<T, O, A, S, N, St, StR>> : std::true_type {};

// fxcorr_unit<T, Allocator, FFT_, TmpFactory_> (fxcorr.hpp:44-147) on the GPU, T = float.
// FFT_ and TmpFactory_ are the reference's policy slots:
//  * FFT_ = this library's fft_unit (the default, any order): the unit runs the fused
//    overlap-save kernels of the C ABI, which compute their own transforms (one kernel per launch
//    for fft_size <= 16384, the batched four-step block path above).
//  * any other FFT_ is driven as the reference drives it (fxcorr.hpp:102-136): one block at a time
//    through FFT_::fft_raw<PT, PS>(tmp, src[, n]), a device-side multiply by conj(kernel) and
//    FFT_::ifft_raw<false, PT>(tmp), with the block offset applied in complex elements (the
//    reference applies it to a float pointer, which is wrong from the second block on).  Such an
//    FFT_ must operate on device memory; if it has set_stream(void*), it is given the unit's
//    stream.
//  * TmpFactory_ = no_tmp_factory (the default) lets the unit own its scratch.  Any other factory
//    (operator() returning a pointer to a device vector of at least fft size, fxcorr.hpp:39-40)
//    supplies the scratch: the block vector of a driven FFT_, or the workspace of the fused path
//    (pcx_fxcorr_exec_batch_ws).
template<typename T,
         typename Allocator   = device_allocator<T>,
         typename FFT_        = fft_unit<T, fft_order::unordered, Allocator>,
         typename TmpFactory_ = no_tmp_factory>
    requires std::same_as<T, float>
class fxcorr_unit {
    static constexpr bool fused       = is_gpu_fft_unit<FFT_>::value;
    static constexpr bool own_factory = std::same_as<TmpFactory_, no_tmp_factory>;
    using factory_t = std::conditional_t<own_factory && !fused, detail_::pseudo_vector_factory<T, device_allocator<T>>,
                                         TmpFactory_>;

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
    , m_factory(make_factory(m_fft->size()))
    , m_overlap(g.size() - 1) {
        init(g);
    }
    // fxcorr.hpp:82-91
    template<device_complex_vector G>
    fxcorr_unit(const G& g, std::shared_ptr<FFT_> fft_unit, TmpFactory_ tmp_factory, allocator_type = allocator_type{})
    : m_fft(std::move(fft_unit))
    , m_factory(std::make_shared<factory_t>(std::move(tmp_factory)))
    , m_overlap(g.size() - 1) {
        init(g);
    }

    fxcorr_unit(const fxcorr_unit&)            = default;
    fxcorr_unit(fxcorr_unit&&) noexcept        = default;
    fxcorr_unit& operator=(const fxcorr_unit&) = delete;
    fxcorr_unit& operator=(fxcorr_unit&&)      = delete;
    ~fxcorr_unit()                             = default;

    void set_stream(void* stream) {
        m_stream = stream;
        if constexpr (requires(FFT_& f) { f.set_stream(stream); })
            m_fft->set_stream(stream);
    }

    // fxcorr.hpp:102-136: in place, out[i] = sum_m f[i+m] * conj(g[m]) for i < f.size()
    template<device_complex_vector Vect_>
    void operator()(Vect_& vector) {
        if constexpr (fused)
            exec(vector.data(), Vect_::pack_size, 0, vector.data(), Vect_::pack_size, 0, vector.size(), 1);
        else
            drive(vector.data(), vector.size(), std::integral_constant<uZ, Vect_::pack_size>{});
    }
    // fxcorr.hpp:138-140 (an empty stub in the reference): out of place
    template<device_complex_vector DestVect_, device_complex_vector SrcVect_>
    void operator()(DestVect_& dest, const SrcVect_& source) {
        if (dest.size() < source.size())
            throw std::invalid_argument("destination size (which is " + std::to_string(dest.size()) +
                                        ") is smaller than source size (which is " + std::to_string(source.size()) + ")");
        if constexpr (fused) {
            exec(dest.data(), DestVect_::pack_size, 0, source.data(), SrcVect_::pack_size, 0, source.size(), 1);
        } else {
            static_assert(DestVect_::pack_size == SrcVect_::pack_size,
                          "a driven FFT_ correlates out of place between vectors of one pack size");
            throw_on(pcx_repack(dest.data(), DestVect_::pack_size, source.data(), SrcVect_::pack_size,
                                static_cast<int64_t>(source.size()), m_stream));
            drive(dest.data(), source.size(), std::integral_constant<uZ, DestVect_::pack_size>{});
        }
    }
    // Every slice of a batch storage (pcx_gpu/mdstorage.hpp) is a channel: one launch, one shared
    // kernel spectrum (the reference correlates one vector per call).
    template<md_batch S>
        requires std::same_as<typename S::real_type, T>
    void operator()(S& storage) {
        static_assert(fused, "batch storages are correlated by the fused path (FFT_ = pcx_gpu::fft_unit)");
        exec(storage.data(), S::pack_size, storage.row_stride(), storage.data(), S::pack_size, storage.row_stride(),
             storage.inner_extent(), storage.size());
    }
    template<md_batch D, md_batch S>
        requires std::same_as<typename D::real_type, T> && std::same_as<typename S::real_type, T>
    void operator()(D& dest, const S& source) {
        static_assert(fused, "batch storages are correlated by the fused path (FFT_ = pcx_gpu::fft_unit)");
        if (dest.size() != source.size() || dest.inner_extent() < source.inner_extent())
            throw std::invalid_argument("destination storage (which is " + std::to_string(dest.size()) + " x " +
                                        std::to_string(dest.inner_extent()) + ") does not hold source storage (which is " +
                                        std::to_string(source.size()) + " x " + std::to_string(source.inner_extent()) +
                                        ")");
        exec(dest.data(), D::pack_size, dest.row_stride(), source.data(), S::pack_size, source.row_stride(),
             source.inner_extent(), source.size());
    }
    // fxcorr.hpp:110: outputs per block for a signal of pack size `pack`
    [[nodiscard]] uZ step(uZ pack = 8) const {
        return fused ? static_cast<uZ>(pcx_fxcorr_step(m_xc.get(), static_cast<int64_t>(pack)))
                     : (m_fft->size() - m_overlap) / pack * pack;
    }

private:
    // fxcorr.hpp:64, 74: the factory is built for the FFT size (none for the fused path's own
    // buffers)
    static std::shared_ptr<factory_t> make_factory(uZ n) {
        if constexpr (own_factory && fused)
            return nullptr;
        else
            return std::make_shared<factory_t>(n);
    }
    template<device_complex_vector G>
    void init(const G& g) {
        if (g.size() == 0 || g.size() > m_fft->size())
            throw std::invalid_argument("source size (which is " + std::to_string(g.size()) +
                                        ") is bigger than fft size (which is " + std::to_string(m_fft->size()) + ")");
        if constexpr (fused) {
            pcx_fxcorr_t h = nullptr;
            throw_on(pcx_fxcorr_create(g.data(), G::pack_size, static_cast<int64_t>(g.size()),
                                       static_cast<int64_t>(m_fft->size()), m_stream, &h));
            m_xc = std::shared_ptr<pcx_fxcorr_s>(h, [](pcx_fxcorr_s* x) { (void)pcx_fxcorr_destroy(x); });
        } else {
            // fxcorr.hpp:67-68: kernel = FFT_(g) / N, in FFT_'s own order
            const uZ n = m_fft->size();
            m_kernel   = std::make_shared<vector<T>>(n);
            m_fft->template fft_raw<default_pack_size<T>, G::pack_size>(m_kernel->data(), g.data(), g.size());
            const pcx_vsource src{m_kernel->data(), default_pack_size<T>};
            const pcx_vinstr  prog[2] = {{PCX_VOP_LOAD, 0, 0, 0, 0.f, 0.f},
                                         {PCX_VOP_MULR, 0, 0, 0, static_cast<float>(1.0 / static_cast<double>(n)), 0.f}};
            throw_on(pcx_vexpr_eval(&src, 1, prog, 2, m_kernel->data(), default_pack_size<T>, static_cast<int64_t>(n),
                                    m_stream));
            throw_on(pcx_synchronize(m_stream));
        }
    }
    void exec(T* dst, uZ dp, uZ dd, const T* src, uZ sp, uZ sd, uZ len, uZ channels) {
        if constexpr (own_factory) {
            throw_on(pcx_fxcorr_exec_batch(m_xc.get(), dst, (int64_t)dp, (int64_t)dd, src, (int64_t)sp, (int64_t)sd,
                                           (int64_t)len, (int64_t)channels, m_stream));
        } else {
            auto* tmp = (*m_factory)();
            using TV  = std::remove_cvref_t<decltype(*tmp)>;
            const int64_t bytes = static_cast<int64_t>((tmp->size() + TV::pack_size - 1) / TV::pack_size *
                                                       TV::pack_size * 2 * sizeof(T));
            throw_on(pcx_fxcorr_exec_batch_ws(m_xc.get(), dst, (int64_t)dp, (int64_t)dd, src, (int64_t)sp, (int64_t)sd,
                                              (int64_t)len, (int64_t)channels, tmp->data(), bytes, m_stream));
        }
    }
    // fxcorr.hpp:102-136 with a driven FFT_, on `data` (pack P) in place
    template<uZ P>
    void drive(T* data, uZ size, std::integral_constant<uZ, P>) {
        auto*        tmp = (*m_factory)();
        using TV         = std::remove_cvref_t<decltype(*tmp)>;
        constexpr uZ PT  = TV::pack_size;
        const uZ     n   = m_fft->size();
        if (tmp->size() < n)
            throw std::invalid_argument("temporary vector size (which is " + std::to_string(tmp->size()) +
                                        ") is smaller than fft size (which is " + std::to_string(n) + ")");
        if (n <= m_overlap)
            throw std::invalid_argument("fft size (which is " + std::to_string(n) +
                                        ") leaves no output per block for a kernel of " + std::to_string(m_overlap + 1) +
                                        " taps");
        const uZ step = (n - m_overlap) / P * P;
        if (step == 0)
            throw std::invalid_argument("fft size (which is " + std::to_string(n) + ") leaves no output per block");
        // tmp = tmp * conj(kernel), both in FFT_'s spectrum layout (pack PT / the kernel's pack)
        const pcx_vsource srcs[2] = {{tmp->data(), (int64_t)PT}, {m_kernel->data(), (int64_t)default_pack_size<T>}};
        const pcx_vinstr  prog[4] = {{PCX_VOP_LOAD, 0, 0, 0, 0.f, 0.f},
                                     {PCX_VOP_LOAD, 1, 1, 0, 0.f, 0.f},
                                     {PCX_VOP_CONJ, 1, 1, 0, 0.f, 0.f},
                                     {PCX_VOP_MUL, 0, 0, 1, 0.f, 0.f}};
        for (uZ offset = 0; offset < size; offset += step) {
            // element `offset` (a multiple of P) starts at float offset 2 * offset of a pack-P vector
            T* const blk  = data + 2 * offset;
            const uZ keep = std::min(step, size - offset);
            if (offset + n < size)
                m_fft->template fft_raw<PT, P>(tmp->data(), blk);
            else
                m_fft->template fft_raw<PT, P>(tmp->data(), blk, size - offset);
            throw_on(pcx_vexpr_eval(srcs, 2, prog, 4, tmp->data(), (int64_t)PT, static_cast<int64_t>(n), m_stream));
            m_fft->template ifft_raw<false, PT>(tmp->data());
            throw_on(pcx_repack(blk, (int64_t)P, tmp->data(), (int64_t)PT, static_cast<int64_t>(keep), m_stream));
        }
    }

    std::shared_ptr<FFT_>         m_fft;
    std::shared_ptr<factory_t>    m_factory;
    std::shared_ptr<pcx_fxcorr_s> m_xc;     // fused path
    std::shared_ptr<vector<T>>    m_kernel; // driven FFT_: FFT_(g) / N
    uZ                            m_overlap;
    void*                         m_stream = nullptr;
};

} // namespace pcx_gpu

#include "vector_arithm.hpp"
