// pcx_gpu/mdstorage.hpp -- a batch of packed vectors in device memory with the layout of the
// reference's two-axis md::dynamic_storage (mdstorage.hpp:1478-1499), and the fft_unit entry
// points that transform every slice of it in one batched launch.
//
// Layout (mdstorage.hpp:879-885 calc_storage_size, 370-391 get_dynamic_slice_offset): the inner
// axis is one packed vector of `inner` complex values with pack P; consecutive slices along the
// outer axis start row_stride() reals apart, with
//   row_stride = ceil(2 * inner / A) * A,   A = lcm(2 * P, Alignment),
// and the storage holds row_stride * outer reals, zero-filled.  Extents are given in layout order
// (inner first), as in a left_basis; a right_basis storage of the same extents has the same
// layout.  Only the two-axis case is covered: it is the batch descriptor of the FFT path.
#pragma once

#include "fft.hpp"

#include <numeric>

namespace pcx_gpu::md {

// A slice along the outer axis: a non-owning packed vector of `size()` values (it satisfies
// device_complex_vector, so fft_unit takes it like a vector, as the reference's fft_unit takes an
// md sslice).
template<typename T, uZ PackSize>
class slice {
public:
    using real_type               = T;
    static constexpr uZ pack_size = PackSize;
    slice(T* data, uZ size)
    : m_data(data)
    , m_size(size) {}
    [[nodiscard]] uZ size() const { return m_size; }
    [[nodiscard]] T* data() const { return m_data; }

private:
    T* m_data;
    uZ m_size;
};

template<gpu_real T, uZ PackSize = default_pack_size<T>, uZ Alignment = default_pack_size<T>,
         typename Allocator = device_allocator<T>>
    requires pack_size<PackSize> && pack_size<Alignment>
class dynamic_storage {
public:
    using md_storage_tag          = void; // marks a storage, not a vector
    using real_type               = T;
    using allocator_type          = Allocator;
    static constexpr uZ pack_size = PackSize;
    static constexpr uZ alignment = std::lcm(PackSize * 2, Alignment);

    dynamic_storage() = default;
    // extents in layout order: the inner (contiguous) axis first
    dynamic_storage(uZ inner_extent, uZ outer_extent, allocator_type alloc = allocator_type{})
    : m_alloc(alloc)
    , m_inner(inner_extent)
    , m_outer(outer_extent) {
        if (storage_size() > 0) {
            m_ptr = m_alloc.allocate(storage_size());
            throw_on(pcx_memset(m_ptr, 0, storage_size() * sizeof(T)));
        }
    }
    dynamic_storage(const dynamic_storage&)            = delete; // as the reference (mdstorage.hpp:763-764)
    dynamic_storage& operator=(const dynamic_storage&) = delete;
    dynamic_storage(dynamic_storage&& o) noexcept
    : m_alloc(o.m_alloc)
    , m_inner(std::exchange(o.m_inner, 0))
    , m_outer(std::exchange(o.m_outer, 0))
    , m_ptr(std::exchange(o.m_ptr, nullptr)) {}
    dynamic_storage& operator=(dynamic_storage&& o) noexcept {
        std::swap(m_inner, o.m_inner);
        std::swap(m_outer, o.m_outer);
        std::swap(m_ptr, o.m_ptr);
        return *this;
    }
    ~dynamic_storage() {
        if (m_ptr)
            m_alloc.deallocate(m_ptr, storage_size());
    }

    // extent of the outer axis (mdstorage.hpp:1466-1468)
    [[nodiscard]] uZ size() const { return m_outer; }
    [[nodiscard]] uZ inner_extent() const { return m_inner; }
    [[nodiscard]] uZ outer_extent() const { return m_outer; }
    // reals between consecutive slices (mdstorage.hpp:879-885)
    [[nodiscard]] uZ row_stride() const { return (m_inner * 2 + alignment - 1) / alignment * alignment; }
    [[nodiscard]] uZ storage_size() const { return row_stride() * m_outer; }
    [[nodiscard]] T* data() { return m_ptr; }
    [[nodiscard]] const T* data() const { return m_ptr; }

    // slice along the outer axis (mdstorage.hpp:1433-1446)
    [[nodiscard]] slice<T, PackSize> operator[](uZ index) { return {m_ptr + index * row_stride(), m_inner}; }
    [[nodiscard]] slice<const T, PackSize> operator[](uZ index) const {
        return {m_ptr + index * row_stride(), m_inner};
    }

    // host <-> device, slice by slice
    void assign(uZ index, const std::complex<T>* host, uZ n) {
        if (index >= m_outer || n > m_inner)
            throw std::out_of_range("assign: slice " + std::to_string(index) + " of " + std::to_string(m_outer) +
                                    ", " + std::to_string(n) + " values into " + std::to_string(m_inner));
        std::vector<T> buf(row_stride(), T{});
        for (uZ i = 0; i < n; ++i) {
            buf[i + i / PackSize * PackSize]            = host[i].real();
            buf[i + i / PackSize * PackSize + PackSize] = host[i].imag();
        }
        throw_on(pcx_memcpy_htod(m_ptr + index * row_stride(), buf.data(), buf.size() * sizeof(T)));
    }
    [[nodiscard]] std::vector<std::complex<T>> to_host(uZ index) const {
        std::vector<T> buf(row_stride());
        throw_on(pcx_memcpy_dtoh(buf.data(), m_ptr + index * row_stride(), buf.size() * sizeof(T)));
        std::vector<std::complex<T>> out(m_inner);
        for (uZ i = 0; i < m_inner; ++i) {
            uZ o   = i + i / PackSize * PackSize;
            out[i] = {buf[o], buf[o + PackSize]};
        }
        return out;
    }

private:
    [[no_unique_address]] allocator_type m_alloc{};
    uZ                                   m_inner = 0;
    uZ                                   m_outer = 0;
    T*                                   m_ptr   = nullptr;
};

} // namespace pcx_gpu::md
