// C++ parity test of the drop-in classes in packed_amd/include/pcx_gpu/fft.hpp, written after the
// reference's tests/test_fft.cpp (test_fft_float 288-435, test_fftu_float 437-639) and
// tests/test_fxcorr.cpp (61-139), with the FP64 oracle (oracle/pcx_oracle.h) as the checker.
// Exit code = number of failed checks.
#include "pcx_gpu/fft.hpp"
#include "pcx_gpu/mdstorage.hpp"
#include "pcx_oracle.h"

#include <cmath>
#include <complex>
#include <cstdio>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

using cxf = std::complex<float>;
using cxd = std::complex<double>;

static int g_fail = 0;
#define CHECK(cond, ...)                                                                           \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            ++g_fail;                                                                              \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                       \
            std::printf(__VA_ARGS__);                                                              \
            std::printf("\n");                                                                     \
        }                                                                                          \
    } while (0)

static double rel_l2(const std::vector<cxf>& a, const std::vector<cxd>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(cxd(a[i]) - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / (den > 0 ? den : 1));
}

// tests/test_pcx.hpp:5-27 equal_eps with k epsilons
static bool equal_eps(cxd a, cxd b, double k) {
    if (a == b)
        return true;
    double largest = std::max(std::abs(a), std::abs(b));
    return std::abs(a - b) < largest * k * std::numeric_limits<float>::epsilon();
}

static std::vector<cxd> oracle_fft(int order, bool inverse, bool normalized, const std::vector<cxf>& in) {
    int64_t            n = (int64_t)in.size();
    std::vector<float> src(2 * n), dst(2 * n);
    for (int64_t i = 0; i < n; ++i) {
        src[orc_pidx(i, n)]     = in[i].real();
        src[orc_pidx(i, n) + n] = in[i].imag();
    }
    std::vector<double> re(n), im(n);
    orc_fft_unit(order, inverse, normalized, n, dst.data(), n, src.data(), n, n, re.data(), im.data());
    std::vector<cxd> out(n);
    for (int64_t i = 0; i < n; ++i)
        out[i] = {re[i], im[i]};
    return out;
}

static std::vector<cxf> leaky_tone(size_t n) { // test_fft.cpp:299-302
    const float      pi = 3.14159265358979323846f;
    std::vector<cxf> v(n);
    for (size_t i = 0; i < n; ++i)
        v[i] = std::exp(cxf(0.F, 2 * pi * i / n * 13.37F));
    return v;
}

template<size_t PackSize>
static void test_fft_float(size_t size) {
    using namespace pcx_gpu;
    double depth = std::log2((double)size);
    double tol   = 2 * depth * std::numeric_limits<float>::epsilon();
    auto   x     = leaky_tone(size);
    auto   want  = oracle_fft(0, false, true, x);

    vector<float, PackSize> vec(x);
    vector<float, PackSize> vec_out(size);
    fft_unit<float, fft_order::normal> unit(size, 2048);
    unit(vec_out, vec); // out of place
    auto got = vec_out.to_host();
    CHECK(rel_l2(got, want) < tol, "normal oop N=%zu P=%zu err=%g", size, PackSize, rel_l2(got, want));
    int bad = 0;
    for (size_t i = 0; i < size; ++i)
        bad += !equal_eps(cxd(got[i]), want[i], (double)size); // reference tolerance k = N (test_fft.cpp:363)
    CHECK(bad == 0, "normal N=%zu P=%zu: %d elements outside N*eps", size, PackSize, bad);

    unit(vec); // in place
    got = vec.to_host();
    CHECK(rel_l2(got, want) < tol, "normal in-place N=%zu P=%zu", size, PackSize);
    unit.ifft(vec); // round trip, normalised
    got = vec.to_host();
    std::vector<cxd> xd(x.begin(), x.end());
    CHECK(rel_l2(got, xd) < 2 * tol, "ifft round trip N=%zu P=%zu err=%g", size, PackSize, rel_l2(got, xd));

    // pack-1 (std::complex interleaved) destination from the pack-P source
    vector<float, 1> svec(size);
    unit(svec, vector<float, PackSize>(x));
    got = svec.to_host();
    CHECK(rel_l2(got, want) < tol, "pack-1 destination N=%zu", size);
}

template<size_t PackSize>
static void test_fftu_float(size_t size) {
    using namespace pcx_gpu;
    double tol = 2 * std::log2((double)size) * std::numeric_limits<float>::epsilon();
    auto   x   = leaky_tone(size);
    {
        auto                                     want = oracle_fft(1, false, true, x);
        vector<float, PackSize>                  vec(x), out(size);
        fft_unit<float, fft_order::bit_reversed> unit(size);
        unit(out, vec);
        auto got = out.to_host();
        CHECK(rel_l2(got, want) < tol, "bit_reversed N=%zu err=%g", size, rel_l2(got, want));
        // in-place fft_raw vs out-of-place operator() (test_fft.cpp:491-505 compares them with
        // equal_eps); the two forms may run different kernels, so compare within the tolerance
        unit.fft_raw<PackSize>(vec.data());
        auto             raw = vec.to_host();
        std::vector<cxd> gotd(got.begin(), got.end());
        CHECK(rel_l2(raw, gotd) < tol, "bit_reversed fft_raw vs operator() N=%zu err=%g", size, rel_l2(raw, gotd));
        unit.template ifft<true>(vec);
        std::vector<cxd> xd(x.begin(), x.end());
        CHECK(rel_l2(vec.to_host(), xd) < 2 * tol, "bit_reversed round trip N=%zu", size);
    }
    {
        auto                                  want = oracle_fft(2, false, true, x);
        vector<float, PackSize>               vec(x), out(size);
        fft_unit<float, fft_order::unordered> unit(size);
        unit(out, vec);
        auto got = out.to_host();
        CHECK(rel_l2(got, want) < tol, "unordered N=%zu err=%g", size, rel_l2(got, want));
        // unnormalised inverse: N * x
        vector<float, PackSize> back(size);
        unit.template ifft<false>(back, out);
        auto             r = back.to_host();
        std::vector<cxd> nx(size);
        for (size_t i = 0; i < size; ++i)
            nx[i] = cxd(x[i]) * (double)size;
        CHECK(rel_l2(r, nx) < 2 * tol, "unordered unnormalised inverse N=%zu", size);
    }
}

static void test_zero_extension() {
    using namespace pcx_gpu;
    size_t                                   n = 1024, m = 700;
    auto                                     x = leaky_tone(m);
    vector<float, 8>                         src(x), dst(n);
    fft_unit<float, fft_order::bit_reversed> unit(n);
    unit(dst, src); // source shorter than the transform: zero-extended
    std::vector<cxf> xz(n, cxf(0, 0));
    for (size_t i = 0; i < m; ++i)
        xz[i] = x[i];
    auto want = oracle_fft(1, false, true, xz);
    CHECK(rel_l2(dst.to_host(), want) < 2 * 10 * std::numeric_limits<float>::epsilon(), "zero extension");
}

static void test_errors() {
    using namespace pcx_gpu;
    try {
        fft_unit<float> u(1000);
        CHECK(false, "size 1000 accepted");
    } catch (const std::invalid_argument& e) {
        CHECK(std::string(e.what()) == "fft_size (which is  1000) is not an integer power of two", "msg: %s", e.what());
    }
    try {
        fft_unit<float> u(1024, 4);
        CHECK(false, "sub_size 4 accepted");
    } catch (const std::invalid_argument& e) {
        CHECK(std::string(e.what()) == "fft_sub_size (which is  4) is smaller than minimum (which is 8)", "msg: %s",
              e.what());
    }
    fft_unit<float>  u(256);
    vector<float, 8> v(128);
    try {
        u(v);
        CHECK(false, "size mismatch accepted");
    } catch (const std::invalid_argument& e) {
        CHECK(std::string(e.what()) == "input size (which is 128) is not equal to fft size (which is 256)", "msg: %s",
              e.what());
    }
    vector<float, 16>  a(256);
    vector<float, 256> b(256);
    try {
        u.fft_raw<16, 256>(a.data(), a.data());
        CHECK(false, "in-place repack accepted");
    } catch (const std::invalid_argument& e) {
        CHECK(std::string(e.what()).find("cannot perform in-place repack") == 0, "msg: %s", e.what());
    }
}

static void test_fxcorr() {
    using namespace pcx_gpu;
    // tests/test_fxcorr.cpp:100-139: Barker-13 kernel, f = i*(Barker at 0 and 109) + rand()%7 - 3
    const double       bark[13] = {1, 1, 1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1};
    size_t             size     = 128;
    std::vector<cxf>   g(size, 0), f(size, 0);
    for (int i = 0; i < 13; ++i) {
        g[i] = (float)bark[i];
        f[i] += (float)bark[i];
        f[109 + i] += (float)bark[i];
    }
    orc_glibc_rand_t rs;
    orc_glibc_srand(&rs, 1);
    for (size_t i = 0; i < size; ++i)
        f[i] = f[i] * cxf(0, 1) + (float)(orc_glibc_rand(&rs) % 7 - 3);
    std::vector<double> fr(size), fi(size), gr(size), gi(size), outr(size), outi(size);
    for (size_t i = 0; i < size; ++i)
        fr[i] = f[i].real(), fi[i] = f[i].imag(), gr[i] = g[i].real(), gi[i] = g[i].imag();
    orc_xcorr_naive(fr.data(), fi.data(), size, gr.data(), gi.data(), size, outr.data(), outi.data());
    vector<float>        gv(g), fv(f);
    fxcorr_unit<float>   xcr(gv, size * 2);
    xcr(fv);
    auto             got = fv.to_host();
    std::vector<cxd> want(size);
    for (size_t i = 0; i < size; ++i)
        want[i] = {outr[i], outi[i]};
    CHECK(rel_l2(got, want) < 1e-6, "fxcorr Barker case err=%g", rel_l2(got, want));
    CHECK(std::abs(got[109].imag() - 13.f) < 0.5f && std::abs(got[0].imag() - 13.f) < 0.5f, "Barker peaks");

    // multi-block, out of place (the reference overload is an empty stub)
    size_t              L = 20000, M = 300, N = 1024;
    std::vector<float>  sr(L), si(L), kr(M), ki(M);
    orc_gen_normal(0x5EED0003, 0.7071067811865476, sr.data(), si.data(), L);
    orc_gen_uniform(0x5EED0013, kr.data(), ki.data(), M);
    std::vector<cxf> sig(L), ker(M);
    for (size_t i = 0; i < L; ++i)
        sig[i] = {sr[i], si[i]};
    for (size_t i = 0; i < M; ++i)
        ker[i] = std::polar(1.0f, 3.14159265f * kr[i]);
    std::vector<double> a(L), b(L), c(M), d(M), o1(L), o2(L);
    for (size_t i = 0; i < L; ++i)
        a[i] = sig[i].real(), b[i] = sig[i].imag();
    for (size_t i = 0; i < M; ++i)
        c[i] = ker[i].real(), d[i] = ker[i].imag();
    orc_xcorr_naive(a.data(), b.data(), L, c.data(), d.data(), M, o1.data(), o2.data());
    vector<float>      sv(sig), out(L), kv(ker);
    fxcorr_unit<float> xc2(kv, N);
    xc2(out, sv);
    auto got2 = out.to_host();
    std::vector<cxd> want2(L);
    for (size_t i = 0; i < L; ++i)
        want2[i] = {o1[i], o2[i]};
    CHECK(rel_l2(got2, want2) < 4e-6, "fxcorr multi-block err=%g", rel_l2(got2, want2));
    CHECK(xc2.step() == (N - (M - 1)) / 8 * 8, "fxcorr step");
}

// tests/test_op.cpp:89-226 test_bin_ops: the 16 vector-vector operations, then complex- and
// real-scalar operations on either side, checked with equal_eps (10 FLT_EPSILON relative).
// tests/test_op.cpp test_bin_ops, for float and double (the reference runs both): element type
// T, host result in a wider type W, every element within 10 epsilon(T) relative.
template<typename T, size_t PackSize>
static void test_bin_ops(size_t length) {
    using namespace pcx_gpu;
    using W             = std::conditional_t<std::is_same_v<T, float>, double, long double>;
    using cT            = std::complex<T>;
    using cW            = std::complex<W>;
    const T         eps = std::numeric_limits<T>::epsilon();
    std::vector<cT> lhs(length), rhs(length);
    for (size_t i = 0; i < length; ++i) {
        lhs[i] = cT((T)i * eps, T(-1) - (T)i);
        rhs[i] = cT((T)i + 1, (T)i * eps);
    }
    vector<T, PackSize> L(lhs), R(rhs), out(length);
    const cT            cx(512 * (eps + 1), -512 * eps);
    const T             re = 1024 * (eps - 1);
    auto check = [&](const char* what, auto ref) {
        auto got = out.to_host();
        for (size_t i = 0; i < length; ++i) {
            cW want = ref(i);
            cW g    = cW(got[i]);
            W  big  = std::max(std::abs(g), std::abs(want));
            if (!(g == want || std::abs(g - want) < big * 10 * (W)eps)) {
                CHECK(false, "%s %s pack %zu length %zu index %zu", sizeof(T) == 4 ? "float" : "double", what, PackSize,
                      length, i);
                return;
            }
        }
    };
    auto l  = [&](size_t i) { return cW(lhs[i]); };
    auto r  = [&](size_t i) { return cW(rhs[i]); };
    auto cs = cW(cx);
    auto rs = (W)re;
    out = L + R;                 check("a+b", [&](size_t i) { return l(i) + r(i); });
    out = L - conj(R);           check("a-conj(b)", [&](size_t i) { return l(i) - std::conj(r(i)); });
    out = conj(L) * R;           check("conj(a)*b", [&](size_t i) { return std::conj(l(i)) * r(i); });
    out = conj(L) / conj(R);     check("conj(a)/conj(b)", [&](size_t i) { return std::conj(l(i)) / std::conj(r(i)); });
    out = L / R;                 check("a/b", [&](size_t i) { return l(i) / r(i); });
    out = L * cx;                check("a*s", [&](size_t i) { return l(i) * cs; });
    out = cx / R;                check("s/b", [&](size_t i) { return cs / r(i); });
    out = cx - conj(R);          check("s-conj(b)", [&](size_t i) { return cs - std::conj(r(i)); });
    out = L + re;                check("a+re", [&](size_t i) { return l(i) + rs; });
    out = re - R;                check("re-b", [&](size_t i) { return rs - r(i); });
    out = re / R;                check("re/b", [&](size_t i) { return rs / r(i); });
    out = L / re;                check("a/re", [&](size_t i) { return l(i) / rs; });
    // a fused pipeline across pack sizes: one kernel
    vector<T, 1> L1(lhs);
    out = (L1 + R) * conj(L - R) / (R + T(2)) - cx * L1;
    check("fused", [&](size_t i) {
        return (l(i) + r(i)) * std::conj(l(i) - r(i)) / (r(i) + W(2)) - cs * l(i);
    });
    // repack through assignment of a vector of another pack size
    store(L1, out);
    check("repack", [&](size_t i) { return l(i); });
}

// fft_unit_par: FFT across 16 vectors of 100 elements (pack 8), natural order, against the FP64
// DFT along the vector index; and the reference's size check (fft.hpp:2198-2211).
static void test_fft_par() {
    using namespace pcx_gpu;
    const size_t n = 16, len = 100;
    std::vector<std::vector<cxf>>          host(n, std::vector<cxf>(len));
    std::vector<vector<float, 8>>          src, dst;
    for (size_t i = 0; i < n; ++i) {
        for (size_t k = 0; k < len; ++k)
            host[i][k] = cxf(std::sin(0.37f * (float)(i * len + k)), std::cos(0.11f * (float)(i + 3 * k)));
        src.emplace_back(host[i]);
        dst.emplace_back(len);
    }
    fft_unit_par<float> par(n);
    par(dst, src);
    double num = 0, den = 0;
    for (size_t j = 0; j < n; ++j) {
        auto got = dst[j].to_host();
        for (size_t k = 0; k < len; ++k) {
            cxd want = 0;
            for (size_t i = 0; i < n; ++i)
                want += cxd(host[i][k]) * std::polar(1.0, -2 * M_PI * (double)((i * j) % n) / (double)n);
            num += std::norm(cxd(got[k]) - want);
            den += std::norm(want);
        }
    }
    double err = std::sqrt(num / den);
    CHECK(err < 2 * 4 * std::numeric_limits<float>::epsilon(), "fft_unit_par rel L2 %g", err);
    std::vector<vector<float, 8>> short_range;
    short_range.emplace_back(len);
    bool thrown = false;
    try {
        par(short_range, src);
    } catch (const std::invalid_argument& e) {
        thrown = std::string(e.what()).find("destination size (which is 1 is not equal to fft size") == 0;
    }
    CHECK(thrown, "fft_unit_par size check");
}

// fft_unit<double, Order> (the reference declares it with pack 4, types.hpp:70-72): forward in
// normal and bit_reversed order against the FP64 oracle, in and out of place, and the round trip.
template<size_t PackSize>
static void test_fft_double(size_t size) {
    std::vector<cxd> in(size);
    for (size_t i = 0; i < size; ++i) // a leaky tone in double precision (test_fft.cpp:299-302)
        in[i] = std::exp(cxd(0., 2 * M_PI * (double)i / (double)size * 13.37));
    std::vector<double> xr(size), xi(size), yr(size), yi(size);
    for (size_t i = 0; i < size; ++i)
        xr[i] = in[i].real(), xi[i] = in[i].imag();
    orc_fft(xr.data(), xi.data(), yr.data(), yi.data(), (int64_t)size, 0);
    int bits = 0;
    while (((size_t)1 << bits) < size)
        ++bits;
    const double tol = 4.0 * bits * std::numeric_limits<double>::epsilon();
    auto         err = [&](const std::vector<cxd>& got, bool brev) {
        double num = 0, den = 0;
        for (size_t i = 0; i < size; ++i) {
            size_t k = brev ? (size_t)orc_brev((int64_t)i, bits) : i;
            num += std::norm(got[i] - cxd(yr[k], yi[k]));
            den += yr[k] * yr[k] + yi[k] * yi[k];
        }
        return std::sqrt(num / den);
    };
    pcx_gpu::vector<double, PackSize>                      v(in);
    pcx_gpu::vector<double, PackSize>                      out(size);
    pcx_gpu::fft_unit<double>                              unit(size);
    pcx_gpu::fft_unit<double, pcx_gpu::fft_order::bit_reversed> unit_br(size);
    unit(out, v);
    unit_br(v);
    pcx_synchronize(nullptr);
    double e1 = err(out.to_host(), false), e2 = err(v.to_host(), true);
    CHECK(e1 < tol, "fft<double> size %zu pack %zu: rel L2 %.3g", size, PackSize, e1);
    CHECK(e2 < tol, "fft<double, bit_reversed> size %zu pack %zu: rel L2 %.3g", size, PackSize, e2);
    unit.ifft(out); // back to the input
    unit_br.ifft(v);
    pcx_synchronize(nullptr);
    auto back1 = out.to_host(), back2 = v.to_host();
    double n1 = 0, n2 = 0, d = 0;
    for (size_t i = 0; i < size; ++i) {
        n1 += std::norm(back1[i] - in[i]);
        n2 += std::norm(back2[i] - in[i]);
        d += std::norm(in[i]);
    }
    CHECK(std::sqrt(n1 / d) < 2 * tol && std::sqrt(n2 / d) < 2 * tol, "ifft<double> round trip size %zu: %.3g %.3g",
          size, std::sqrt(n1 / d), std::sqrt(n2 / d));
}

// Batches as two-axis md::dynamic_storage (mdstorage.hpp:1478-1499): the row stride follows
// calc_storage_size (mdstorage.hpp:879-885); one fft_unit call transforms every slice, with the
// same results as one call per slice, and zero-extends from a shorter source storage.
static void test_md_storage() {
    using pcx_gpu::md::dynamic_storage;
    // stride = ceil(2 * inner / lcm(2P, A)) * lcm(2P, A), values worked by hand
    CHECK((dynamic_storage<float, 8, 8>(1001, 0).row_stride() == 2016), "stride 1001/8/8");
    CHECK((dynamic_storage<float, 1, 8>(5, 0).row_stride() == 16), "stride 5/1/8");
    CHECK((dynamic_storage<float, 16, 8>(17, 0).row_stride() == 64), "stride 17/16/8");
    CHECK((dynamic_storage<double, 4, 4>(3, 0).row_stride() == 8), "stride 3/4/4");
    CHECK((dynamic_storage<double, 2, 32>(33, 0).row_stride() == 96), "stride 33/2/32");

    const size_t n = 1024, rows = 37, short_n = 1001;
    dynamic_storage<float, 8, 8> src(short_n, rows); // padded rows: 2016 reals for 2002 used
    dynamic_storage<float, 8, 8> dst(n, rows);
    std::vector<std::vector<cxf>> host(rows);
    for (size_t r = 0; r < rows; ++r) {
        std::vector<float> re(short_n), im(short_n);
        orc_gen_uniform(0x5EED0700 + r, re.data(), im.data(), (int64_t)short_n);
        host[r].resize(short_n);
        for (size_t i = 0; i < short_n; ++i)
            host[r][i] = {re[i], im[i]};
        src.assign(r, host[r].data(), short_n);
    }
    pcx_gpu::fft_unit<float> unit(n);
    unit(dst, src); // every slice, zero-extended
    pcx_synchronize(nullptr);
    double worst = 0;
    for (size_t r = 0; r < rows; ++r) {
        std::vector<cxf> in(n, cxf(0, 0));
        std::copy(host[r].begin(), host[r].end(), in.begin());
        worst = std::max(worst, rel_l2(dst.to_host(r), oracle_fft(0, false, true, in)));
    }
    CHECK(worst < 2 * 10 * std::numeric_limits<float>::epsilon(), "md storage batch vs oracle: %.3g", worst);
    // the batched call equals one call per slice, bit for bit
    dynamic_storage<float, 8, 8> one(n, rows);
    for (size_t r = 0; r < rows; ++r) {
        auto d = one[r];
        auto s = src[r];
        unit(d, s);
    }
    pcx_synchronize(nullptr);
    bool same = true;
    for (size_t r = 0; r < rows; ++r)
        same = same && one.to_host(r) == dst.to_host(r);
    CHECK(same, "md storage batch differs from per-slice calls");
    unit.ifft(dst); // back, in place, every slice
    pcx_synchronize(nullptr);
    worst = 0;
    for (size_t r = 0; r < rows; ++r) {
        auto back = dst.to_host(r);
        double num = 0, den = 0;
        for (size_t i = 0; i < n; ++i) {
            cxd want = i < short_n ? cxd(host[r][i]) : cxd(0, 0);
            num += std::norm(cxd(back[i]) - want);
            den += std::norm(want);
        }
        worst = std::max(worst, std::sqrt(num / den));
    }
    CHECK(worst < 4 * 10 * std::numeric_limits<float>::epsilon(), "md storage round trip: %.3g", worst);
    // double slices through the four-step
    const size_t nd = 1 << 16, rd = 3;
    dynamic_storage<double> sd(nd, rd);
    std::vector<std::vector<cxd>> hd(rd, std::vector<cxd>(nd));
    for (size_t r = 0; r < rd; ++r) {
        for (size_t i = 0; i < nd; ++i)
            hd[r][i] = std::exp(cxd(0., 2 * M_PI * (double)i * (13.37 + (double)r) / (double)nd));
        sd.assign(r, hd[r].data(), nd);
    }
    pcx_gpu::fft_unit<double> ud(nd);
    ud(sd);
    pcx_synchronize(nullptr);
    for (size_t r = 0; r < rd; ++r) {
        std::vector<double> xr(nd), xi(nd), yr(nd), yi(nd);
        for (size_t i = 0; i < nd; ++i)
            xr[i] = hd[r][i].real(), xi[i] = hd[r][i].imag();
        orc_fft(xr.data(), xi.data(), yr.data(), yi.data(), (int64_t)nd, 0);
        auto   got = sd.to_host(r);
        double num = 0, den = 0;
        for (size_t i = 0; i < nd; ++i) {
            num += std::norm(got[i] - cxd(yr[i], yi[i]));
            den += yr[i] * yr[i] + yi[i] * yi[i];
        }
        CHECK(std::sqrt(num / den) < 4 * 16 * std::numeric_limits<double>::epsilon(), "md storage<double> row %zu: %.3g",
              r, std::sqrt(num / den));
    }
    bool threw = false;
    try {
        dynamic_storage<float> wrong(512, 2);
        unit(wrong);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw, "slice size mismatch must throw");

    // correlation: every slice a channel, one shared kernel spectrum, in place and out of place
    const size_t L = 5001, M = 100, ch = 4;
    std::vector<float> gr(M), gi(M);
    orc_gen_uniform(0x5EED0B00, gr.data(), gi.data(), (int64_t)M);
    std::vector<cxf> g(M);
    for (size_t i = 0; i < M; ++i)
        g[i] = {gr[i], gi[i]};
    pcx_gpu::vector<float>       gv(g);
    pcx_gpu::fxcorr_unit<float>  xc(gv, 512);
    dynamic_storage<float, 8, 8> sig(L, ch), outp(L, ch);
    std::vector<std::vector<cxf>> hs(ch);
    for (size_t c = 0; c < ch; ++c) {
        std::vector<float> fr(L), fi(L);
        orc_gen_normal(0x5EED0C00 + c, 1.0, fr.data(), fi.data(), (int64_t)L);
        hs[c].resize(L);
        for (size_t i = 0; i < L; ++i)
            hs[c][i] = {fr[i], fi[i]};
        sig.assign(c, hs[c].data(), L);
    }
    xc(outp, sig);
    xc(sig);
    pcx_synchronize(nullptr);
    for (size_t c = 0; c < ch; ++c) {
        std::vector<double> fr(L), fi(L), g2r(M), g2i(M), outr(L), outi(L);
        for (size_t i = 0; i < L; ++i)
            fr[i] = hs[c][i].real(), fi[i] = hs[c][i].imag();
        for (size_t i = 0; i < M; ++i)
            g2r[i] = gr[i], g2i[i] = gi[i];
        orc_xcorr_naive(fr.data(), fi.data(), (int64_t)L, g2r.data(), g2i.data(), (int64_t)M, outr.data(), outi.data());
        std::vector<cxd> want(L);
        for (size_t i = 0; i < L; ++i)
            want[i] = {outr[i], outi[i]};
        double e1 = rel_l2(outp.to_host(c), want), e2 = rel_l2(sig.to_host(c), want);
        CHECK(e1 < 4e-6 && e2 < 4e-6, "md storage fxcorr channel %zu: %.3g %.3g", c, e1, e2);
    }
}

// The reference's template surface (fft.hpp:565-605, 2168-2172; fxcorr.hpp:44-91): a unit with a
// fixed SubSize and its one-argument constructor, non-default node-size strategies, fft_unit_par
// with all four parameters and with double, and fxcorr_unit with a driven FFT_ and with a
// TmpFactory_.  Every form is checked against the default unit or the FP64 oracle.
// a pack-4 view of device memory (any type with real_type, pack_size, size() and data() is a vector)
struct pack4_view {
    float* p;
    size_t n;
    using real_type                   = float;
    static constexpr size_t pack_size = 4;
    size_t       size() const { return n; }
    float*       data() { return p; }
    const float* data() const { return p; }
};

struct driven_fft : pcx_gpu::fft_unit<float, pcx_gpu::fft_order::bit_reversed> {
    using pcx_gpu::fft_unit<float, pcx_gpu::fft_order::bit_reversed>::fft_unit;
};

static void test_template_surface() {
    using namespace pcx_gpu;
    const size_t n = 4096;
    std::vector<cxf> x(n);
    for (size_t i = 0; i < n; ++i)
        x[i] = cxf(std::sin(0.3f * (float)i), std::cos(0.7f * (float)i + 1.f));
    vector<float> a(x), b(x), c(x);
    fft_unit<float, fft_order::normal> dyn(n, 256);
    fft_unit<float, fft_order::normal, device_allocator<float>, 2048> fixed(n);
    fft_unit<float, fft_order::normal, device_allocator<float>, dynamic_size, 8, strategy8, strategy48> strat(n, 64);
    dyn(a);
    fixed(b);
    strat(c);
    auto ha = a.to_host(), hb = b.to_host(), hc = c.to_host();
    bool same = true;
    for (size_t i = 0; i < n; ++i)
        same = same && ha[i] == hb[i] && ha[i] == hc[i];
    CHECK(same, "fixed SubSize and strategy units equal the default unit");
    CHECK(rel_l2(ha, oracle_fft(0, false, true, x)) < 2 * 12 * std::numeric_limits<float>::epsilon(),
          "template surface unit vs oracle");

    // fft_unit_par<double, Order, Allocator, 4>, across 32 vectors of 50 elements (pack 4)
    const size_t                    m = 32, len = 50;
    std::vector<std::vector<cxd>>   host(m, std::vector<cxd>(len));
    std::vector<vector<double>>     src, dst;
    std::vector<vector<double>*>    dptr;
    for (size_t i = 0; i < m; ++i) {
        for (size_t k = 0; k < len; ++k)
            host[i][k] = cxd(std::sin(0.37 * (double)(i * len + k)), std::cos(0.11 * (double)(i + 3 * k))) *
                         (1.0 + std::ldexp(1.0, -40));
        src.emplace_back(host[i]);
        dst.emplace_back(len);
    }
    for (auto& v : dst)
        dptr.push_back(&v); // a range of pointers to vectors (fft.hpp:2221-2228)
    fft_unit_par<double, fft_order::normal, device_allocator<double>, 4> par(m);
    par(dptr, src);
    pcx_synchronize(nullptr);
    double num = 0, den = 0;
    for (size_t j = 0; j < m; ++j) {
        auto got = dst[j].to_host();
        for (size_t k = 0; k < len; ++k) {
            long double wr = 0, wi = 0;
            for (size_t i = 0; i < m; ++i) {
                long double ang = -2.0L * 3.141592653589793238462643383279L * (long double)((i * j) % m) / (long double)m;
                wr += (long double)host[i][k].real() * std::cos(ang) - (long double)host[i][k].imag() * std::sin(ang);
                wi += (long double)host[i][k].real() * std::sin(ang) + (long double)host[i][k].imag() * std::cos(ang);
            }
            num += std::norm(got[k] - cxd((double)wr, (double)wi));
            den += (double)(wr * wr + wi * wi);
        }
    }
    CHECK(std::sqrt(num / den) < 4 * 5 * std::numeric_limits<double>::epsilon(), "fft_unit_par<double> rel L2 %g",
          std::sqrt(num / den));
    // differing packs with a destination on a source: the reference's in-place repack refusal
    std::vector<vector<float, 8>> s8;
    for (size_t i = 0; i < 8; ++i)
        s8.emplace_back(16);
    bool threw = false;
    try {
        std::vector<pack4_view> dv;
        for (auto& v : s8)
            dv.push_back({v.data(), 16});
        fft_unit_par<float> p8(8);
        p8(dv, s8);
    } catch (const std::invalid_argument& e) {
        threw = std::string(e.what()).rfind("cannot perform in-place repack", 0) == 0;
    }
    CHECK(threw, "fft_unit_par refuses a destination that overlaps a source at another pack");

    // fxcorr_unit: fused (default FFT_), driven FFT_ (a user type with fft_raw / ifft_raw), and a
    // TmpFactory_ supplying the fused path's scratch; in place and out of place; block sizes
    // 1024 and 2^15 (above the fused kernel's 16384)
    for (size_t N : {size_t(1024), size_t(1) << 15}) {
        const size_t L = 70000, M = N == 1024 ? 300 : 20000;
        std::vector<float> sr(L), si(L), kr(M), ki(M);
        orc_gen_normal(0x5EED0D00 + N, 0.7071067811865476, sr.data(), si.data(), (int64_t)L);
        orc_gen_uniform(0x5EED0D01 + N, kr.data(), ki.data(), (int64_t)M);
        std::vector<cxf> sig(L), ker(M);
        for (size_t i = 0; i < L; ++i)
            sig[i] = {sr[i], si[i]};
        for (size_t i = 0; i < M; ++i)
            ker[i] = std::polar(1.0f, 3.14159265f * kr[i]);
        std::vector<double> fa(L), fb(L), ga(M), gb(M), o1(L), o2(L);
        for (size_t i = 0; i < L; ++i)
            fa[i] = sig[i].real(), fb[i] = sig[i].imag();
        for (size_t i = 0; i < M; ++i)
            ga[i] = ker[i].real(), gb[i] = ker[i].imag();
        orc_xcorr_naive(fa.data(), fb.data(), (int64_t)L, ga.data(), gb.data(), (int64_t)M, o1.data(), o2.data());
        std::vector<cxd> want(L);
        for (size_t i = 0; i < L; ++i)
            want[i] = {o1[i], o2[i]};
        vector<float> kv(ker);
        using factory = detail_::pseudo_vector_factory<float>;
        fxcorr_unit<float> fused(kv, N);
        fxcorr_unit<float, device_allocator<float>, driven_fft> driven(kv, N);
        fxcorr_unit<float, device_allocator<float>, fft_unit<float, fft_order::unordered>, factory> with_factory(kv, N);
        vector<float> v1(sig), v2(sig), v3(sig), o(L);
        fused(v1);
        driven(v2);
        with_factory(v3);
        with_factory(o, vector<float>(sig));
        pcx_synchronize(nullptr);
        auto h1 = v1.to_host(), h2 = v2.to_host(), h3 = v3.to_host(), h4 = o.to_host();
        CHECK(rel_l2(h1, want) < 4e-6, "fxcorr fused N=%zu err %g", N, rel_l2(h1, want));
        CHECK(rel_l2(h2, want) < 4e-6, "fxcorr driven FFT_ N=%zu err %g", N, rel_l2(h2, want));
        CHECK(rel_l2(h3, want) < 4e-6, "fxcorr TmpFactory_ in place N=%zu err %g", N, rel_l2(h3, want));
        CHECK(rel_l2(h4, want) < 4e-6, "fxcorr TmpFactory_ out of place N=%zu err %g", N, rel_l2(h4, want));
        bool eq = true;
        for (size_t i = 0; i < L; ++i)
            eq = eq && h1[i] == h3[i];
        CHECK(eq, "fxcorr TmpFactory_ scratch gives the own-buffer result bit for bit (N=%zu)", N);
    }
}

int main() {
    for (size_t n = 64; n <= 16384; n *= 2) { // test_fft.cpp:779-798 runs 2^6 .. 2^14
        test_fft_float<8>(n);
        test_fft_float<1>(n);
        test_fftu_float<8>(n);
    }
    test_fft_float<8>(1 << 16);
    test_fftu_float<8>(1 << 16);
    for (size_t n = 2; n <= (1 << 20); n *= 4) {
        test_fft_double<4>(n);
        test_fft_double<1>(2 * n);
    }
    test_md_storage();
    test_zero_extension();
    test_errors();
    test_fxcorr();
    test_fft_par();
    test_template_surface();
    for (size_t n = 1; n < 32; ++n) { // test_op.cpp:241-252
        test_bin_ops<float, 1>(n);
        test_bin_ops<float, 8>(n);
        test_bin_ops<float, 32>(n);
        test_bin_ops<double, 1>(n);
        test_bin_ops<double, 4>(n);
        test_bin_ops<double, 32>(n);
    }
    std::printf("%s: %d failed checks\n", g_fail ? "FAIL" : "OK", g_fail);
    return g_fail;
}
