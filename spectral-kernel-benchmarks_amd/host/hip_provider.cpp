// C++20 host mirror of the skbench providers over the C ABI of libskb_hip.so.
#include "hip_provider.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstddef>
#include <cstring>
#include <iomanip>
#include <new>
#include <sstream>
#include <stdexcept>

namespace skbhip {

static_assert(sizeof(Complex) == 2 * sizeof(double), "Complex must be two doubles");
static_assert(sizeof(Complex) == sizeof(std::complex<double>), "Complex must match std::complex<double>");
static_assert(sizeof(skb_mode_group) == sizeof(VerticalModeGroup), "group records must be layout-identical");

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

std::string last_error() {
    char buf[1024];
    skb_last_error(buf, sizeof buf);
    return buf;
}

// Status -> the reference's exception classes (src/fftw_provider.cpp:468-487).
void check(int status) {
    switch (status) {
        case SKB_OK: return;
        case SKB_INVALID_ARGUMENT: throw std::invalid_argument(last_error());
        case SKB_LOGIC_ERROR: throw std::logic_error(last_error());
        case SKB_BAD_ALLOC: throw std::bad_alloc();
        default: throw std::runtime_error(last_error());
    }
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

skb_workload to_c(const Workload& w) {
    skb_workload c{};
    c.nx = w.nx;
    c.ny = w.ny;
    c.nz = w.nz;
    c.fields = w.fields;
    c.lx = w.lx;
    c.ly = w.ly;
    c.antialias = w.antialias ? 1 : 0;
    return c;
}

std::vector<skb_retained_mode> to_c(const std::vector<RetainedMode>& modes) {
    std::vector<skb_retained_mode> out(modes.size());
    for (std::size_t i = 0; i < modes.size(); ++i) {
        out[i].k = modes[i].k;
        out[i].l = modes[i].l;
        out[i].stored_kx = modes[i].storedKx;
        out[i].stored_ky = modes[i].storedKy;
        out[i].conjugates_stored_value = modes[i].conjugatesStoredValue ? 1 : 0;
        out[i].radial = modes[i].radialMode;
    }
    return out;
}

int c_precision(Precision p) { return p == Precision::float64 ? SKB_F64 : SKB_F32; }

std::string fnv(std::uint64_t h) {
    std::ostringstream s;
    s << "fnv1a64:" << std::hex << std::setfill('0') << std::setw(16) << h;
    return s.str();
}

// Device buffer plus host staging for the adapter scope.  FP64 host data is
// converted to the plan precision on the host before upload.
struct Staged {
    void* dev = nullptr;
    std::size_t elements = 0;  // real scalars
    bool fp32 = false;
    std::vector<float> host32;

    Staged() = default;
    Staged(const Staged&) = delete;
    Staged& operator=(const Staged&) = delete;
    ~Staged() {
        if (dev) (void)hipFree(dev);
    }
    void allocate(std::size_t scalars, bool single) {
        elements = scalars;
        fp32 = single;
        if (hipMalloc(&dev, scalars * (single ? 4 : 8)) != hipSuccess) {
            dev = nullptr;
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        if (single) host32.resize(scalars);
    }
    void upload(const double* src, hipStream_t s) {
        if (fp32) {
            for (std::size_t i = 0; i < elements; ++i) host32[i] = static_cast<float>(src[i]);
            hip_check(hipMemcpyAsync(dev, host32.data(), elements * 4, hipMemcpyHostToDevice, s), "upload");
            hip_check(hipStreamSynchronize(s), "upload sync");
        } else {
            hip_check(hipMemcpyAsync(dev, src, elements * 8, hipMemcpyHostToDevice, s), "upload");
            hip_check(hipStreamSynchronize(s), "upload sync");  // the caller may reuse src on return
        }
    }
    void download(double* dst, hipStream_t s) {
        if (fp32) {
            hip_check(hipMemcpyAsync(host32.data(), dev, elements * 4, hipMemcpyDeviceToHost, s), "download");
            hip_check(hipStreamSynchronize(s), "download sync");
            for (std::size_t i = 0; i < elements; ++i) dst[i] = host32[i];
        } else {
            hip_check(hipMemcpyAsync(dst, dev, elements * 8, hipMemcpyDeviceToHost, s), "download");
            hip_check(hipStreamSynchronize(s), "download sync");
        }
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
};

}  // namespace

// ---- model ----------------------------------------------------------------------
std::vector<RetainedMode> retainedHorizontalModes(const Workload& workload) {
    const auto w = to_c(workload);
    std::uint64_t n = 0;
    check(skb_retained_modes(&w, nullptr, 0, &n));
    std::vector<skb_retained_mode> c(n);
    check(skb_retained_modes(&w, c.data(), n, &n));
    std::vector<RetainedMode> out(n);
    for (std::size_t i = 0; i < n; ++i)
        out[i] = {c[i].k, c[i].l, c[i].stored_kx, c[i].stored_ky, c[i].conjugates_stored_value != 0, c[i].radial};
    return out;
}

std::vector<VerticalModeGroup> squaredWavenumberGroups(const std::vector<RetainedMode>& modes) {
    const auto c = to_c(modes);
    std::vector<VerticalModeGroup> out(modes.size());
    std::uint64_t n = 0;
    check(skb_k2_groups(c.data(), c.size(), reinterpret_cast<skb_mode_group*>(out.data()), out.size(), &n));
    out.resize(n);
    return out;
}

std::string modeOrderHash(const std::vector<RetainedMode>& modes) {
    const auto c = to_c(modes);
    return fnv(skb_mode_order_hash(c.data(), c.size()));
}

std::string verticalModeGroupHash(const std::vector<VerticalModeGroup>& groups) {
    return fnv(skb_group_order_hash(reinterpret_cast<const skb_mode_group*>(groups.data()), groups.size()));
}

std::string wvmSpectrumOrderHash(const Workload& workload) {
    const auto w = to_c(workload);
    return fnv(skb_wvm_order_hash(&w));
}

GroupedVerticalOperators squaredWavenumberVerticalFixture(const Workload& workload,
                                                          const std::vector<RetainedMode>& modes) {
    if (workload.lx != workload.ly)
        throw std::invalid_argument("The exact integer K-squared fixture currently requires a square horizontal domain.");
    GroupedVerticalOperators ops;
    ops.id = "synthetic-k2-grouped-orthonormal-dct2-pair-rotation-v1";
    ops.nz = workload.nz;
    ops.nj = workload.retainedVerticalModes();
    ops.groups = squaredWavenumberGroups(modes);
    ops.forward.resize(ops.groups.size() * ops.nj * ops.nz);
    ops.inverse.resize(ops.forward.size());
    check(skb_k2_vertical_family(ops.nz, ops.nj, reinterpret_cast<const skb_mode_group*>(ops.groups.data()),
                                 ops.groups.size(), ops.forward.data(), ops.inverse.data()));
    return ops;
}

GroupedVerticalOperators binnedQrVerticalFixture(const Workload& workload, const std::vector<RetainedMode>& modes,
                                                 std::size_t bins, std::uint64_t seed) {
    GroupedVerticalOperators ops;
    ops.id = "k2-binned-qr-gaussian-v1";
    ops.nz = workload.nz;
    ops.nj = workload.nz;
    const auto c = to_c(modes);
    ops.groups.resize(modes.size());
    std::uint64_t n = 0;
    check(skb_k2_binned_groups(c.data(), c.size(), bins, reinterpret_cast<skb_mode_group*>(ops.groups.data()),
                               ops.groups.size(), &n));
    ops.groups.resize(n);
    ops.forward.resize(n * ops.nz * ops.nz);
    ops.inverse.resize(ops.forward.size());
    check(skb_qr_gaussian_family(ops.nz, n, seed, ops.forward.data(), ops.inverse.data()));
    return ops;
}

std::vector<double> makeFixture(const Workload& workload, FixtureKind fixture, std::uint64_t seed) {
    const auto w = to_c(workload);
    std::vector<double> out(workload.realElements());
    check(skb_make_fixture(&w, static_cast<int>(fixture), seed, out.data()));
    return out;
}

// ---- HipProvider -------------------------------------------------------------------
struct HipProvider::Impl {
    Workload w;
    Precision precision;
    DataLayout layout;
    skb_fft_plan* plan = nullptr;
    bool supported = true;
    std::string capability = "hip provider: gfx950 Stockham r2c/c2r in WVM order";
    double allocation_seconds = 0, planning_seconds = 0;
    Stream stream;
    Staged real, spectrum, retained, spectrum_re, spectrum_im, retained_re, retained_im;
    std::size_t retained_modes = 0;

    Impl(const Workload& wl, DataLayout l, Precision p) : w(wl), precision(p), layout(l) {
        const auto c = to_c(wl);
        const auto t0 = Clock::now();
        const int st = skb_fft_plan_create(&c, c_precision(p), &plan);
        planning_seconds = seconds_since(t0);
        if (st == SKB_UNSUPPORTED) {
            supported = false;
            capability = "unsupported: " + last_error();
            return;
        }
        check(st);
        const auto t1 = Clock::now();
        const bool single = p == Precision::float32;
        real.allocate(w.realElements(), single);
        spectrum.allocate(2 * w.spectrumElements(), single);
        if (layout == DataLayout::split) {
            spectrum_re.allocate(w.spectrumElements(), single);
            spectrum_im.allocate(w.spectrumElements(), single);
        }
        allocation_seconds = seconds_since(t1);
    }
    ~Impl() {
        if (plan) skb_fft_plan_destroy(plan);
    }
    void require() const {
        if (!supported) throw std::logic_error(capability);
    }
    // src/fftw_provider.cpp:468-472
    void require_layout(DataLayout expected) const {
        require();
        if (layout != expected)
            throw std::logic_error("FFTW execution entry point does not match the planned data layout.");
    }
    // Retained staging follows the mode list of the call (a provider serves
    // any retained set); it is reallocated only when the mode count changes.
    void ensure_retained(std::size_t modes) {
        if (retained_modes == modes) return;
        for (Staged* b : {&retained, &retained_re, &retained_im}) {
            b->~Staged();
            new (b) Staged();
        }
        const bool single = precision == Precision::float32;
        retained.allocate(2 * modes * w.planes(), single);
        retained_re.allocate(modes * w.planes(), single);
        retained_im.allocate(modes * w.planes(), single);
        retained_modes = modes;
    }
    int cp() const { return c_precision(precision); }
};

HipProvider::HipProvider(const Workload& workload, Precision precision)
    : impl_(std::make_unique<Impl>(workload, DataLayout::interleaved, precision)) {}
HipProvider::HipProvider(const Workload& workload, DataLayout layout, Precision precision)
    : impl_(std::make_unique<Impl>(workload, layout, precision)) {}
HipProvider::~HipProvider() = default;
HipProvider::HipProvider(HipProvider&&) noexcept = default;
HipProvider& HipProvider::operator=(HipProvider&&) noexcept = default;

void HipProvider::forward(const double* input, Complex* wvmSpectrum) {
    auto& m = *impl_;
    m.require_layout(DataLayout::interleaved);
    m.real.upload(input, m.stream.s);
    check(skb_fft_forward(m.plan, m.real.dev, m.spectrum.dev, m.stream.s));
    m.spectrum.download(reinterpret_cast<double*>(wvmSpectrum), m.stream.s);
}

void HipProvider::inverse(Complex* wvmSpectrum, double* output) {
    auto& m = *impl_;
    m.require_layout(DataLayout::interleaved);
    m.spectrum.upload(reinterpret_cast<const double*>(wvmSpectrum), m.stream.s);
    check(skb_fft_inverse(m.plan, m.spectrum.dev, m.real.dev, m.stream.s));
    m.real.download(output, m.stream.s);
}

void HipProvider::gatherRetainedOuter(const std::vector<RetainedMode>& modes, const Complex* wvmSpectrum,
                                      Complex* retainedSpectrum) {
    auto& m = *impl_;
    m.require();
    const auto c = to_c(modes);
    m.ensure_retained(modes.size());
    m.spectrum.upload(reinterpret_cast<const double*>(wvmSpectrum), m.stream.s);
    check(skb_fft_gather_retained(m.plan, c.data(), c.size(), m.spectrum.dev, m.retained.dev, m.stream.s));
    m.retained.download(reinterpret_cast<double*>(retainedSpectrum), m.stream.s);
}

void HipProvider::embedRetainedOuter(const std::vector<RetainedMode>& modes, const Complex* retainedSpectrum,
                                     Complex* wvmSpectrum) {
    auto& m = *impl_;
    m.require();
    const auto c = to_c(modes);
    m.ensure_retained(modes.size());
    m.retained.upload(reinterpret_cast<const double*>(retainedSpectrum), m.stream.s);
    check(skb_fft_embed_retained(m.plan, c.data(), c.size(), m.retained.dev, m.spectrum.dev, m.stream.s));
    m.spectrum.download(reinterpret_cast<double*>(wvmSpectrum), m.stream.s);
}

// Split forms: the device kernels read and write interleaved arrays, so the
// split arrays are (de)interleaved on the device around them (skb_interleave /
// skb_deinterleave, which also applies the gather's scale).
void HipProvider::gatherRetainedToSplitOuter(const std::vector<RetainedMode>& modes, const Complex* spectrum,
                                             double* retainedReal, double* retainedImag, double scale) {
    auto& m = *impl_;
    m.require();
    const auto c = to_c(modes);
    m.ensure_retained(modes.size());
    const std::uint64_t n = modes.size() * m.w.planes();
    m.spectrum.upload(reinterpret_cast<const double*>(spectrum), m.stream.s);
    check(skb_fft_gather_retained(m.plan, c.data(), c.size(), m.spectrum.dev, m.retained.dev, m.stream.s));
    check(skb_deinterleave(m.cp(), n, m.retained.dev, m.retained_re.dev, m.retained_im.dev, scale, m.stream.s));
    m.retained_re.download(retainedReal, m.stream.s);
    m.retained_im.download(retainedImag, m.stream.s);
}

void HipProvider::embedRetainedFromSplitOuter(const std::vector<RetainedMode>& modes, const double* retainedReal,
                                              const double* retainedImag, Complex* spectrum) {
    auto& m = *impl_;
    m.require();
    const auto c = to_c(modes);
    m.ensure_retained(modes.size());
    const std::uint64_t n = modes.size() * m.w.planes();
    m.retained_re.upload(retainedReal, m.stream.s);
    m.retained_im.upload(retainedImag, m.stream.s);
    check(skb_interleave(m.cp(), n, m.retained_re.dev, m.retained_im.dev, m.retained.dev, m.stream.s));
    check(skb_fft_embed_retained(m.plan, c.data(), c.size(), m.retained.dev, m.spectrum.dev, m.stream.s));
    m.spectrum.download(reinterpret_cast<double*>(spectrum), m.stream.s);
}

void HipProvider::scaleRetainedSplitOuter(const std::vector<RetainedMode>& modes, double* retainedReal,
                                          double* retainedImag, double scale) {
    auto& m = *impl_;
    m.require();
    m.ensure_retained(modes.size());
    const std::uint64_t n = modes.size() * m.w.planes();
    m.retained_re.upload(retainedReal, m.stream.s);
    m.retained_im.upload(retainedImag, m.stream.s);
    check(skb_scale_split(m.cp(), n, m.retained_re.dev, m.retained_im.dev, scale, m.stream.s));
    m.retained_re.download(retainedReal, m.stream.s);
    m.retained_im.download(retainedImag, m.stream.s);
}

void HipProvider::forwardSplit(const double* input, double* wvmSpectrumReal, double* wvmSpectrumImag) {
    auto& m = *impl_;
    m.require_layout(DataLayout::split);
    m.real.upload(input, m.stream.s);
    check(skb_fft_forward(m.plan, m.real.dev, m.spectrum.dev, m.stream.s));
    check(skb_deinterleave(m.cp(), m.w.spectrumElements(), m.spectrum.dev, m.spectrum_re.dev, m.spectrum_im.dev, 1.0,
                           m.stream.s));
    m.spectrum_re.download(wvmSpectrumReal, m.stream.s);
    m.spectrum_im.download(wvmSpectrumImag, m.stream.s);
}

void HipProvider::inverseSplit(double* wvmSpectrumReal, double* wvmSpectrumImag, double* output) {
    auto& m = *impl_;
    m.require_layout(DataLayout::split);
    m.spectrum_re.upload(wvmSpectrumReal, m.stream.s);
    m.spectrum_im.upload(wvmSpectrumImag, m.stream.s);
    check(skb_interleave(m.cp(), m.w.spectrumElements(), m.spectrum_re.dev, m.spectrum_im.dev, m.spectrum.dev,
                         m.stream.s));
    check(skb_fft_inverse(m.plan, m.spectrum.dev, m.real.dev, m.stream.s));
    m.real.download(output, m.stream.s);
}

void HipProvider::gatherRetainedSplitOuter(const std::vector<RetainedMode>& modes, const double* spectrumReal,
                                           const double* spectrumImag, double* retainedReal, double* retainedImag) {
    auto& m = *impl_;
    m.require_layout(DataLayout::split);
    const auto c = to_c(modes);
    m.ensure_retained(modes.size());
    const std::uint64_t n = modes.size() * m.w.planes();
    m.spectrum_re.upload(spectrumReal, m.stream.s);
    m.spectrum_im.upload(spectrumImag, m.stream.s);
    check(skb_interleave(m.cp(), m.w.spectrumElements(), m.spectrum_re.dev, m.spectrum_im.dev, m.spectrum.dev,
                         m.stream.s));
    check(skb_fft_gather_retained(m.plan, c.data(), c.size(), m.spectrum.dev, m.retained.dev, m.stream.s));
    check(skb_deinterleave(m.cp(), n, m.retained.dev, m.retained_re.dev, m.retained_im.dev, 1.0, m.stream.s));
    m.retained_re.download(retainedReal, m.stream.s);
    m.retained_im.download(retainedImag, m.stream.s);
}

void HipProvider::embedRetainedSplitOuter(const std::vector<RetainedMode>& modes, const double* retainedReal,
                                          const double* retainedImag, double* spectrumReal, double* spectrumImag) {
    auto& m = *impl_;
    m.require_layout(DataLayout::split);
    const auto c = to_c(modes);
    m.ensure_retained(modes.size());
    const std::uint64_t n = modes.size() * m.w.planes();
    m.retained_re.upload(retainedReal, m.stream.s);
    m.retained_im.upload(retainedImag, m.stream.s);
    check(skb_interleave(m.cp(), n, m.retained_re.dev, m.retained_im.dev, m.retained.dev, m.stream.s));
    check(skb_fft_embed_retained(m.plan, c.data(), c.size(), m.retained.dev, m.spectrum.dev, m.stream.s));
    check(skb_deinterleave(m.cp(), m.w.spectrumElements(), m.spectrum.dev, m.spectrum_re.dev, m.spectrum_im.dev, 1.0,
                           m.stream.s));
    m.spectrum_re.download(spectrumReal, m.stream.s);
    m.spectrum_im.download(spectrumImag, m.stream.s);
}

DataLayout HipProvider::layout() const noexcept { return impl_->layout; }
std::string HipProvider::splitInPlaceWvmOrderCapability() const {
    return "unsupported: the device transforms read and write interleaved spectra; the split entry points stage "
           "through the provider's interleaved spectrum out of place";
}

void HipProvider::forwardDevice(const void* real, void* wvm, void* stream) {
    impl_->require();
    check(skb_fft_forward(impl_->plan, real, wvm, stream));
}

void HipProvider::inverseDevice(void* wvm, void* real, void* stream) {
    impl_->require();
    check(skb_fft_inverse(impl_->plan, wvm, real, stream));
}

void HipProvider::executeSchedulerNoop() { hip_check(hipStreamSynchronize(impl_->stream.s), "sync"); }
bool HipProvider::supported() const noexcept { return impl_->supported; }
std::string HipProvider::capability() const { return impl_->capability; }
double HipProvider::allocationSeconds() const noexcept { return impl_->allocation_seconds; }
double HipProvider::planningSeconds() const noexcept { return impl_->planning_seconds; }
std::size_t HipProvider::planningBytes() const noexcept { return (impl_->w.nx + impl_->w.ny) * 16; }
std::size_t HipProvider::minimumAlignmentBytes() const noexcept { return 256; }
std::string HipProvider::libraryIdentity() const { return "libskb_hip.so"; }
std::string HipProvider::version() const { return skb_version(); }

// ---- HipPrunedProvider ----------------------------------------------------------------
struct HipPrunedProvider::Impl {
    Workload w;
    std::size_t modes;
    skb_pruned_plan* plan = nullptr;
    Stream stream;
    Staged real, retained, split_re, split_im;
    std::vector<double> host_split;  // interleave / deinterleave of the split adapters

    Impl(const Workload& wl, const std::vector<RetainedMode>& m, Precision p) : w(wl), modes(m.size()) {
        const auto c = to_c(m);
        const auto cw = to_c(wl);
        check(skb_pruned_plan_create(&cw, c.data(), c.size(), c_precision(p), &plan));
        const bool single = p == Precision::float32;
        real.allocate(w.realElements(), single);
        retained.allocate(2 * modes * w.planes(), single);
        split_re.allocate(modes * w.planes(), single);
        split_im.allocate(modes * w.planes(), single);
        host_split.resize(2 * modes * w.planes());
    }
    ~Impl() {
        if (plan) skb_pruned_plan_destroy(plan);
    }
};

HipPrunedProvider::HipPrunedProvider(const Workload& workload, const std::vector<RetainedMode>& modes,
                                     Precision precision)
    : impl_(std::make_unique<Impl>(workload, modes, precision)) {}
HipPrunedProvider::~HipPrunedProvider() = default;
HipPrunedProvider::HipPrunedProvider(HipPrunedProvider&&) noexcept = default;
HipPrunedProvider& HipPrunedProvider::operator=(HipPrunedProvider&&) noexcept = default;

void HipPrunedProvider::forward(const double* input, Complex* retainedSpectrum) {
    executeForwardRows(input);
    executeForwardColumns();
    gatherForward(retainedSpectrum);
}

void HipPrunedProvider::inverse(const Complex* retainedSpectrum, double* output) {
    embedInverse(retainedSpectrum);
    executeInverseColumns();
    executeInverseRows(output);
}

void HipPrunedProvider::executeForwardRows(const double* input) {
    auto& m = *impl_;
    m.real.upload(input, m.stream.s);
    check(skb_pruned_forward_rows(m.plan, m.real.dev, m.stream.s));
    hip_check(hipStreamSynchronize(m.stream.s), "sync");
}

void HipPrunedProvider::executeForwardColumns() {
    auto& m = *impl_;
    check(skb_pruned_forward_columns(m.plan, m.retained.dev, m.stream.s));
    hip_check(hipStreamSynchronize(m.stream.s), "sync");
}

void HipPrunedProvider::gatherForward(Complex* retainedSpectrum) {
    auto& m = *impl_;
    m.retained.download(reinterpret_cast<double*>(retainedSpectrum), m.stream.s);
}

void HipPrunedProvider::embedInverse(const Complex* retainedSpectrum) {
    auto& m = *impl_;
    m.retained.upload(reinterpret_cast<const double*>(retainedSpectrum), m.stream.s);
}

void HipPrunedProvider::executeInverseColumns() {
    auto& m = *impl_;
    check(skb_pruned_inverse_columns(m.plan, m.retained.dev, m.stream.s));
    hip_check(hipStreamSynchronize(m.stream.s), "sync");
}

void HipPrunedProvider::executeInverseRows(double* output) {
    auto& m = *impl_;
    check(skb_pruned_inverse_rows(m.plan, m.real.dev, m.stream.s));
    m.real.download(output, m.stream.s);
}

// gatherForwardSplit / embedInverseSplit act on the interleaved device
// result of executeForwardColumns / input of executeInverseColumns; the split
// is done on the host side of the adapter (src/pruned_fftw_provider.cpp:430-440).
void HipPrunedProvider::gatherForwardSplit(double* retainedReal, double* retainedImag, double scale) {
    auto& m = *impl_;
    m.retained.download(m.host_split.data(), m.stream.s);
    const std::size_t n = m.modes * m.w.planes();
    for (std::size_t i = 0; i < n; ++i) {
        retainedReal[i] = scale * m.host_split[2 * i];
        retainedImag[i] = scale * m.host_split[2 * i + 1];
    }
}

void HipPrunedProvider::embedInverseSplit(const double* retainedReal, const double* retainedImag) {
    auto& m = *impl_;
    const std::size_t n = m.modes * m.w.planes();
    for (std::size_t i = 0; i < n; ++i) {
        m.host_split[2 * i] = retainedReal[i];
        m.host_split[2 * i + 1] = retainedImag[i];
    }
    m.retained.upload(m.host_split.data(), m.stream.s);
}

void HipPrunedProvider::forwardSplit(const double* input, double* retainedReal, double* retainedImag, double scale) {
    auto& m = *impl_;
    m.real.upload(input, m.stream.s);
    check(skb_pruned_forward_split(m.plan, m.real.dev, m.split_re.dev, m.split_im.dev, scale, m.stream.s));
    m.split_re.download(retainedReal, m.stream.s);
    m.split_im.download(retainedImag, m.stream.s);
}

void HipPrunedProvider::inverseSplit(const double* retainedReal, const double* retainedImag, double* output) {
    auto& m = *impl_;
    m.split_re.upload(retainedReal, m.stream.s);
    m.split_im.upload(retainedImag, m.stream.s);
    check(skb_pruned_inverse_split(m.plan, m.split_re.dev, m.split_im.dev, m.real.dev, m.stream.s));
    m.real.download(output, m.stream.s);
}

void HipPrunedProvider::forwardSplitDevice(const void* real, void* retainedReal, void* retainedImag, double scale,
                                           void* stream) {
    check(skb_pruned_forward_split(impl_->plan, real, retainedReal, retainedImag, scale, stream));
}

void HipPrunedProvider::inverseSplitDevice(const void* retainedReal, const void* retainedImag, void* real,
                                           void* stream) {
    check(skb_pruned_inverse_split(impl_->plan, retainedReal, retainedImag, real, stream));
}

void HipPrunedProvider::forwardDevice(const void* real, void* retained, void* stream) {
    check(skb_pruned_forward(impl_->plan, real, retained, stream));
}

void HipPrunedProvider::inverseDevice(const void* retained, void* real, void* stream) {
    check(skb_pruned_inverse(impl_->plan, retained, real, stream));
}

void HipPrunedProvider::executeSchedulerNoop() { hip_check(hipStreamSynchronize(impl_->stream.s), "sync"); }
std::size_t HipPrunedProvider::activeKxCount() const noexcept { return skb_pruned_active_kx(impl_->plan); }
std::size_t HipPrunedProvider::fullKxCount() const noexcept { return impl_->w.nxHalf(); }
std::size_t HipPrunedProvider::rowTransformsPerDirection() const noexcept { return impl_->w.ny * impl_->w.planes(); }
std::size_t HipPrunedProvider::columnTransformsPerDirection() const noexcept {
    return activeKxCount() * impl_->w.planes();
}
std::size_t HipPrunedProvider::omittedColumnTransformsPerDirection() const noexcept {
    return (fullKxCount() - activeKxCount()) * impl_->w.planes();
}
std::size_t HipPrunedProvider::scratchBytes() const noexcept { return skb_pruned_scratch_bytes(impl_->plan); }
std::string HipPrunedProvider::libraryIdentity() const { return "libskb_hip.so"; }
std::string HipPrunedProvider::version() const { return skb_version(); }

// ---- HipWvmDirectVerticalGemmProvider ----------------------------------------------
struct HipWvmDirectVerticalGemmProvider::Impl {
    Workload w;
    std::size_t nj, modal_elements;
    skb_vertical_plan* plan = nullptr;
    Stream stream;
    Staged spectrum, modal;  // full WVM arrays (the adapter transfers them whole)

    Impl(const Workload& wl, const std::vector<RetainedMode>& modes, const GroupedVerticalOperators& ops, Precision p)
        : w(wl), nj(ops.nj) {
        // src/vertical_gemm.cpp:965-972
        if (ops.nz != wl.nz || ops.nj != wl.retainedVerticalModes())
            throw std::invalid_argument("Direct WVM vertical GEMM operator dimensions do not match the workload.");
        const auto cw = to_c(wl);
        const auto cm = to_c(modes);
        check(skb_vertical_wvm_plan_create(&cw, cm.data(), cm.size(), nj,
                                           reinterpret_cast<const skb_mode_group*>(ops.groups.data()),
                                           ops.groups.size(), ops.forward.data(), ops.inverse.data(), c_precision(p),
                                           &plan));
        modal_elements = skb_vertical_wvm_modal_elements(plan);
        const bool single = p == Precision::float32;
        spectrum.allocate(2 * wl.spectrumElements(), single);
        modal.allocate(2 * modal_elements, single);
    }
    ~Impl() {
        if (plan) skb_vertical_plan_destroy(plan);
    }
};

HipWvmDirectVerticalGemmProvider::HipWvmDirectVerticalGemmProvider(const Workload& workload,
                                                                   const std::vector<RetainedMode>& modes,
                                                                   const GroupedVerticalOperators& operators,
                                                                   Precision precision)
    : impl_(std::make_unique<Impl>(workload, modes, operators, precision)) {}
HipWvmDirectVerticalGemmProvider::~HipWvmDirectVerticalGemmProvider() = default;
HipWvmDirectVerticalGemmProvider::HipWvmDirectVerticalGemmProvider(HipWvmDirectVerticalGemmProvider&&) noexcept =
    default;
HipWvmDirectVerticalGemmProvider& HipWvmDirectVerticalGemmProvider::operator=(
    HipWvmDirectVerticalGemmProvider&&) noexcept = default;

std::string HipWvmDirectVerticalGemmProvider::capability() const {
    return "supported: frequency-major direct grouped MFMA GEMM (gfx950)";
}
void HipWvmDirectVerticalGemmProvider::initializeModalOutput(Complex* fullModalSpectrum) const {
    std::fill_n(fullModalSpectrum, impl_->modal_elements, Complex{});
}
void HipWvmDirectVerticalGemmProvider::initializeSpectrumOutput(Complex* fullSpectrum) const {
    std::fill_n(fullSpectrum, impl_->w.spectrumElements(), Complex{});
}
// The adapter stages whole arrays, so entries the GEMM does not write keep
// the caller's values, as in the reference.
void HipWvmDirectVerticalGemmProvider::executeForward(const Complex* fullSpectrum, Complex* fullModalSpectrum) {
    auto& m = *impl_;
    m.spectrum.upload(reinterpret_cast<const double*>(fullSpectrum), m.stream.s);
    m.modal.upload(reinterpret_cast<const double*>(fullModalSpectrum), m.stream.s);
    check(skb_vertical_wvm_forward(m.plan, m.spectrum.dev, m.modal.dev, m.stream.s));
    m.modal.download(reinterpret_cast<double*>(fullModalSpectrum), m.stream.s);
}
void HipWvmDirectVerticalGemmProvider::executeInverse(const Complex* fullModalSpectrum, Complex* fullSpectrum) {
    auto& m = *impl_;
    m.modal.upload(reinterpret_cast<const double*>(fullModalSpectrum), m.stream.s);
    m.spectrum.upload(reinterpret_cast<const double*>(fullSpectrum), m.stream.s);
    check(skb_vertical_wvm_inverse(m.plan, m.modal.dev, m.spectrum.dev, m.stream.s));
    m.spectrum.download(reinterpret_cast<double*>(fullSpectrum), m.stream.s);
}
void HipWvmDirectVerticalGemmProvider::initializeModalOutputDevice(void* fullModalSpectrum, void* stream) const {
    check(skb_vertical_wvm_initialize_modal(impl_->plan, fullModalSpectrum, stream));
}
void HipWvmDirectVerticalGemmProvider::initializeSpectrumOutputDevice(void* fullSpectrum, void* stream) const {
    check(skb_vertical_wvm_initialize_spectrum(impl_->plan, fullSpectrum, stream));
}
void HipWvmDirectVerticalGemmProvider::executeForwardDevice(const void* fullSpectrum, void* fullModalSpectrum,
                                                            void* stream) {
    check(skb_vertical_wvm_forward(impl_->plan, fullSpectrum, fullModalSpectrum, stream));
}
void HipWvmDirectVerticalGemmProvider::executeInverseDevice(const void* fullModalSpectrum, void* fullSpectrum,
                                                            void* stream) {
    check(skb_vertical_wvm_inverse(impl_->plan, fullModalSpectrum, fullSpectrum, stream));
}
void HipWvmDirectVerticalGemmProvider::executeSchedulerNoop() {
    hip_check(hipStreamSynchronize(impl_->stream.s), "sync");
}
std::size_t HipWvmDirectVerticalGemmProvider::modalSpectrumElements() const noexcept { return impl_->modal_elements; }
std::size_t HipWvmDirectVerticalGemmProvider::matrixBytesPerDirection() const noexcept {
    return skb_vertical_matrix_bytes(impl_->plan);
}
std::string HipWvmDirectVerticalGemmProvider::libraryIdentity() const { return "libskb_hip.so"; }

// ---- HipStreamingPrunedSplitProvider ------------------------------------------------
static_assert(sizeof(HipStreamingPrunedSplitProvider::FieldView) == sizeof(skb_field_view) &&
                  offsetof(HipStreamingPrunedSplitProvider::FieldView, imaginary) ==
                      offsetof(skb_field_view, imaginary) &&
                  offsetof(HipStreamingPrunedSplitProvider::FieldView, modeStride) ==
                      offsetof(skb_field_view, mode_stride),
              "FieldView must match skb_field_view");

struct HipStreamingPrunedSplitProvider::Impl {
    Workload w;
    std::size_t modes;
    std::size_t tile;
    StreamingInversePreparationPolicy policy;
    skb_pruned_plan* plan = nullptr;
    Stream stream;
    Staged real, split_re, split_im;
    std::vector<double> host_re, host_im;

    Impl(const Workload& wl, const std::vector<RetainedMode>& m, Precision p, std::size_t tileWidth,
         StreamingInversePreparationPolicy pol)
        : w(wl), modes(m.size()), tile(tileWidth), policy(pol) {
        // src/pruned_fftw_provider.cpp:668-677
        if (tileWidth == 0 || tileWidth > 16)
            throw std::invalid_argument("Streaming pruned FFTW tile width must lie in [1, 16].");
        if (tileWidth == 1 && pol != StreamingInversePreparationPolicy::fullZero)
            throw std::invalid_argument(
                "Alternative streaming inverse preparation requires tile width greater than one.");
        const auto c = to_c(m);
        const auto cw = to_c(wl);
        check(skb_pruned_plan_create(&cw, c.data(), c.size(), c_precision(p), &plan));
        const bool single = p == Precision::float32;
        real.allocate(w.realElements(), single);
        split_re.allocate(modes * w.planes(), single);
        split_im.allocate(modes * w.planes(), single);
        host_re.resize(modes * w.planes());
        host_im.resize(modes * w.planes());
    }
    ~Impl() {
        if (plan) skb_pruned_plan_destroy(plan);
    }
    // src/pruned_fftw_provider.cpp:929-958
    template <typename V>
    void validate(const V* fields, std::size_t count, bool output) const {
        if (fields == nullptr || count != w.fields)
            throw std::invalid_argument(output
                                            ? "Streaming pruned output field views must match the workload field count."
                                            : "Streaming pruned input field views must match the workload field count.");
        for (std::size_t f = 0; f < count; ++f)
            if (fields[f].real == nullptr || fields[f].imaginary == nullptr || fields[f].modeStride < w.nz)
                throw std::invalid_argument(output ? "Streaming pruned output field view is invalid."
                                                   : "Streaming pruned input field view is invalid.");
    }
};

HipStreamingPrunedSplitProvider::HipStreamingPrunedSplitProvider(const Workload& workload,
                                                                 const std::vector<RetainedMode>& modes,
                                                                 Precision precision, std::size_t tileWidth,
                                                                 StreamingInversePreparationPolicy policy)
    : impl_(std::make_unique<Impl>(workload, modes, precision, tileWidth, policy)) {}
HipStreamingPrunedSplitProvider::~HipStreamingPrunedSplitProvider() = default;
HipStreamingPrunedSplitProvider::HipStreamingPrunedSplitProvider(HipStreamingPrunedSplitProvider&&) noexcept = default;
HipStreamingPrunedSplitProvider& HipStreamingPrunedSplitProvider::operator=(
    HipStreamingPrunedSplitProvider&&) noexcept = default;

void HipStreamingPrunedSplitProvider::forwardSplit(const double* input, double* retainedReal, double* retainedImag,
                                                   double scale) {
    auto& m = *impl_;
    m.real.upload(input, m.stream.s);
    check(skb_pruned_forward_split(m.plan, m.real.dev, m.split_re.dev, m.split_im.dev, scale, m.stream.s));
    m.split_re.download(retainedReal, m.stream.s);
    m.split_im.download(retainedImag, m.stream.s);
}

void HipStreamingPrunedSplitProvider::inverseSplit(const double* retainedReal, const double* retainedImag,
                                                   double* output) {
    auto& m = *impl_;
    m.split_re.upload(retainedReal, m.stream.s);
    m.split_im.upload(retainedImag, m.stream.s);
    check(skb_pruned_inverse_split(m.plan, m.split_re.dev, m.split_im.dev, m.real.dev, m.stream.s));
    m.real.download(output, m.stream.s);
}

// Host field views are gathered into / scattered from the plain split staging
// arrays (plane + P * mode, plane = z + Nz * field).
void HipStreamingPrunedSplitProvider::forwardSplitFields(const double* input, const FieldView* fields,
                                                         std::size_t fieldCount, double scale) {
    auto& m = *impl_;
    m.validate(fields, fieldCount, true);
    m.real.upload(input, m.stream.s);
    check(skb_pruned_forward_split(m.plan, m.real.dev, m.split_re.dev, m.split_im.dev, scale, m.stream.s));
    m.split_re.download(m.host_re.data(), m.stream.s);
    m.split_im.download(m.host_im.data(), m.stream.s);
    const std::size_t P = m.w.planes(), nz = m.w.nz;
    for (std::size_t f = 0; f < fieldCount; ++f)
        for (std::size_t mode = 0; mode < m.modes; ++mode)
            for (std::size_t z = 0; z < nz; ++z) {
                const std::size_t src = z + nz * f + P * mode, dst = z + fields[f].modeStride * mode;
                fields[f].real[dst] = m.host_re[src];
                fields[f].imaginary[dst] = m.host_im[src];
            }
}

void HipStreamingPrunedSplitProvider::inverseSplitFields(const ConstFieldView* fields, std::size_t fieldCount,
                                                         double* output) {
    auto& m = *impl_;
    m.validate(fields, fieldCount, false);
    const std::size_t P = m.w.planes(), nz = m.w.nz;
    for (std::size_t f = 0; f < fieldCount; ++f)
        for (std::size_t mode = 0; mode < m.modes; ++mode)
            for (std::size_t z = 0; z < nz; ++z) {
                const std::size_t dst = z + nz * f + P * mode, src = z + fields[f].modeStride * mode;
                m.host_re[dst] = fields[f].real[src];
                m.host_im[dst] = fields[f].imaginary[src];
            }
    m.split_re.upload(m.host_re.data(), m.stream.s);
    m.split_im.upload(m.host_im.data(), m.stream.s);
    check(skb_pruned_inverse_split(m.plan, m.split_re.dev, m.split_im.dev, m.real.dev, m.stream.s));
    m.real.download(output, m.stream.s);
}

void HipStreamingPrunedSplitProvider::forwardSplitDevice(const void* real, void* retainedReal, void* retainedImag,
                                                         double scale, void* stream) {
    check(skb_pruned_forward_split(impl_->plan, real, retainedReal, retainedImag, scale, stream));
}

void HipStreamingPrunedSplitProvider::inverseSplitDevice(const void* retainedReal, const void* retainedImag,
                                                         void* real, void* stream) {
    check(skb_pruned_inverse_split(impl_->plan, retainedReal, retainedImag, real, stream));
}

void HipStreamingPrunedSplitProvider::forwardSplitFieldsDevice(const void* real, const skb_field_view* fields,
                                                               std::size_t fieldCount, double scale, void* stream) {
    check(skb_pruned_forward_split_fields(impl_->plan, real, fields, fieldCount, scale, stream));
}

void HipStreamingPrunedSplitProvider::inverseSplitFieldsDevice(const skb_field_view* fields, std::size_t fieldCount,
                                                               void* real, void* stream) {
    check(skb_pruned_inverse_split_fields(impl_->plan, fields, fieldCount, real, stream));
}

void HipStreamingPrunedSplitProvider::executeSchedulerNoop() {
    hip_check(hipStreamSynchronize(impl_->stream.s), "sync");
}
std::size_t HipStreamingPrunedSplitProvider::activeKxCount() const noexcept {
    return skb_pruned_active_kx(impl_->plan);
}
std::size_t HipStreamingPrunedSplitProvider::fullKxCount() const noexcept { return impl_->w.nxHalf(); }
std::size_t HipStreamingPrunedSplitProvider::rowTransformsPerDirection() const noexcept {
    return impl_->w.ny * impl_->w.planes();
}
std::size_t HipStreamingPrunedSplitProvider::columnTransformsPerDirection() const noexcept {
    return activeKxCount() * impl_->w.planes();
}
std::size_t HipStreamingPrunedSplitProvider::omittedColumnTransformsPerDirection() const noexcept {
    return (fullKxCount() - activeKxCount()) * impl_->w.planes();
}
std::size_t HipStreamingPrunedSplitProvider::scratchBytes() const noexcept {
    return skb_pruned_scratch_bytes(impl_->plan);
}
std::size_t HipStreamingPrunedSplitProvider::tileWidth() const noexcept { return impl_->tile; }
StreamingInversePreparationPolicy HipStreamingPrunedSplitProvider::inversePreparationPolicy() const noexcept {
    return impl_->policy;
}
std::string HipStreamingPrunedSplitProvider::libraryIdentity() const { return "libskb_hip.so"; }
std::string HipStreamingPrunedSplitProvider::version() const { return skb_version(); }

// ---- HipInterleavedWvmFieldViewProvider ------------------------------------------------
// Host views are packed into (and unpacked from) a contiguous one-field
// spectrum on the host side of the adapter; device views go to the library's
// strided entry points unchanged.
struct HipInterleavedWvmFieldViewProvider::Impl {
    Workload w;
    VerticalGemmBufferPolicy policy;
    std::vector<std::size_t> strides;
    skb_fft_plan* plan = nullptr;
    Stream stream;
    Staged real, spectrum;
    std::vector<double> host;  // packed one-field spectrum (interleaved doubles)
    double allocation_seconds = 0, planning_seconds = 0;

    // src/fftw_provider.cpp:1066-1110
    Impl(const Workload& wl, std::vector<std::size_t> st, VerticalGemmBufferPolicy bp, Precision p)
        : w(wl), policy(bp), strides(std::move(st)) {
        if (w.fields != 1) throw std::invalid_argument("WVM field-view FFTW requires a one-field logical workload.");
        if (strides.empty())
            throw std::invalid_argument("WVM field-view FFTW requires at least one native frequency stride.");
        for (const auto s : strides)
            if (s < w.nz)
                throw std::invalid_argument("WVM field-view frequency strides must contain every vertical plane.");
        std::sort(strides.begin(), strides.end());
        strides.erase(std::unique(strides.begin(), strides.end()), strides.end());
        const auto cw = to_c(w);
        const auto t0 = Clock::now();
        check(skb_fft_plan_create(&cw, c_precision(p), &plan));
        planning_seconds = seconds_since(t0);
        const auto t1 = Clock::now();
        const bool single = p == Precision::float32;
        real.allocate(w.realElements(), single);
        spectrum.allocate(2 * w.spectrumElements(), single);
        host.resize(2 * w.spectrumElements());
        allocation_seconds = seconds_since(t1);
    }
    ~Impl() {
        if (plan) skb_fft_plan_destroy(plan);
    }
    void require_forward() const {
        if (policy == VerticalGemmBufferPolicy::inverseOnly)
            throw std::logic_error("Forward transform requested from an inverse-only WVM field-view provider.");
    }
    void require_inverse() const {
        if (policy == VerticalGemmBufferPolicy::forwardOnly)
            throw std::logic_error("Inverse transform requested from a forward-only WVM field-view provider.");
    }
    void require_stride(std::size_t s) const {
        if (!std::binary_search(strides.begin(), strides.end(), s))
            throw std::invalid_argument("WVM field view uses a frequency stride that was not planned.");
    }
    template <typename V>
    void validate_view(const V& v) const {
        if (v.data == nullptr) throw std::invalid_argument("WVM FFTW field view is null.");
        require_stride(v.frequencyStride);
    }
    std::size_t frequencies() const { return w.nxHalf() * w.ny; }
};

HipInterleavedWvmFieldViewProvider::HipInterleavedWvmFieldViewProvider(const Workload& singleFieldWorkload,
                                                                       std::vector<std::size_t> frequencyStrides,
                                                                       VerticalGemmBufferPolicy bufferPolicy,
                                                                       Precision precision)
    : impl_(std::make_unique<Impl>(singleFieldWorkload, std::move(frequencyStrides), bufferPolicy, precision)) {}
HipInterleavedWvmFieldViewProvider::~HipInterleavedWvmFieldViewProvider() = default;
HipInterleavedWvmFieldViewProvider::HipInterleavedWvmFieldViewProvider(HipInterleavedWvmFieldViewProvider&&) noexcept =
    default;
HipInterleavedWvmFieldViewProvider& HipInterleavedWvmFieldViewProvider::operator=(
    HipInterleavedWvmFieldViewProvider&&) noexcept = default;

void HipInterleavedWvmFieldViewProvider::inverseFields(MutableFieldView* fields, std::size_t fieldCount,
                                                       double* output) {
    auto& m = *impl_;
    m.require_inverse();
    if (fields == nullptr || fieldCount == 0)
        throw std::invalid_argument("WVM field-view inverse requires at least one field.");
    for (std::size_t f = 0; f < fieldCount; ++f) m.validate_view(fields[f]);
    if (output == nullptr) throw std::invalid_argument("WVM field-view inverse output is null.");
    const std::size_t nz = m.w.nz, freqs = m.frequencies(), vol = m.w.realElements();
    for (std::size_t f = 0; f < fieldCount; ++f) {
        const Complex* v = fields[f].data;
        const std::size_t s = fields[f].frequencyStride;
        for (std::size_t q = 0; q < freqs; ++q)
            for (std::size_t z = 0; z < nz; ++z) {
                m.host[2 * (z + nz * q)] = v[z + s * q].real;
                m.host[2 * (z + nz * q) + 1] = v[z + s * q].imag;
            }
        m.spectrum.upload(m.host.data(), m.stream.s);
        check(skb_fft_inverse(m.plan, m.spectrum.dev, m.real.dev, m.stream.s));
        m.real.download(output + f * vol, m.stream.s);
    }
}

void HipInterleavedWvmFieldViewProvider::forwardField(const double* input, MutableFieldView output) {
    auto& m = *impl_;
    m.require_forward();
    m.validate_view(output);
    if (input == nullptr) throw std::invalid_argument("WVM field-view forward input is null.");
    m.real.upload(input, m.stream.s);
    check(skb_fft_forward(m.plan, m.real.dev, m.spectrum.dev, m.stream.s));
    m.spectrum.download(m.host.data(), m.stream.s);
    const std::size_t nz = m.w.nz, freqs = m.frequencies(), s = output.frequencyStride;
    for (std::size_t q = 0; q < freqs; ++q)
        for (std::size_t z = 0; z < nz; ++z)
            output.data[z + s * q] = {m.host[2 * (z + nz * q)], m.host[2 * (z + nz * q) + 1]};
}

void HipInterleavedWvmFieldViewProvider::inverseFieldsDevice(const DeviceFieldView* fields, std::size_t fieldCount,
                                                             void* output, void* stream) {
    auto& m = *impl_;
    m.require_inverse();
    if (fields == nullptr || fieldCount == 0)
        throw std::invalid_argument("WVM field-view inverse requires at least one field.");
    for (std::size_t f = 0; f < fieldCount; ++f) m.validate_view(fields[f]);
    if (output == nullptr) throw std::invalid_argument("WVM field-view inverse output is null.");
    const std::size_t bytes = m.w.realElements() * (m.real.fp32 ? 4 : 8);
    for (std::size_t f = 0; f < fieldCount; ++f)
        check(skb_fft_inverse_view(m.plan, fields[f].data, fields[f].frequencyStride,
                                   static_cast<char*>(output) + f * bytes, stream));
}

void HipInterleavedWvmFieldViewProvider::forwardFieldDevice(const void* input, DeviceFieldView output, void* stream) {
    auto& m = *impl_;
    m.require_forward();
    m.validate_view(output);
    if (input == nullptr) throw std::invalid_argument("WVM field-view forward input is null.");
    check(skb_fft_forward_view(m.plan, input, output.data, output.frequencyStride, stream));
}

void HipInterleavedWvmFieldViewProvider::executeSchedulerNoop() {
    hip_check(hipStreamSynchronize(impl_->stream.s), "sync");
}
double HipInterleavedWvmFieldViewProvider::allocationSeconds() const noexcept { return impl_->allocation_seconds; }
double HipInterleavedWvmFieldViewProvider::planningSeconds() const noexcept { return impl_->planning_seconds; }
std::size_t HipInterleavedWvmFieldViewProvider::planningBytes() const noexcept {
    return impl_->w.realElements() * 8 + impl_->strides.back() * impl_->frequencies() * 16;
}
std::string HipInterleavedWvmFieldViewProvider::libraryIdentity() const { return "libskb_hip.so"; }
std::string HipInterleavedWvmFieldViewProvider::version() const { return skb_version(); }

// ---- HipVerticalGemmProvider ----------------------------------------------------------
std::string verticalGemmLayoutName(VerticalGemmLayout layout) {
    return layout == VerticalGemmLayout::split ? "split" : "complex-interleaved";
}

// Host arrays of the layout (the reference's Accelerate buffers,
// src/vertical_gemm.cpp:360-390) plus device staging of the plan precision.
// Forward direction: physical input -> modal output; inverse: modal input ->
// physical output.  A direction the buffer policy leaves out has no arrays.
struct HipVerticalGemmProvider::Impl {
    Workload w;
    std::size_t nj = 0, modes = 0, groups = 0;
    VerticalGemmLayout layout;
    VerticalGemmStrategy strategy;
    VerticalGemmBufferPolicy policy;
    Precision precision;
    skb_vertical_plan* plan = nullptr;
    bool available = true;
    std::string capability_text;
    double allocation_seconds = 0, preparation_seconds = 0;
    Stream stream;
    std::size_t physical_count = 0, modal_count = 0;
    // interleaved layout
    std::vector<Complex> physical_in, modal_out, modal_in, physical_out;
    Staged d_physical, d_modal;
    // split layout
    std::vector<double> physical_in_re, physical_in_im, modal_out_re, modal_out_im;
    std::vector<double> modal_in_re, modal_in_im, physical_out_re, physical_out_im;
    Staged d_physical_re, d_physical_im, d_modal_re, d_modal_im;

    Impl(const Workload& wl, const GroupedVerticalOperators& ops, VerticalGemmLayout l, VerticalGemmStrategy st,
         VerticalGemmBufferPolicy bp, Precision p)
        : w(wl), nj(ops.nj), groups(ops.groups.size()), layout(l), strategy(st), policy(bp), precision(p) {
        if (ops.nz != wl.nz) throw std::invalid_argument("Vertical operator dimensions do not match the workload.");
        if (ops.groups.empty()) throw std::invalid_argument("Vertical GEMM requires at least one group.");
        if (strategy.outerWorkers == 0) throw std::invalid_argument("Vertical GEMM outer workers must be positive.");
        modes = ops.groups.back().firstMode + ops.groups.back().modeCount;
        physical_count = modes * wl.planes();
        modal_count = modes * wl.fields * nj;
        const auto cw = to_c(wl);
        const auto t0 = Clock::now();
        const int st_create =
            skb_vertical_plan_create(&cw, nj, reinterpret_cast<const skb_mode_group*>(ops.groups.data()), groups,
                                     ops.forward.data(), ops.inverse.data(), c_precision(p), &plan);
        preparation_seconds = seconds_since(t0);
        if (st_create == SKB_UNSUPPORTED) {
            available = false;
            capability_text = "unsupported: " + last_error();
            return;
        }
        check(st_create);
        const auto t1 = Clock::now();
        const bool single = p == Precision::float32;
        const bool fwd = policy != VerticalGemmBufferPolicy::inverseOnly;
        const bool inv = policy != VerticalGemmBufferPolicy::forwardOnly;
        if (layout == VerticalGemmLayout::complexInterleaved) {
            if (fwd) {
                physical_in.resize(physical_count);
                modal_out.resize(modal_count);
            }
            if (inv) {
                modal_in.resize(modal_count);
                physical_out.resize(physical_count);
            }
            d_physical.allocate(2 * physical_count, single);
            d_modal.allocate(2 * modal_count, single);
        } else {
            check(skb_vertical_plan_reserve_split(plan));
            if (fwd) {
                physical_in_re.resize(physical_count);
                physical_in_im.resize(physical_count);
                modal_out_re.resize(modal_count);
                modal_out_im.resize(modal_count);
            }
            if (inv) {
                modal_in_re.resize(modal_count);
                modal_in_im.resize(modal_count);
                physical_out_re.resize(physical_count);
                physical_out_im.resize(physical_count);
            }
            d_physical_re.allocate(physical_count, single);
            d_physical_im.allocate(physical_count, single);
            d_modal_re.allocate(modal_count, single);
            d_modal_im.allocate(modal_count, single);
        }
        allocation_seconds = seconds_since(t1);
        capability_text = "supported: hip grouped MFMA GEMM (gfx950), " + verticalGemmLayoutName(layout) +
                          " layout, one grouped launch per direction";
    }
    ~Impl() {
        if (plan) skb_vertical_plan_destroy(plan);
    }
    // src/vertical_gemm.cpp:433-467
    void require_available() const {
        if (!available) throw std::runtime_error(capability_text);
    }
    void require_forward() const {
        require_available();
        if (policy == VerticalGemmBufferPolicy::inverseOnly)
            throw std::logic_error("Forward GEMM component requested from an inverse-only provider.");
    }
    void require_inverse() const {
        require_available();
        if (policy == VerticalGemmBufferPolicy::forwardOnly)
            throw std::logic_error("Inverse GEMM component requested from a forward-only provider.");
    }
    void require_split() const {
        require_available();
        if (layout != VerticalGemmLayout::split)
            throw std::logic_error("Split GEMM component requested from the complex GEMM provider.");
    }
    void require_interleaved() const {
        require_available();
        if (layout != VerticalGemmLayout::complexInterleaved)
            throw std::logic_error("Interleaved GEMM component requested from the split GEMM provider.");
    }
    void validate_modes(const std::vector<RetainedMode>& m) const {
        if (m.size() != modes)
            throw std::invalid_argument("Retained-mode count does not match the vertical GEMM columns.");
    }
    // One direction of the split product over the host arrays; a null host
    // pair leaves that part out.
    void split_product(bool forward, const double* in_re, const double* in_im, double* out_re, double* out_im) {
        Staged& i_re = forward ? d_physical_re : d_modal_re;
        Staged& i_im = forward ? d_physical_im : d_modal_im;
        Staged& o_re = forward ? d_modal_re : d_physical_re;
        Staged& o_im = forward ? d_modal_im : d_physical_im;
        if (in_re) i_re.upload(in_re, stream.s);
        if (in_im) i_im.upload(in_im, stream.s);
        const void* dir = in_re ? i_re.dev : nullptr;
        const void* dii = in_im ? i_im.dev : nullptr;
        void* dor = in_re ? o_re.dev : nullptr;
        void* doi = in_im ? o_im.dev : nullptr;
        if (forward) check(skb_vertical_forward_split(plan, dir, dii, dor, doi, stream.s));
        else check(skb_vertical_inverse_split(plan, dir, dii, dor, doi, stream.s));
        if (out_re) o_re.download(out_re, stream.s);
        if (out_im) o_im.download(out_im, stream.s);
    }
};

HipVerticalGemmProvider::HipVerticalGemmProvider(const Workload& workload, const GroupedVerticalOperators& operators,
                                                 Precision precision)
    : impl_(std::make_unique<Impl>(workload, operators, VerticalGemmLayout::complexInterleaved,
                                   VerticalGemmStrategy{}, VerticalGemmBufferPolicy::bidirectional, precision)) {}
HipVerticalGemmProvider::HipVerticalGemmProvider(const Workload& workload, const GroupedVerticalOperators& operators,
                                                 VerticalGemmLayout layout, Precision precision)
    : impl_(std::make_unique<Impl>(workload, operators, layout, VerticalGemmStrategy{},
                                   VerticalGemmBufferPolicy::bidirectional, precision)) {}
HipVerticalGemmProvider::HipVerticalGemmProvider(const Workload& workload, const GroupedVerticalOperators& operators,
                                                 VerticalGemmLayout layout, VerticalGemmStrategy strategy,
                                                 VerticalGemmBufferPolicy bufferPolicy, Precision precision)
    : impl_(std::make_unique<Impl>(workload, operators, layout, strategy, bufferPolicy, precision)) {}
HipVerticalGemmProvider::~HipVerticalGemmProvider() = default;
HipVerticalGemmProvider::HipVerticalGemmProvider(HipVerticalGemmProvider&&) noexcept = default;
HipVerticalGemmProvider& HipVerticalGemmProvider::operator=(HipVerticalGemmProvider&&) noexcept = default;

bool HipVerticalGemmProvider::supported() const noexcept { return impl_->available; }
std::string HipVerticalGemmProvider::capability() const { return impl_->capability_text; }
VerticalGemmLayout HipVerticalGemmProvider::layout() const noexcept { return impl_->layout; }
VerticalGemmStrategy HipVerticalGemmProvider::strategy() const noexcept { return impl_->strategy; }
VerticalGemmBufferPolicy HipVerticalGemmProvider::bufferPolicy() const noexcept { return impl_->policy; }
std::size_t HipVerticalGemmProvider::outerWorkers() const noexcept { return impl_->strategy.outerWorkers; }
std::size_t HipVerticalGemmProvider::columns() const noexcept { return impl_->modes * impl_->w.fields; }
std::size_t HipVerticalGemmProvider::physicalElements() const noexcept { return impl_->physical_count; }
std::size_t HipVerticalGemmProvider::modalElements() const noexcept { return impl_->modal_count; }
std::size_t HipVerticalGemmProvider::groupCount() const noexcept { return impl_->groups; }
std::size_t HipVerticalGemmProvider::gemmCallsPerExecution() const noexcept { return 1; }
std::size_t HipVerticalGemmProvider::matrixBytesPerDirection() const noexcept {
    return impl_->plan ? skb_vertical_matrix_bytes(impl_->plan) : 0;
}
std::size_t HipVerticalGemmProvider::persistentBytes() const noexcept {
    const auto& m = *impl_;
    const std::size_t e = m.precision == Precision::float32 ? 4 : 8;
    const std::size_t host = 16 * (m.physical_in.size() + m.modal_out.size() + m.modal_in.size() +
                                   m.physical_out.size()) +
                             8 * (m.physical_in_re.size() + m.physical_in_im.size() + m.modal_out_re.size() +
                                  m.modal_out_im.size() + m.modal_in_re.size() + m.modal_in_im.size() +
                                  m.physical_out_re.size() + m.physical_out_im.size());
    const std::size_t device = m.available ? 2 * e * (m.physical_count + m.modal_count) * (m.layout ==
                                   VerticalGemmLayout::split ? 2 : 1) : 0;
    return host + device + 2 * matrixBytesPerDirection();
}
double HipVerticalGemmProvider::allocationSeconds() const noexcept { return impl_->allocation_seconds; }
double HipVerticalGemmProvider::matrixPreparationSeconds() const noexcept { return impl_->preparation_seconds; }

// src/vertical_gemm.cpp:698-729
void HipVerticalGemmProvider::loadPhysicalInput(const Complex* input) {
    auto& m = *impl_;
    m.require_forward();
    if (m.layout == VerticalGemmLayout::complexInterleaved) {
        std::copy_n(input, m.physical_count, m.physical_in.data());
        return;
    }
    for (std::size_t i = 0; i < m.physical_count; ++i) {
        m.physical_in_re[i] = input[i].real;
        m.physical_in_im[i] = input[i].imag;
    }
}
void HipVerticalGemmProvider::loadModalInput(const Complex* input) {
    auto& m = *impl_;
    m.require_inverse();
    if (m.layout == VerticalGemmLayout::complexInterleaved) {
        std::copy_n(input, m.modal_count, m.modal_in.data());
        return;
    }
    for (std::size_t i = 0; i < m.modal_count; ++i) {
        m.modal_in_re[i] = input[i].real;
        m.modal_in_im[i] = input[i].imag;
    }
}

// WVM spectrum (z + Nz*field + P*(kx + (Nx/2+1)*ky)) -> retained physical
// input (z + Nz*field + P*mode), conjugating modes stored through their
// Hermitian partner (src/vertical_gemm.cpp:470-490).
void HipVerticalGemmProvider::packPhysicalInputFromWvm(const std::vector<RetainedMode>& modes,
                                                       const Complex* wvmSpectrum) {
    auto& m = *impl_;
    m.require_forward();
    m.validate_modes(modes);
    const std::size_t P = m.w.planes(), half = m.w.nxHalf();
    for (std::size_t mode = 0; mode < modes.size(); ++mode) {
        const auto& r = modes[mode];
        const Complex* src = wvmSpectrum + P * (r.storedKx + half * r.storedKy);
        const double sign = r.conjugatesStoredValue ? -1.0 : 1.0;
        for (std::size_t plane = 0; plane < P; ++plane) {
            const std::size_t dst = plane + P * mode;
            if (m.layout == VerticalGemmLayout::complexInterleaved) {
                m.physical_in[dst] = {src[plane].real, sign * src[plane].imag};
            } else {
                m.physical_in_re[dst] = src[plane].real;
                m.physical_in_im[dst] = sign * src[plane].imag;
            }
        }
    }
}

// Retained physical output -> zero-filled WVM spectrum; a k_x = 0 mode also
// writes its conjugate at (0, Ny - k_y) unless k_y is 0 or Ny/2
// (src/vertical_gemm.cpp:492-516).
void HipVerticalGemmProvider::embedPhysicalOutputToWvm(const std::vector<RetainedMode>& modes,
                                                       Complex* wvmSpectrum) const {
    const auto& m = *impl_;
    m.require_inverse();
    m.validate_modes(modes);
    std::fill_n(wvmSpectrum, m.w.spectrumElements(), Complex{});
    const std::size_t P = m.w.planes(), half = m.w.nxHalf(), ny = m.w.ny;
    for (std::size_t mode = 0; mode < modes.size(); ++mode) {
        const auto& r = modes[mode];
        Complex* dst = wvmSpectrum + P * (r.storedKx + half * r.storedKy);
        const bool partner = r.storedKx == 0 && r.storedKy != 0 && 2 * r.storedKy != ny;
        Complex* mirror = partner ? wvmSpectrum + P * (half * ((ny - r.storedKy) % ny)) : nullptr;
        for (std::size_t plane = 0; plane < P; ++plane) {
            const std::size_t src = plane + P * mode;
            Complex v = m.layout == VerticalGemmLayout::complexInterleaved
                            ? m.physical_out[src]
                            : Complex{m.physical_out_re[src], m.physical_out_im[src]};
            if (r.conjugatesStoredValue) v.imag = -v.imag;
            dst[plane] = v;
            if (mirror) mirror[plane] = {v.real, -v.imag};
        }
    }
}

void HipVerticalGemmProvider::executeForward() {
    auto& m = *impl_;
    m.require_forward();
    if (m.layout == VerticalGemmLayout::complexInterleaved) {
        m.d_physical.upload(reinterpret_cast<const double*>(m.physical_in.data()), m.stream.s);
        check(skb_vertical_forward(m.plan, m.d_physical.dev, m.d_modal.dev, m.stream.s));
        m.d_modal.download(reinterpret_cast<double*>(m.modal_out.data()), m.stream.s);
        return;
    }
    m.split_product(true, m.physical_in_re.data(), m.physical_in_im.data(), m.modal_out_re.data(),
                    m.modal_out_im.data());
}
void HipVerticalGemmProvider::executeInverse() {
    auto& m = *impl_;
    m.require_inverse();
    if (m.layout == VerticalGemmLayout::complexInterleaved) {
        m.d_modal.upload(reinterpret_cast<const double*>(m.modal_in.data()), m.stream.s);
        check(skb_vertical_inverse(m.plan, m.d_modal.dev, m.d_physical.dev, m.stream.s));
        m.d_physical.download(reinterpret_cast<double*>(m.physical_out.data()), m.stream.s);
        return;
    }
    m.split_product(false, m.modal_in_re.data(), m.modal_in_im.data(), m.physical_out_re.data(),
                    m.physical_out_im.data());
}
void HipVerticalGemmProvider::executeForwardReal() {
    auto& m = *impl_;
    m.require_forward();
    m.require_split();
    m.split_product(true, m.physical_in_re.data(), nullptr, m.modal_out_re.data(), nullptr);
}
void HipVerticalGemmProvider::executeForwardImaginary() {
    auto& m = *impl_;
    m.require_forward();
    m.require_split();
    m.split_product(true, nullptr, m.physical_in_im.data(), nullptr, m.modal_out_im.data());
}
void HipVerticalGemmProvider::executeInverseReal() {
    auto& m = *impl_;
    m.require_inverse();
    m.require_split();
    m.split_product(false, m.modal_in_re.data(), nullptr, m.physical_out_re.data(), nullptr);
}
void HipVerticalGemmProvider::executeInverseImaginary() {
    auto& m = *impl_;
    m.require_inverse();
    m.require_split();
    m.split_product(false, nullptr, m.modal_in_im.data(), nullptr, m.physical_out_im.data());
}
void HipVerticalGemmProvider::executeSchedulerNoop() {
    impl_->require_available();
    hip_check(hipStreamSynchronize(impl_->stream.s), "sync");
}

// src/vertical_gemm.cpp:826-900
Complex* HipVerticalGemmProvider::interleavedPhysicalInputData() {
    impl_->require_forward();
    impl_->require_interleaved();
    return impl_->physical_in.data();
}
Complex* HipVerticalGemmProvider::interleavedModalInputData() {
    impl_->require_inverse();
    impl_->require_interleaved();
    return impl_->modal_in.data();
}
const Complex* HipVerticalGemmProvider::interleavedModalOutputData() const {
    impl_->require_forward();
    impl_->require_interleaved();
    return impl_->modal_out.data();
}
const Complex* HipVerticalGemmProvider::interleavedPhysicalOutputData() const {
    impl_->require_inverse();
    impl_->require_interleaved();
    return impl_->physical_out.data();
}
double* HipVerticalGemmProvider::splitPhysicalInputRealData() {
    impl_->require_forward();
    impl_->require_split();
    return impl_->physical_in_re.data();
}
double* HipVerticalGemmProvider::splitPhysicalInputImaginaryData() {
    impl_->require_forward();
    impl_->require_split();
    return impl_->physical_in_im.data();
}
double* HipVerticalGemmProvider::splitModalInputRealData() {
    impl_->require_inverse();
    impl_->require_split();
    return impl_->modal_in_re.data();
}
double* HipVerticalGemmProvider::splitModalInputImaginaryData() {
    impl_->require_inverse();
    impl_->require_split();
    return impl_->modal_in_im.data();
}
const double* HipVerticalGemmProvider::splitModalOutputRealData() const {
    impl_->require_forward();
    impl_->require_split();
    return impl_->modal_out_re.data();
}
const double* HipVerticalGemmProvider::splitModalOutputImaginaryData() const {
    impl_->require_forward();
    impl_->require_split();
    return impl_->modal_out_im.data();
}
const double* HipVerticalGemmProvider::splitPhysicalOutputRealData() const {
    impl_->require_inverse();
    impl_->require_split();
    return impl_->physical_out_re.data();
}
const double* HipVerticalGemmProvider::splitPhysicalOutputImaginaryData() const {
    impl_->require_inverse();
    impl_->require_split();
    return impl_->physical_out_im.data();
}

void HipVerticalGemmProvider::copyForwardOutput(Complex* output) const {
    const auto& m = *impl_;
    m.require_forward();
    if (m.layout == VerticalGemmLayout::complexInterleaved) {
        std::copy_n(m.modal_out.data(), m.modal_count, output);
        return;
    }
    for (std::size_t i = 0; i < m.modal_count; ++i) output[i] = {m.modal_out_re[i], m.modal_out_im[i]};
}
void HipVerticalGemmProvider::copyInverseOutput(Complex* output) const {
    const auto& m = *impl_;
    m.require_inverse();
    if (m.layout == VerticalGemmLayout::complexInterleaved) {
        std::copy_n(m.physical_out.data(), m.physical_count, output);
        return;
    }
    for (std::size_t i = 0; i < m.physical_count; ++i) output[i] = {m.physical_out_re[i], m.physical_out_im[i]};
}

void HipVerticalGemmProvider::forwardDevice(const void* physical, void* modal, void* stream) {
    impl_->require_forward();
    check(skb_vertical_forward(impl_->plan, physical, modal, stream));
}
void HipVerticalGemmProvider::inverseDevice(const void* modal, void* physical, void* stream) {
    impl_->require_inverse();
    check(skb_vertical_inverse(impl_->plan, modal, physical, stream));
}
void HipVerticalGemmProvider::forwardSplitDevice(const void* physicalReal, const void* physicalImag,
                                                 void* modalReal, void* modalImag, void* stream) {
    impl_->require_forward();
    impl_->require_split();
    check(skb_vertical_forward_split(impl_->plan, physicalReal, physicalImag, modalReal, modalImag, stream));
}
void HipVerticalGemmProvider::inverseSplitDevice(const void* modalReal, const void* modalImag, void* physicalReal,
                                                 void* physicalImag, void* stream) {
    impl_->require_inverse();
    impl_->require_split();
    check(skb_vertical_inverse_split(impl_->plan, modalReal, modalImag, physicalReal, physicalImag, stream));
}
std::string HipVerticalGemmProvider::libraryIdentity() const { return "libskb_hip.so"; }

// ---- HipPointwiseAdvectionExecutor ------------------------------------------------------
struct HipPointwiseAdvectionExecutor::Impl {
    std::size_t volume;
    double scale;
    Precision precision;
    Stream stream;
    Staged shared, derivative, target;

    Impl(std::size_t v, double s, Precision p) : volume(v), scale(s), precision(p) {
        if (v == 0) throw std::invalid_argument("Pointwise advection volume must be positive.");
        const bool single = p == Precision::float32;
        shared.allocate(3 * v, single);
        derivative.allocate(3 * v, single);
        target.allocate(v, single);
    }
};

HipPointwiseAdvectionExecutor::HipPointwiseAdvectionExecutor(std::size_t volumeElements, double scale,
                                                             Precision precision)
    : impl_(std::make_unique<Impl>(volumeElements, scale, precision)) {}
HipPointwiseAdvectionExecutor::~HipPointwiseAdvectionExecutor() = default;
HipPointwiseAdvectionExecutor::HipPointwiseAdvectionExecutor(HipPointwiseAdvectionExecutor&&) noexcept = default;
HipPointwiseAdvectionExecutor& HipPointwiseAdvectionExecutor::operator=(HipPointwiseAdvectionExecutor&&) noexcept =
    default;

void HipPointwiseAdvectionExecutor::execute(const double* shared, const double* derivative, double* target) {
    auto& m = *impl_;
    m.shared.upload(shared, m.stream.s);
    m.derivative.upload(derivative, m.stream.s);
    check(skb_pointwise_advection(m.volume, m.scale, m.shared.dev, m.derivative.dev, m.target.dev,
                                  c_precision(m.precision), m.stream.s));
    m.target.download(target, m.stream.s);
}

void HipPointwiseAdvectionExecutor::executeDevice(const void* shared, const void* derivative, void* target,
                                                  void* stream) {
    auto& m = *impl_;
    check(skb_pointwise_advection(m.volume, m.scale, shared, derivative, target, c_precision(m.precision), stream));
}

void HipPointwiseAdvectionExecutor::executeSchedulerNoop() {
    hip_check(hipStreamSynchronize(impl_->stream.s), "sync");
}
std::size_t HipPointwiseAdvectionExecutor::persistentBytes() const noexcept {
    const std::size_t e = impl_->precision == Precision::float32 ? 4 : 8;
    return 7 * impl_->volume * e;
}

}  // namespace skbhip
