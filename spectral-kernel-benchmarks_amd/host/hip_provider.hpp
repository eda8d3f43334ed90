// C++20 host mirror of skbench's provider classes, backed by the gfx950 HIP
// kernels of libskb_hip.so through the C ABI in include/skb_hip.h.
//
// Every class keeps the reference's method names, argument meaning, buffer
// ownership and error behaviour:
//   HipProvider                  <- skbench::FFTWProvider          (include/skbench/skbench.hpp:364-421)
//   HipPrunedProvider            <- skbench::FFTWPrunedProvider    (include/skbench/skbench.hpp:467-520)
//   HipVerticalGemmProvider      <- skbench::VerticalGemmProvider  (include/skbench/skbench.hpp:152-218)
//   HipPointwiseAdvectionExecutor<- skbench::PointwiseAdvectionExecutor (src/pointwise_advection.hpp:18-43)
//   HipWvmDirectVerticalGemmProvider <- skbench::WvmDirectVerticalGemmProvider
//                                      (include/skbench/skbench.hpp:220-257)
//   HipStreamingPrunedSplitProvider <- skbench::FFTWStreamingPrunedSplitProvider
//                                      (include/skbench/skbench.hpp:534-621)
//   HipInterleavedWvmFieldViewProvider <- skbench::FFTWInterleavedWvmFieldViewProvider
//                                      (include/skbench/skbench.hpp:423-465)
// Host-pointer methods are the reference's "adapter" scope: they stage the
// caller's FP64 arrays through device buffers owned by the provider (allocated
// at construction; nothing is allocated per call) and return only after the
// provider's stream has finished the work, so host results are complete and
// a host timer around a call measures the call (the reference's CPU calls are
// synchronous).  The *Device methods are the "primitive" scope on
// caller-owned device memory: they enqueue on the caller's stream and return
// at once.
//
// The value types below are layout-identical to the reference's
// (skbench::Complex, Workload, RetainedMode, VerticalModeGroup); the static
// asserts in hip_provider.cpp pin that, so a skbench build can pass its own
// objects through reinterpret_cast exactly as it does with fftw_complex.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "skb_hip.h"

namespace skbhip {

struct Complex {
    double real = 0.0;
    double imag = 0.0;
};

struct Workload {
    std::size_t nx = 0, ny = 0, nz = 0, fields = 0;
    double lx = 1.0, ly = 1.0;
    bool antialias = true;

    std::size_t planes() const { return nz * fields; }
    std::size_t nxHalf() const { return nx / 2 + 1; }
    std::size_t realElements() const { return nx * ny * planes(); }
    std::size_t spectrumElements() const { return nxHalf() * ny * planes(); }
    std::size_t retainedVerticalModes() const { return nz < 2 ? 0 : 2 * (nz - 1) / 3; }
};

struct RetainedMode {
    std::int64_t k = 0, l = 0;
    std::size_t storedKx = 0, storedKy = 0;
    bool conjugatesStoredValue = false;
    double radialMode = 0.0;
};

struct VerticalModeGroup {
    std::uint64_t squaredModeKey = 0;
    std::size_t firstMode = 0;
    std::size_t modeCount = 0;
    bool operator==(const VerticalModeGroup&) const = default;
};

struct GroupedVerticalOperators {
    std::string id;
    std::size_t nz = 0, nj = 0;
    std::vector<VerticalModeGroup> groups;
    std::vector<double> forward;  // groups x (nj x nz)
    std::vector<double> inverse;  // groups x (nz x nj)
};

enum class Precision { float64, float32 };

enum class FixtureKind { impulse, sinusoid, random, dc, nyquist, sinusoid5x17 };

// Data layout of the FFT provider's spectra (FFTWDataLayout,
// include/skbench/skbench.hpp:337-340): interleaved Complex arrays, or
// separate real and imaginary arrays in the same WVM element order.
enum class DataLayout { interleaved, split };

// Vertical GEMM options (include/skbench/skbench.hpp:125-151).  The device
// product is one grouped launch, so the schedule and worker count are kept
// for API parity and reported back; the buffer policy decides which
// direction's staging exists, as in the reference.
enum class VerticalGemmLayout { complexInterleaved, split };
std::string verticalGemmLayoutName(VerticalGemmLayout layout);
enum class VerticalGemmBufferPolicy { bidirectional, forwardOnly, inverseOnly };
enum class VerticalGemmSchedule { serial, outerStatic, outerDynamic };
struct VerticalGemmStrategy {
    VerticalGemmSchedule schedule = VerticalGemmSchedule::serial;
    std::size_t outerWorkers = 1;
};

// ---- model (src/model.cpp) ------------------------------------------------
std::vector<RetainedMode> retainedHorizontalModes(const Workload& workload);
std::vector<VerticalModeGroup> squaredWavenumberGroups(const std::vector<RetainedMode>& modes);
std::string modeOrderHash(const std::vector<RetainedMode>& modes);
std::string verticalModeGroupHash(const std::vector<VerticalModeGroup>& groups);
std::string wvmSpectrumOrderHash(const Workload& workload);
GroupedVerticalOperators squaredWavenumberVerticalFixture(const Workload& workload,
                                                          const std::vector<RetainedMode>& modes);
// 32 contiguous K^2 bins of 65x65-style orthogonal matrices (no reference counterpart).
GroupedVerticalOperators binnedQrVerticalFixture(const Workload& workload, const std::vector<RetainedMode>& modes,
                                                 std::size_t bins, std::uint64_t seed);
std::vector<double> makeFixture(const Workload& workload, FixtureKind fixture, std::uint64_t seed);

// ---- providers --------------------------------------------------------------
class HipProvider {
public:
    HipProvider(const Workload& workload, Precision precision = Precision::float64);
    HipProvider(const Workload& workload, DataLayout layout, Precision precision = Precision::float64);
    ~HipProvider();
    HipProvider(HipProvider&&) noexcept;
    HipProvider& operator=(HipProvider&&) noexcept;
    HipProvider(const HipProvider&) = delete;
    HipProvider& operator=(const HipProvider&) = delete;

    // adapter scope (host FP64 arrays, WVM order)
    void forward(const double* input, Complex* wvmSpectrum);
    void inverse(Complex* wvmSpectrum, double* output);  // may treat wvmSpectrum as scratch
    void gatherRetainedOuter(const std::vector<RetainedMode>& modes, const Complex* wvmSpectrum,
                             Complex* retainedSpectrum);
    void embedRetainedOuter(const std::vector<RetainedMode>& modes, const Complex* retainedSpectrum,
                            Complex* wvmSpectrum);
    // split forms (include/skbench/skbench.hpp:380-398, src/fftw_provider.cpp:331-395, 701-830):
    // retained split arrays hold (plane, mode) at plane + P * mode; split
    // spectra hold the WVM order of the interleaved spectrum.  forwardSplit,
    // inverseSplit, gatherRetainedSplitOuter and embedRetainedSplitOuter need
    // the split layout; forward and inverse need the interleaved one.
    void gatherRetainedToSplitOuter(const std::vector<RetainedMode>& modes, const Complex* spectrum,
                                    double* retainedReal, double* retainedImag, double scale = 1.0);
    void embedRetainedFromSplitOuter(const std::vector<RetainedMode>& modes, const double* retainedReal,
                                     const double* retainedImag, Complex* spectrum);
    void scaleRetainedSplitOuter(const std::vector<RetainedMode>& modes, double* retainedReal, double* retainedImag,
                                 double scale);
    void forwardSplit(const double* input, double* wvmSpectrumReal, double* wvmSpectrumImag);
    void inverseSplit(double* wvmSpectrumReal, double* wvmSpectrumImag, double* output);
    void gatherRetainedSplitOuter(const std::vector<RetainedMode>& modes, const double* spectrumReal,
                                  const double* spectrumImag, double* retainedReal, double* retainedImag);
    void embedRetainedSplitOuter(const std::vector<RetainedMode>& modes, const double* retainedReal,
                                 const double* retainedImag, double* spectrumReal, double* spectrumImag);
    // primitive scope (device arrays in the plan's precision)
    void forwardDevice(const void* real, void* wvm, void* stream);
    void inverseDevice(void* wvm, void* real, void* stream);

    void executeSchedulerNoop();
    DataLayout layout() const noexcept;
    bool splitInPlaceWvmOrderSupported() const noexcept { return false; }
    std::string splitInPlaceWvmOrderCapability() const;
    bool supported() const noexcept;
    std::string capability() const;
    double allocationSeconds() const noexcept;
    double planningSeconds() const noexcept;
    std::size_t planningBytes() const noexcept;
    std::size_t minimumAlignmentBytes() const noexcept;
    std::string libraryIdentity() const;
    std::string version() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

class HipPrunedProvider {
public:
    HipPrunedProvider(const Workload& workload, const std::vector<RetainedMode>& modes,
                      Precision precision = Precision::float64);
    ~HipPrunedProvider();
    HipPrunedProvider(HipPrunedProvider&&) noexcept;
    HipPrunedProvider& operator=(HipPrunedProvider&&) noexcept;
    HipPrunedProvider(const HipPrunedProvider&) = delete;
    HipPrunedProvider& operator=(const HipPrunedProvider&) = delete;

    void forward(const double* input, Complex* retainedSpectrum);
    void inverse(const Complex* retainedSpectrum, double* output);
    void executeForwardRows(const double* input);
    void executeForwardColumns();  // fused with gatherForward on the device
    void gatherForward(Complex* retainedSpectrum);
    void embedInverse(const Complex* retainedSpectrum);
    void executeInverseColumns();
    void executeInverseRows(double* output);
    // split compact form (include/skbench/skbench.hpp:481-492)
    void gatherForwardSplit(double* retainedReal, double* retainedImag, double scale = 1.0);
    void forwardSplit(const double* input, double* retainedReal, double* retainedImag, double scale = 1.0);
    void embedInverseSplit(const double* retainedReal, const double* retainedImag);
    void inverseSplit(const double* retainedReal, const double* retainedImag, double* output);
    void forwardDevice(const void* real, void* retained, void* stream);
    void inverseDevice(const void* retained, void* real, void* stream);
    void forwardSplitDevice(const void* real, void* retainedReal, void* retainedImag, double scale, void* stream);
    void inverseSplitDevice(const void* retainedReal, const void* retainedImag, void* real, void* stream);
    void executeSchedulerNoop();

    std::size_t activeKxCount() const noexcept;
    std::size_t fullKxCount() const noexcept;
    std::size_t rowTransformsPerDirection() const noexcept;
    std::size_t columnTransformsPerDirection() const noexcept;
    std::size_t omittedColumnTransformsPerDirection() const noexcept;
    std::size_t scratchBytes() const noexcept;
    bool completeHalfSpectrumOutputMaterialized() const noexcept { return false; }
    std::string libraryIdentity() const;
    std::string version() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// The reference's CPU scratch-clearing strategies for the streaming inverse
// (include/skbench/skbench.hpp:522-527).  On the GPU the inverse column pass
// reads retained values through a (k_x, k_y) -> mode map and never clears a
// half-spectrum, so every policy runs the same kernels; the policy is kept
// for API parity and reported back.
enum class StreamingInversePreparationPolicy { fullZero, activeReset, compactPreserved, fullPreserved };

class HipStreamingPrunedSplitProvider {
public:
    // Layout-identical to skb_field_view (static_assert in hip_provider.cpp).
    struct ConstFieldView {
        const double* real = nullptr;
        const double* imaginary = nullptr;
        std::size_t modeStride = 0;
    };
    struct FieldView {
        double* real = nullptr;
        double* imaginary = nullptr;
        std::size_t modeStride = 0;
    };

    HipStreamingPrunedSplitProvider(const Workload& workload, const std::vector<RetainedMode>& modes,
                                    Precision precision = Precision::float64, std::size_t tileWidth = 1,
                                    StreamingInversePreparationPolicy inversePreparationPolicy =
                                        StreamingInversePreparationPolicy::fullZero);
    ~HipStreamingPrunedSplitProvider();
    HipStreamingPrunedSplitProvider(HipStreamingPrunedSplitProvider&&) noexcept;
    HipStreamingPrunedSplitProvider& operator=(HipStreamingPrunedSplitProvider&&) noexcept;
    HipStreamingPrunedSplitProvider(const HipStreamingPrunedSplitProvider&) = delete;
    HipStreamingPrunedSplitProvider& operator=(const HipStreamingPrunedSplitProvider&) = delete;

    // adapter scope (host FP64 arrays)
    void forwardSplit(const double* input, double* retainedReal, double* retainedImag, double scale = 1.0);
    void inverseSplit(const double* retainedReal, const double* retainedImag, double* output);
    void forwardSplitFields(const double* input, const FieldView* fields, std::size_t fieldCount,
                            double scale = 1.0);
    void inverseSplitFields(const ConstFieldView* fields, std::size_t fieldCount, double* output);
    // primitive scope (device arrays in the plan's precision)
    void forwardSplitDevice(const void* real, void* retainedReal, void* retainedImag, double scale, void* stream);
    void inverseSplitDevice(const void* retainedReal, const void* retainedImag, void* real, void* stream);
    void forwardSplitFieldsDevice(const void* real, const skb_field_view* fields, std::size_t fieldCount,
                                  double scale, void* stream);
    void inverseSplitFieldsDevice(const skb_field_view* fields, std::size_t fieldCount, void* real, void* stream);
    void executeSchedulerNoop();

    std::size_t activeKxCount() const noexcept;
    std::size_t fullKxCount() const noexcept;
    std::size_t rowTransformsPerDirection() const noexcept;
    std::size_t columnTransformsPerDirection() const noexcept;
    std::size_t omittedColumnTransformsPerDirection() const noexcept;
    std::size_t scratchBytes() const noexcept;
    std::size_t tileWidth() const noexcept;
    StreamingInversePreparationPolicy inversePreparationPolicy() const noexcept;
    bool rowInversePreservesInput() const noexcept { return true; }
    bool completeHalfSpectrumMaterialized() const noexcept { return false; }
    std::string libraryIdentity() const;
    std::string version() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// Full transforms of single fields that live inside interleaved WVM arrays
// with their native frequency stride (include/skbench/skbench.hpp:423-465,
// src/fftw_provider.cpp:919-1260).  A view holds element (kx, ky, z) at
// data[z + frequencyStride * (kx + (Nx/2+1) * ky)]; the provider is built for
// a one-field workload and the list of strides its views may use.
class HipInterleavedWvmFieldViewProvider {
public:
    struct MutableFieldView {
        Complex* data = nullptr;
        std::size_t frequencyStride = 0;
    };
    struct DeviceFieldView {
        void* data = nullptr;  // device array of the plan's precision
        std::size_t frequencyStride = 0;
    };

    HipInterleavedWvmFieldViewProvider(const Workload& singleFieldWorkload, std::vector<std::size_t> frequencyStrides,
                                       VerticalGemmBufferPolicy bufferPolicy = VerticalGemmBufferPolicy::bidirectional,
                                       Precision precision = Precision::float64);
    ~HipInterleavedWvmFieldViewProvider();
    HipInterleavedWvmFieldViewProvider(HipInterleavedWvmFieldViewProvider&&) noexcept;
    HipInterleavedWvmFieldViewProvider& operator=(HipInterleavedWvmFieldViewProvider&&) noexcept;
    HipInterleavedWvmFieldViewProvider(const HipInterleavedWvmFieldViewProvider&) = delete;
    HipInterleavedWvmFieldViewProvider& operator=(const HipInterleavedWvmFieldViewProvider&) = delete;

    // adapter scope (host arrays): the output holds field-major real volumes
    // in the order supplied; the inverse may overwrite each view
    void inverseFields(MutableFieldView* fields, std::size_t fieldCount, double* output);
    void forwardField(const double* input, MutableFieldView output);
    // primitive scope (device views and volumes, caller's stream)
    void inverseFieldsDevice(const DeviceFieldView* fields, std::size_t fieldCount, void* output, void* stream);
    void forwardFieldDevice(const void* input, DeviceFieldView output, void* stream);
    void executeSchedulerNoop();
    bool inverseInputIsDestructive() const noexcept { return true; }
    double otherSetupSeconds() const noexcept { return 0.0; }
    double allocationSeconds() const noexcept;
    double planningSeconds() const noexcept;
    std::size_t planningBytes() const noexcept;
    std::size_t internalWorkers() const noexcept { return 1; }
    std::size_t outerWorkers() const noexcept { return 1; }
    std::string libraryIdentity() const;
    std::string version() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

class HipVerticalGemmProvider {
public:
    HipVerticalGemmProvider(const Workload& workload, const GroupedVerticalOperators& operators,
                            Precision precision = Precision::float64);
    HipVerticalGemmProvider(const Workload& workload, const GroupedVerticalOperators& operators,
                            VerticalGemmLayout layout, Precision precision = Precision::float64);
    HipVerticalGemmProvider(const Workload& workload, const GroupedVerticalOperators& operators,
                            VerticalGemmLayout layout, VerticalGemmStrategy strategy,
                            VerticalGemmBufferPolicy bufferPolicy = VerticalGemmBufferPolicy::bidirectional,
                            Precision precision = Precision::float64);
    ~HipVerticalGemmProvider();
    HipVerticalGemmProvider(HipVerticalGemmProvider&&) noexcept;
    HipVerticalGemmProvider& operator=(HipVerticalGemmProvider&&) noexcept;
    HipVerticalGemmProvider(const HipVerticalGemmProvider&) = delete;
    HipVerticalGemmProvider& operator=(const HipVerticalGemmProvider&) = delete;

    // false (with the library's reason in capability()) when the GEMM is not
    // instantiated for these dimensions; every execution call then throws
    // std::runtime_error(capability()), as the reference's unavailable BLAS does.
    bool supported() const noexcept;
    std::string capability() const;
    VerticalGemmLayout layout() const noexcept;
    VerticalGemmStrategy strategy() const noexcept;
    VerticalGemmBufferPolicy bufferPolicy() const noexcept;
    std::size_t outerWorkers() const noexcept;
    std::size_t columns() const noexcept;
    std::size_t physicalElements() const noexcept;
    std::size_t modalElements() const noexcept;
    std::size_t groupCount() const noexcept;
    std::size_t gemmCallsPerExecution() const noexcept;  // one grouped launch
    std::size_t matrixBytesPerDirection() const noexcept;
    std::size_t persistentBytes() const noexcept;  // host staging + device staging + device matrices
    std::size_t schedulerPersistentBytes() const noexcept { return 0; }
    std::size_t minimumAlignmentBytes() const noexcept { return 256; }
    double allocationSeconds() const noexcept;
    double matrixPreparationSeconds() const noexcept;  // plan creation: matrix split and upload
    double schedulerSetupSeconds() const noexcept { return 0.0; }
    bool hasOpaqueSchedulerMemory() const noexcept { return false; }

    // adapter scope: the provider's host arrays (FP64, the layout's element
    // order: retained index z + Nz * field + P * mode, modal index
    // j + Nj * field + Nj * fields * mode); execute* uploads the input array,
    // runs the grouped product and downloads the output array before returning.
    void loadPhysicalInput(const Complex* input);
    void loadModalInput(const Complex* input);
    void packPhysicalInputFromWvm(const std::vector<RetainedMode>& modes, const Complex* wvmSpectrum);
    void executeForward();
    void executeInverse();
    void executeForwardReal();  // split layout: the real (imaginary) part of the product alone
    void executeForwardImaginary();
    void executeInverseReal();
    void executeInverseImaginary();
    void executeSchedulerNoop();
    Complex* interleavedPhysicalInputData();
    Complex* interleavedModalInputData();
    const Complex* interleavedModalOutputData() const;
    const Complex* interleavedPhysicalOutputData() const;
    double* splitPhysicalInputRealData();
    double* splitPhysicalInputImaginaryData();
    double* splitModalInputRealData();
    double* splitModalInputImaginaryData();
    const double* splitModalOutputRealData() const;
    const double* splitModalOutputImaginaryData() const;
    const double* splitPhysicalOutputRealData() const;
    const double* splitPhysicalOutputImaginaryData() const;
    void copyForwardOutput(Complex* output) const;
    void copyInverseOutput(Complex* output) const;
    void embedPhysicalOutputToWvm(const std::vector<RetainedMode>& modes, Complex* wvmSpectrum) const;

    // primitive scope (device arrays in the plan's precision, caller's stream)
    void forwardDevice(const void* physical, void* modal, void* stream);
    void inverseDevice(const void* modal, void* physical, void* stream);
    void forwardSplitDevice(const void* physicalReal, const void* physicalImag, void* modalReal, void* modalImag,
                            void* stream);
    void inverseSplitDevice(const void* modalReal, const void* modalImag, void* physicalReal, void* physicalImag,
                            void* stream);
    std::string libraryIdentity() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

class HipWvmDirectVerticalGemmProvider {
public:
    HipWvmDirectVerticalGemmProvider(const Workload& workload, const std::vector<RetainedMode>& modes,
                                     const GroupedVerticalOperators& operators,
                                     Precision precision = Precision::float64);
    ~HipWvmDirectVerticalGemmProvider();
    HipWvmDirectVerticalGemmProvider(HipWvmDirectVerticalGemmProvider&&) noexcept;
    HipWvmDirectVerticalGemmProvider& operator=(HipWvmDirectVerticalGemmProvider&&) noexcept;
    HipWvmDirectVerticalGemmProvider(const HipWvmDirectVerticalGemmProvider&) = delete;
    HipWvmDirectVerticalGemmProvider& operator=(const HipWvmDirectVerticalGemmProvider&) = delete;

    bool supported() const noexcept { return true; }
    std::string capability() const;
    // adapter scope (host FP64 full WVM arrays)
    void initializeModalOutput(Complex* fullModalSpectrum) const;
    void initializeSpectrumOutput(Complex* fullSpectrum) const;
    void executeForward(const Complex* fullSpectrum, Complex* fullModalSpectrum);
    void executeInverse(const Complex* fullModalSpectrum, Complex* fullSpectrum);
    // primitive scope (device full WVM arrays in the plan's precision)
    void initializeModalOutputDevice(void* fullModalSpectrum, void* stream) const;
    void initializeSpectrumOutputDevice(void* fullSpectrum, void* stream) const;
    void executeForwardDevice(const void* fullSpectrum, void* fullModalSpectrum, void* stream);
    void executeInverseDevice(const void* fullModalSpectrum, void* fullSpectrum, void* stream);
    void executeSchedulerNoop();
    std::size_t modalSpectrumElements() const noexcept;
    std::size_t gemmCallsPerExecution() const noexcept { return 1; }  // one grouped launch
    std::size_t matrixBytesPerDirection() const noexcept;
    std::string libraryIdentity() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

class HipPointwiseAdvectionExecutor {
public:
    HipPointwiseAdvectionExecutor(std::size_t volumeElements, double scale,
                                  Precision precision = Precision::float64);
    ~HipPointwiseAdvectionExecutor();
    HipPointwiseAdvectionExecutor(HipPointwiseAdvectionExecutor&&) noexcept;
    HipPointwiseAdvectionExecutor& operator=(HipPointwiseAdvectionExecutor&&) noexcept;
    HipPointwiseAdvectionExecutor(const HipPointwiseAdvectionExecutor&) = delete;
    HipPointwiseAdvectionExecutor& operator=(const HipPointwiseAdvectionExecutor&) = delete;

    void execute(const double* shared, const double* derivative, double* target);
    void executeDevice(const void* shared, const void* derivative, void* target, void* stream);
    void executeSchedulerNoop();
    std::size_t persistentBytes() const noexcept;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace skbhip
