// GPU tests of the C++ provider classes (skbench-style API), in the style of
// the reference's tests/skbench_tests.cpp: one main() of require() checks.
// The oracle (oracle/skb_oracle.c) is linked as the checker.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "hip_provider.hpp"

extern "C" {
typedef struct {
    double re, im;
} or_complex;
void or_direct_r2c(uint64_t, uint64_t, uint64_t, uint64_t, const double*, or_complex*);
void or_direct_c2r(uint64_t, uint64_t, uint64_t, uint64_t, const or_complex*, double*);
void or_direct_retained(uint64_t, uint64_t, uint64_t, uint64_t, const void*, uint64_t, const double*, or_complex*);
void or_embed_retained(uint64_t, uint64_t, uint64_t, uint64_t, const void*, uint64_t, const or_complex*, or_complex*);
void or_vertical_forward(uint64_t, uint64_t, uint64_t, uint64_t, const uint64_t*, const uint64_t*, const double*,
                         const or_complex*, or_complex*);
void or_gather_retained(uint64_t, uint64_t, uint64_t, uint64_t, const void*, uint64_t, const or_complex*,
                        or_complex*);
double or_max_relative_error_complex(const void*, const void*, uint64_t);
double or_max_relative_error_real(const double*, const double*, uint64_t, double);
}

namespace {

int failures = 0;

void require(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what.c_str());
        ++failures;
    }
}

// or_mode records are layout-identical to skb_retained_mode
struct OrMode {
    int64_t k, l;
    uint64_t skx, sky;
    int32_t conj, pad;
    double radial;
};

bool same_values(const std::vector<skbhip::Complex>& a, const std::vector<skbhip::Complex>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].real != b[i].real || a[i].imag != b[i].imag) return false;
    return true;
}

std::vector<OrMode> to_or(const std::vector<skbhip::RetainedMode>& modes) {
    std::vector<OrMode> out(modes.size());
    for (size_t i = 0; i < modes.size(); ++i)
        out[i] = {modes[i].k, modes[i].l, modes[i].storedKx, modes[i].storedKy, modes[i].conjugatesStoredValue ? 1 : 0,
                  0, modes[i].radialMode};
    return out;
}

}  // namespace

int main() {
    try {
        using namespace skbhip;
        const Workload w{8, 8, 7, 2, 1.0, 1.0, true};
        require(w.planes() == 14 && w.nxHalf() == 5 && w.retainedVerticalModes() == 4, "smoke geometry");
        const auto modes = retainedHorizontalModes(w);
        require(modes.front().k == 0 && modes.front().l == 0, "DC is first");
        const auto groups = squaredWavenumberGroups(modes);
        require(groups.size() == 5, "smoke K-squared group count");
        require(verticalModeGroupHash(groups) == "fnv1a64:69843c1f93cead00", "K-squared group hash");
        const Workload quick{256, 256, 65, 3, 1.0, 1.0, true};
        const auto quickModes = retainedHorizontalModes(quick);
        require(quickModes.size() == 11439, "256^2 retained-mode count");
        require(modeOrderHash(quickModes) == "fnv1a64:22e98f1dcda40ad5", "256^2 mode order hash");
        require(wvmSpectrumOrderHash(quick) == "fnv1a64:b1cdd8f25ebcbb72", "256^2 WVM order hash");

        const auto om = to_or(modes);
        HipProvider full(w);
        require(full.supported(), "hip provider supported at 8x8");
        HipPrunedProvider pruned(w, modes);
        require(pruned.activeKxCount() < pruned.fullKxCount(), "pruned provider omits kx columns");
        require(pruned.columnTransformsPerDirection() + pruned.omittedColumnTransformsPerDirection() ==
                    w.nxHalf() * w.planes(),
                "pruned column accounting");
        std::vector<Complex> spec(w.spectrumElements()), oracle(w.spectrumElements()), embedded(w.spectrumElements());
        std::vector<Complex> retained(modes.size() * w.planes()), retainedOracle(retained.size());
        std::vector<double> real(w.realElements()), realOracle(w.realElements());
        for (int f = 0; f < 5; ++f) {
            const auto kind = static_cast<FixtureKind>(f);
            const auto input = makeFixture(w, kind, 700 + f);
            const auto preserved = input;
            or_direct_r2c(8, 8, 7, 2, input.data(), reinterpret_cast<or_complex*>(oracle.data()));
            full.forward(input.data(), spec.data());
            require(or_max_relative_error_complex(spec.data(), oracle.data(), spec.size()) <= 1e-12,
                    "hip forward vs direct DFT, fixture " + std::to_string(f));
            auto destroyed = oracle;
            full.inverse(destroyed.data(), real.data());
            require(or_max_relative_error_real(real.data(), input.data(), real.size(), 1.0 / 64.0) <= 1e-12,
                    "hip inverse round trip, fixture " + std::to_string(f));
            or_direct_retained(8, 8, 7, 2, om.data(), om.size(), input.data(),
                               reinterpret_cast<or_complex*>(retainedOracle.data()));
            pruned.forward(input.data(), retained.data());
            require(or_max_relative_error_complex(retained.data(), retainedOracle.data(), retained.size()) < 1e-12,
                    "pruned forward vs direct retained DFT");
            require(input == preserved, "pruned forward modified caller input");
            or_embed_retained(8, 8, 7, 2, om.data(), om.size(), reinterpret_cast<or_complex*>(retainedOracle.data()),
                              reinterpret_cast<or_complex*>(embedded.data()));
            or_direct_c2r(8, 8, 7, 2, reinterpret_cast<or_complex*>(embedded.data()), realOracle.data());
            pruned.inverse(retainedOracle.data(), real.data());
            require(or_max_relative_error_real(real.data(), realOracle.data(), real.size(), 1.0) < 1e-12,
                    "pruned inverse vs direct DFT");
            std::vector<Complex> gathered(retained.size());
            full.gatherRetainedOuter(modes, oracle.data(), gathered.data());
            require(or_max_relative_error_complex(gathered.data(), retainedOracle.data(), retained.size()) < 1e-12,
                    "gatherRetainedOuter");
        }

        // Split compact forms (tests/skbench_tests.cpp:1341-1460): direct split,
        // 0.25 normalisation, field views with mode stride Nz + 3.
        {
            HipStreamingPrunedSplitProvider streaming(w, modes);
            require(streaming.activeKxCount() == pruned.activeKxCount() && streaming.tileWidth() == 1,
                    "streaming split accounting");
            const std::size_t n = modes.size() * w.planes();
            std::vector<double> re(n), im(n), oracleRe(n), oracleIm(n), re4(n), im4(n);
            const std::size_t stride = w.nz + 3;
            const std::size_t per = w.nz + stride * (modes.size() - 1);
            std::vector<double> viewRe(w.fields * per, 7.25), viewIm(w.fields * per, 7.25);
            std::vector<HipStreamingPrunedSplitProvider::FieldView> out(w.fields);
            std::vector<HipStreamingPrunedSplitProvider::ConstFieldView> in(w.fields);
            for (std::size_t f = 0; f < w.fields; ++f) {
                out[f] = {viewRe.data() + f * per, viewIm.data() + f * per, stride};
                in[f] = {viewRe.data() + f * per, viewIm.data() + f * per, stride};
            }
            for (int f = 0; f < 5; ++f) {
                const auto input = makeFixture(w, static_cast<FixtureKind>(f), 900 + f);
                or_direct_retained(8, 8, 7, 2, om.data(), om.size(), input.data(),
                                   reinterpret_cast<or_complex*>(retainedOracle.data()));
                for (std::size_t i = 0; i < n; ++i) {
                    oracleRe[i] = retainedOracle[i].real;
                    oracleIm[i] = retainedOracle[i].imag;
                }
                const std::string tag = ", fixture " + std::to_string(f);
                pruned.forwardSplit(input.data(), re.data(), im.data());
                require(or_max_relative_error_real(re.data(), oracleRe.data(), n, 1.0) < 1e-12 &&
                            or_max_relative_error_real(im.data(), oracleIm.data(), n, 1.0) < 1e-12,
                        "pruned retained-split forward vs direct retained DFT" + tag);
                pruned.forwardSplit(input.data(), re4.data(), im4.data(), 0.25);
                bool scaled = true;
                for (std::size_t i = 0; i < n; ++i) scaled = scaled && re4[i] == 0.25 * re[i] && im4[i] == 0.25 * im[i];
                require(scaled, "retained-split normalisation equals a separate scale" + tag);
                streaming.forwardSplit(input.data(), re4.data(), im4.data());
                require(re4 == re && im4 == im, "streaming split forward equals pruned split forward" + tag);
                or_embed_retained(8, 8, 7, 2, om.data(), om.size(), reinterpret_cast<or_complex*>(retainedOracle.data()),
                                  reinterpret_cast<or_complex*>(embedded.data()));
                or_direct_c2r(8, 8, 7, 2, reinterpret_cast<or_complex*>(embedded.data()), realOracle.data());
                pruned.inverseSplit(oracleRe.data(), oracleIm.data(), real.data());
                require(or_max_relative_error_real(real.data(), realOracle.data(), real.size(), 1.0) < 1e-12,
                        "pruned retained-split inverse vs direct DFT" + tag);
                streaming.inverseSplit(oracleRe.data(), oracleIm.data(), real.data());
                require(or_max_relative_error_real(real.data(), realOracle.data(), real.size(), 1.0) < 1e-12,
                        "streaming split inverse vs direct DFT" + tag);
                streaming.forwardSplitFields(input.data(), out.data(), out.size());
                bool views = true, gaps = true;
                for (std::size_t fld = 0; fld < w.fields; ++fld)
                    for (std::size_t m = 0; m < modes.size(); ++m)
                        for (std::size_t z = 0; z < (m + 1 < modes.size() ? stride : w.nz); ++z) {
                            const std::size_t at = fld * per + z + stride * m;
                            if (z < w.nz) {
                                const std::size_t src = z + w.nz * fld + w.planes() * m;
                                views = views && viewRe[at] == re[src] && viewIm[at] == im[src];
                            } else {
                                gaps = gaps && viewRe[at] == 7.25 && viewIm[at] == 7.25;
                            }
                        }
                require(views, "field views hold the split forward" + tag);
                require(gaps, "field-view gaps untouched" + tag);
                std::vector<double> real2(real.size());
                streaming.inverseSplit(re.data(), im.data(), real.data());
                streaming.inverseSplitFields(in.data(), in.size(), real2.data());
                require(real == real2, "split-field inverse equals split inverse" + tag);
            }
            bool threwCount = false, threwStride = false, threwTile = false;
            try {
                streaming.forwardSplitFields(makeFixture(w, FixtureKind::dc, 1).data(), out.data(), 1);
            } catch (const std::invalid_argument&) {
                threwCount = true;
            }
            auto badStride = out;
            badStride[1].modeStride = w.nz - 1;
            try {
                streaming.forwardSplitFields(makeFixture(w, FixtureKind::dc, 1).data(), badStride.data(), badStride.size());
            } catch (const std::invalid_argument&) {
                threwStride = true;
            }
            try {
                HipStreamingPrunedSplitProvider wide(w, modes, Precision::float64, 17);
            } catch (const std::invalid_argument&) {
                threwTile = true;
            }
            require(threwCount && threwStride && threwTile, "split field-view and tile-width validation");
        }

        const auto ops = squaredWavenumberVerticalFixture(w, modes);
        HipVerticalGemmProvider vertical(w, ops);
        require(vertical.groupCount() == 5 && vertical.gemmCallsPerExecution() == 1, "vertical accounting");
        std::vector<Complex> phys(vertical.physicalElements()), modal(vertical.modalElements()),
            modalOracle(vertical.modalElements());
        for (size_t i = 0; i < phys.size(); ++i)
            phys[i] = {static_cast<double>(i % 31) / 31.0, -static_cast<double>(i % 37) / 37.0};
        vertical.loadPhysicalInput(phys.data());
        vertical.executeForward();
        vertical.copyForwardOutput(modal.data());
        std::vector<uint64_t> first, count;
        for (const auto& g : ops.groups) {
            first.push_back(g.firstMode);
            count.push_back(g.modeCount);
        }
        or_vertical_forward(w.nz, ops.nj, w.fields, first.size(), first.data(), count.data(), ops.forward.data(),
                            reinterpret_cast<or_complex*>(phys.data()), reinterpret_cast<or_complex*>(modalOracle.data()));
        require(or_max_relative_error_complex(modal.data(), modalOracle.data(), modal.size()) <= 1e-12,
                "vertical forward vs scalar products");

        // WVM-direct vertical GEMM (src/vertical_gemm.cpp:1133-1157): on stored
        // values; with real matrices gather(out) equals the compact product of
        // gather(in) (gather conjugates k < 0 modes), and the k_x = 0 partner
        // holds the conjugate.
        {
            HipWvmDirectVerticalGemmProvider wvm(w, modes, ops);
            std::vector<Complex> full(w.spectrumElements()), fullModal(wvm.modalSpectrumElements());
            for (size_t i = 0; i < full.size(); ++i)
                full[i] = {std::sin(0.37 * static_cast<double>(i)), std::cos(0.11 * static_cast<double>(i) + 0.5)};
            wvm.initializeModalOutput(fullModal.data());
            wvm.executeForward(full.data(), fullModal.data());
            std::vector<Complex> gathered(modes.size() * w.planes()), compact(modes.size() * w.fields * ops.nj),
                gatheredModal(compact.size());
            or_gather_retained(8, 8, 7, 2, om.data(), om.size(), reinterpret_cast<const or_complex*>(full.data()),
                               reinterpret_cast<or_complex*>(gathered.data()));
            or_vertical_forward(w.nz, ops.nj, w.fields, first.size(), first.data(), count.data(), ops.forward.data(),
                                reinterpret_cast<or_complex*>(gathered.data()),
                                reinterpret_cast<or_complex*>(compact.data()));
            or_gather_retained(8, 8, ops.nj, 2, om.data(), om.size(),
                               reinterpret_cast<const or_complex*>(fullModal.data()),
                               reinterpret_cast<or_complex*>(gatheredModal.data()));
            require(or_max_relative_error_complex(gatheredModal.data(), compact.data(), compact.size()) <= 1e-12,
                    "WVM-direct forward vs compact product of gathered input");
            bool partner = true;
            const std::size_t nxh = w.nxHalf(), per = w.fields * ops.nj;
            for (const auto& md : modes)
                if (md.storedKx == 0 && md.storedKy != 0 && 2 * md.storedKy != w.ny)
                    for (std::size_t e = 0; e < per; ++e) {
                        const auto a = fullModal[nxh * md.storedKy * per + e];
                        const auto b = fullModal[nxh * ((w.ny - md.storedKy) % w.ny) * per + e];
                        partner = partner && a.real == b.real && a.imag == -b.imag;
                    }
            require(partner, "WVM-direct k_x = 0 partners hold the conjugate");
            bool threw = false;
            try {
                auto bad = ops;
                bad.nj = w.nz;
                HipWvmDirectVerticalGemmProvider wrong(w, modes, bad);
            } catch (const std::invalid_argument&) {
                threw = true;
            }
            require(threw, "WVM-direct requires Nj = 2(Nz-1)/3");
        }

        // Split-layout vertical GEMM (VerticalGemmLayout::split,
        // src/vertical_gemm.cpp:360-390, 698-935).  Every output below is read
        // right after the call returns, with no device synchronisation of the
        // test's own: the adapter calls finish their stream before returning.
        {
            HipVerticalGemmProvider split(w, ops, VerticalGemmLayout::split);
            require(split.supported() && split.layout() == VerticalGemmLayout::split, "split vertical provider");
            double* inRe = split.splitPhysicalInputRealData();
            double* inIm = split.splitPhysicalInputImaginaryData();
            for (size_t i = 0; i < phys.size(); ++i) {
                inRe[i] = phys[i].real;
                inIm[i] = phys[i].imag;
            }
            split.executeForward();
            const double* outRe = split.splitModalOutputRealData();
            const double* outIm = split.splitModalOutputImaginaryData();
            bool same = true;
            for (size_t i = 0; i < modal.size(); ++i) same = same && outRe[i] == modal[i].real && outIm[i] == modal[i].imag;
            require(same, "split forward equals the interleaved forward");
            std::vector<double> keepRe(outRe, outRe + modal.size()), keepIm(outIm, outIm + modal.size());
            // the parts one at a time: the real part depends on the real input alone
            for (size_t i = 0; i < phys.size(); ++i) inIm[i] = 1.0e6;
            split.executeForwardReal();
            bool realPart = true;
            for (size_t i = 0; i < modal.size(); ++i) realPart = realPart && outRe[i] == keepRe[i] && outIm[i] == keepIm[i];
            require(realPart, "executeForwardReal writes the real part alone");
            for (size_t i = 0; i < phys.size(); ++i) inIm[i] = phys[i].imag;
            for (size_t i = 0; i < phys.size(); ++i) inRe[i] = -3.0;
            split.executeForwardImaginary();
            bool imagPart = true;
            for (size_t i = 0; i < modal.size(); ++i) imagPart = imagPart && outRe[i] == keepRe[i] && outIm[i] == keepIm[i];
            require(imagPart, "executeForwardImaginary writes the imaginary part alone");
            std::vector<Complex> copied(modal.size());
            split.copyForwardOutput(copied.data());
            require(or_max_relative_error_complex(copied.data(), modalOracle.data(), copied.size()) <= 1e-12,
                    "split forward vs scalar products");

            // inverse, interleaved and split, with the modal output as input
            vertical.loadModalInput(modal.data());
            vertical.executeInverse();
            std::vector<Complex> back(phys.size());
            vertical.copyInverseOutput(back.data());
            double* mRe = split.splitModalInputRealData();
            double* mIm = split.splitModalInputImaginaryData();
            for (size_t i = 0; i < modal.size(); ++i) {
                mRe[i] = modal[i].real;
                mIm[i] = modal[i].imag;
            }
            split.executeInverseReal();
            split.executeInverseImaginary();
            const double* pRe = split.splitPhysicalOutputRealData();
            const double* pIm = split.splitPhysicalOutputImaginaryData();
            bool inv = true;
            for (size_t i = 0; i < back.size(); ++i) inv = inv && pRe[i] == back[i].real && pIm[i] == back[i].imag;
            require(inv, "split inverse parts equal the interleaved inverse");
            split.executeInverse();
            bool inv2 = true;
            for (size_t i = 0; i < back.size(); ++i) inv2 = inv2 && pRe[i] == back[i].real && pIm[i] == back[i].imag;
            require(inv2, "split inverse equals the interleaved inverse");
            // K^2 family: orthonormal rows, so inverse(forward(x)) keeps the
            // projection of x; compare against the scalar inverse product.
            const Complex* ip = vertical.interleavedPhysicalOutputData();
            require(ip[0].real == back[0].real && ip[0].imag == back[0].imag, "interleaved physical output accessor");

            // WVM pack / embed (src/vertical_gemm.cpp:470-516) against the
            // oracle's gather / embed
            std::vector<Complex> full(w.spectrumElements());
            for (size_t i = 0; i < full.size(); ++i)
                full[i] = {std::cos(0.23 * static_cast<double>(i)), std::sin(0.41 * static_cast<double>(i) + 0.2)};
            std::vector<Complex> gathered(phys.size());
            or_gather_retained(8, 8, 7, 2, om.data(), om.size(), reinterpret_cast<const or_complex*>(full.data()),
                               reinterpret_cast<or_complex*>(gathered.data()));
            for (auto* provider : {&vertical, &split}) {
                provider->packPhysicalInputFromWvm(modes, full.data());
                provider->executeForward();
                std::vector<Complex> out(modal.size()), expect(modal.size());
                provider->copyForwardOutput(out.data());
                vertical.loadPhysicalInput(gathered.data());
                vertical.executeForward();
                vertical.copyForwardOutput(expect.data());
                require(same_values(out, expect), "packPhysicalInputFromWvm equals the gathered load");
                provider->executeInverse();
                std::vector<Complex> compactOut(phys.size()), wvmOut(full.size(), Complex{9.0, 9.0}),
                    wvmOracle(full.size());
                provider->copyInverseOutput(compactOut.data());
                provider->embedPhysicalOutputToWvm(modes, wvmOut.data());
                or_embed_retained(8, 8, 7, 2, om.data(), om.size(), reinterpret_cast<or_complex*>(compactOut.data()),
                                  reinterpret_cast<or_complex*>(wvmOracle.data()));
                bool embedOk = true;
                for (size_t i = 0; i < full.size(); ++i)
                    embedOk = embedOk && wvmOut[i].real == wvmOracle[i].real && wvmOut[i].imag == wvmOracle[i].imag;
                require(embedOk, std::string("embedPhysicalOutputToWvm vs oracle embed, ") +
                                     verticalGemmLayoutName(provider->layout()));
            }

            // component / policy errors (src/vertical_gemm.cpp:433-467)
            int logic = 0;
            auto expectLogic = [&](auto&& f) {
                try {
                    f();
                } catch (const std::logic_error&) {
                    ++logic;
                }
            };
            expectLogic([&] { (void)split.interleavedPhysicalInputData(); });
            expectLogic([&] { (void)vertical.splitModalOutputRealData(); });
            expectLogic([&] { vertical.executeForwardReal(); });
            HipVerticalGemmProvider forwardOnly(w, ops, VerticalGemmLayout::split, VerticalGemmStrategy{},
                                                VerticalGemmBufferPolicy::forwardOnly);
            expectLogic([&] { forwardOnly.executeInverse(); });
            expectLogic([&] { (void)forwardOnly.splitModalInputRealData(); });
            require(logic == 5, "layout and buffer-policy errors are std::logic_error");
            bool wrongModes = false;
            try {
                split.packPhysicalInputFromWvm(std::vector<RetainedMode>(modes.begin(), modes.end() - 1), full.data());
            } catch (const std::invalid_argument&) {
                wrongModes = true;
            }
            require(wrongModes, "packPhysicalInputFromWvm checks the retained-mode count");

            // dimensions the FP64 GEMM does not instantiate: supported() false,
            // execution throws std::runtime_error with the capability text
            const Workload deep{8, 8, 257, 1, 1.0, 1.0, true};
            const auto deepModes = retainedHorizontalModes(deep);
            HipVerticalGemmProvider unsupported(deep, squaredWavenumberVerticalFixture(deep, deepModes));
            require(!unsupported.supported() && unsupported.capability().rfind("unsupported", 0) == 0,
                    "FP64 Nz = 257 reports unsupported");
            bool runtime = false;
            try {
                unsupported.executeForward();
            } catch (const std::runtime_error&) {
                runtime = true;
            }
            require(runtime, "unsupported vertical provider throws std::runtime_error on execution");
        }

        // FFT provider split forms (src/fftw_provider.cpp:331-395, 701-830)
        // against the interleaved forms already checked against the oracle.
        {
            HipProvider splitFft(w, DataLayout::split);
            require(splitFft.layout() == DataLayout::split && !splitFft.splitInPlaceWvmOrderSupported(),
                    "split FFT provider layout");
            const auto input = makeFixture(w, FixtureKind::random, 4242);
            std::vector<Complex> ref(w.spectrumElements());
            full.forward(input.data(), ref.data());
            const std::size_t ns = ref.size(), nr = modes.size() * w.planes();
            std::vector<double> sRe(ns), sIm(ns);
            splitFft.forwardSplit(input.data(), sRe.data(), sIm.data());
            bool fwd = true;
            for (size_t i = 0; i < ns; ++i) fwd = fwd && sRe[i] == ref[i].real && sIm[i] == ref[i].imag;
            require(fwd, "forwardSplit equals forward");
            std::vector<Complex> refRet(nr), embedRef(ns);
            full.gatherRetainedOuter(modes, ref.data(), refRet.data());
            std::vector<double> rRe(nr), rIm(nr);
            splitFft.gatherRetainedSplitOuter(modes, sRe.data(), sIm.data(), rRe.data(), rIm.data());
            bool gat = true;
            for (size_t i = 0; i < nr; ++i) gat = gat && rRe[i] == refRet[i].real && rIm[i] == refRet[i].imag;
            require(gat, "gatherRetainedSplitOuter equals gatherRetainedOuter");
            splitFft.gatherRetainedToSplitOuter(modes, ref.data(), rRe.data(), rIm.data(), 0.5);
            bool gat2 = true;
            for (size_t i = 0; i < nr; ++i)
                gat2 = gat2 && rRe[i] == 0.5 * refRet[i].real && rIm[i] == 0.5 * refRet[i].imag;
            require(gat2, "gatherRetainedToSplitOuter applies the scale");
            splitFft.scaleRetainedSplitOuter(modes, rRe.data(), rIm.data(), 2.0);
            bool sc = true;
            for (size_t i = 0; i < nr; ++i) sc = sc && rRe[i] == refRet[i].real && rIm[i] == refRet[i].imag;
            require(sc, "scaleRetainedSplitOuter");
            full.embedRetainedOuter(modes, refRet.data(), embedRef.data());
            std::vector<Complex> emb(ns, Complex{5.0, 5.0});
            splitFft.embedRetainedFromSplitOuter(modes, rRe.data(), rIm.data(), emb.data());
            require(same_values(emb, embedRef), "embedRetainedFromSplitOuter equals embedRetainedOuter");
            std::vector<double> eRe(ns, 5.0), eIm(ns, 5.0);
            splitFft.embedRetainedSplitOuter(modes, rRe.data(), rIm.data(), eRe.data(), eIm.data());
            bool emb2 = true;
            for (size_t i = 0; i < ns; ++i) emb2 = emb2 && eRe[i] == embedRef[i].real && eIm[i] == embedRef[i].imag;
            require(emb2, "embedRetainedSplitOuter equals embedRetainedOuter");
            std::vector<double> outA(w.realElements()), outB(w.realElements());
            auto destroyed = embedRef;
            full.inverse(destroyed.data(), outA.data());
            splitFft.inverseSplit(eRe.data(), eIm.data(), outB.data());
            require(outA == outB, "inverseSplit equals inverse");
            int layoutErrors = 0;
            try {
                splitFft.forward(input.data(), ref.data());
            } catch (const std::logic_error&) {
                ++layoutErrors;
            }
            try {
                full.forwardSplit(input.data(), sRe.data(), sIm.data());
            } catch (const std::logic_error&) {
                ++layoutErrors;
            }
            require(layoutErrors == 2, "FFT entry points check the planned data layout");
        }

        // Native-stride WVM field views (include/skbench/skbench.hpp:423-465):
        // the two fields of the 8x8x7x2 WVM array as views of stride 14
        {
            const Workload one{8, 8, 7, 1, 1.0, 1.0, true};
            HipInterleavedWvmFieldViewProvider views(one, {14, 7});
            require(views.inverseInputIsDestructive(), "field-view inverse may overwrite its views");
            const auto input = makeFixture(w, FixtureKind::random, 77);
            std::vector<Complex> expect(w.spectrumElements()), got(w.spectrumElements(), Complex{3.0, 3.0});
            or_direct_r2c(8, 8, 7, 2, input.data(), reinterpret_cast<or_complex*>(expect.data()));
            const std::size_t vol = one.realElements();
            views.forwardField(input.data() + vol, {got.data() + 7, 14});
            bool untouched = true;
            for (std::size_t q = 0; q < w.nxHalf() * w.ny; ++q)
                for (std::size_t z = 0; z < 7; ++z)
                    untouched = untouched && got[z + 14 * q].real == 3.0 && got[z + 14 * q].imag == 3.0;
            require(untouched, "forwardField writes its own view only");
            views.forwardField(input.data(), {got.data(), 14});
            require(or_max_relative_error_complex(got.data(), expect.data(), got.size()) <= 1e-12,
                    "forwardField views vs direct DFT");
            std::vector<double> back(w.realElements());
            HipInterleavedWvmFieldViewProvider::MutableFieldView fv[2] = {{got.data(), 14}, {got.data() + 7, 14}};
            views.inverseFields(fv, 2, back.data());
            require(or_max_relative_error_real(back.data(), input.data(), back.size(), 1.0 / 64.0) <= 1e-12,
                    "inverseFields round trip");
            int errors = 0;
            try {
                views.forwardField(input.data(), {got.data(), 9});
            } catch (const std::invalid_argument&) {
                ++errors;
            }
            try {
                HipInterleavedWvmFieldViewProvider bad(w, {14});
            } catch (const std::invalid_argument&) {
                ++errors;
            }
            try {
                HipInterleavedWvmFieldViewProvider inv(one, {14}, VerticalGemmBufferPolicy::inverseOnly);
                inv.forwardField(input.data(), {got.data(), 14});
            } catch (const std::logic_error&) {
                ++errors;
            }
            require(errors == 3, "field-view stride, workload and policy checks");
        }

        HipPointwiseAdvectionExecutor advection(37, 0.03125);
        std::vector<double> shared(111), derivative(111), target(37), expected(37);
        for (size_t i = 0; i < 111; ++i) {
            shared[i] = static_cast<double>(i % 17) - 8.0;
            derivative[i] = static_cast<double>((3 * i + 5) % 23) - 11.0;
        }
        for (size_t p = 0; p < 37; ++p)
            expected[p] = -0.03125 * (shared[p] * derivative[p] + shared[37 + p] * derivative[37 + p] +
                                      shared[74 + p] * derivative[74 + p]);
        advection.execute(shared.data(), derivative.data(), target.data());
        require(or_max_relative_error_real(target.data(), expected.data(), 37, 1.0) <= 1e-15, "pointwise tail");

        bool threw = false;
        try {
            HipPrunedProvider bad(w, {});
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        require(threw, "empty mode list must throw std::invalid_argument");
        HipProvider odd(Workload{12, 8, 2, 1, 1.0, 1.0, true});
        require(!odd.supported() && !odd.capability().empty(), "non-power-of-two grid reports unsupported");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 2;
    }
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("skbench_hip_tests: all checks passed\n");
    return 0;
}
