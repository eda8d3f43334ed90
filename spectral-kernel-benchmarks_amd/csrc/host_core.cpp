// Host-side model of the retained-mode operator: workload validation,
// retained-mode enumeration, K^2 grouping, FNV-1a hashes, vertical matrix
// families and input fixtures.  This is the product's own C++20 statement of
// src/model.cpp (skbench); the GPU plans consume its outputs.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numbers>
#include <string>
#include <unordered_set>
#include <vector>

#include "common.hpp"

namespace skb {
namespace {

thread_local std::string g_last_error;

constexpr double kPi = std::numbers::pi;

uint64_t half_width(uint64_t nx) { return nx / 2 + 1; }

// The DFT mode stored at list position `index` (src/model.cpp:34-45):
// 0..ceil(n/2)-1 then -(n/2)..-1.
int64_t list_mode(uint64_t index, uint64_t n) {
    const uint64_t positive = (n + 1) / 2;
    return index < positive ? static_cast<int64_t>(index) : -static_cast<int64_t>(n / 2 - (index - positive));
}

bool self_conjugate(int64_t m, uint64_t n) { return m == 0 || (n % 2 == 0 && m == -static_cast<int64_t>(n / 2)); }

// src/model.cpp:56-65
bool primary(int64_t k, int64_t l, uint64_t nx, uint64_t ny) {
    return l > 0 || (self_conjugate(l, ny) && (k > 0 || self_conjugate(k, nx)));
}
bool nyquist(int64_t k, int64_t l, uint64_t nx, uint64_t ny) {
    return (nx % 2 == 0 && k == -static_cast<int64_t>(nx / 2)) || (ny % 2 == 0 && l == -static_cast<int64_t>(ny / 2));
}

// src/model.cpp:67-71: the squares are rounded separately so the ordering
// of equal-|k| modes is the same on every compiler.
double radial(double a, double b) {
    volatile double a2 = a * a;
    volatile double b2 = b * b;
    return std::sqrt(a2 + b2);
}

uint64_t wrap(int64_t m, uint64_t n) { return m >= 0 ? static_cast<uint64_t>(m) : static_cast<uint64_t>(static_cast<int64_t>(n) + m); }

uint64_t magnitude(int64_t v) { return v < 0 ? static_cast<uint64_t>(-(v + 1)) + 1 : static_cast<uint64_t>(v); }

uint64_t squared_key(const skb_retained_mode& m) {
    const uint64_t a = magnitude(m.k), b = magnitude(m.l);
    const uint64_t top = std::numeric_limits<uint64_t>::max();
    if ((a != 0 && a > top / a) || (b != 0 && b > top / b) || b * b > top - a * a)
        fail(SKB_INVALID_ARGUMENT, "squared horizontal mode key overflows uint64_t");
    return a * a + b * b;
}

struct Fnv {
    uint64_t h = 14695981039346656037ull;
    void add(uint64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    }
};

// Portable generator (src/benchmark.cpp:187-196).
struct SignedUnit {
    uint64_t state;
    explicit SignedUnit(uint64_t seed) : state(seed == 0 ? 0x9e3779b97f4a7c15ull : seed) {}
    double next() {
        state ^= state >> 12;
        state ^= state << 25;
        state ^= state >> 27;
        const uint64_t v = state * 2685821657736338717ull;
        return 2.0 * static_cast<double>(v >> 11) / static_cast<double>(1ull << 53) - 1.0;
    }
};

// Box-Muller pairs: pair i consumes draws 2i and 2i+1 (cosine, then sine).
std::vector<double> gaussians(uint64_t seed, uint64_t count) {
    SignedUnit rng(seed);
    std::vector<double> out((count + 1) / 2 * 2);
    for (uint64_t i = 0; i < out.size(); i += 2) {
        const double u1 = (rng.next() + 1.0) * 0.5;
        const double u2 = (rng.next() + 1.0) * 0.5;
        const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
        out[i] = r * std::cos(2.0 * kPi * u2);
        out[i + 1] = r * std::sin(2.0 * kPi * u2);
    }
    out.resize(count);
    return out;
}

void check_groups(const skb_mode_group* g, uint64_t n) {
    if (!g && n) fail(SKB_INVALID_ARGUMENT, "null group table");
}

}  // namespace

void set_last_error(const std::string& message) { g_last_error = message; }

void validate_workload(const skb_workload* w) {
    // src/model.cpp:25-32
    if (!w) fail(SKB_INVALID_ARGUMENT, "null workload");
    if (w->nx < 2 || w->ny < 2 || w->nz < 1 || w->fields < 1)
        fail(SKB_INVALID_ARGUMENT, "Nx and Ny must be at least two; Nz and fields must be positive.");
    if (!(w->lx > 0.0) || !(w->ly > 0.0)) fail(SKB_INVALID_ARGUMENT, "Horizontal domain lengths must be positive.");
}

std::vector<skb_retained_mode> retained_modes(const skb_workload& w) {
    validate_workload(&w);
    const double kmax = 2.0 * kPi * (static_cast<double>(w.nx / 2) / w.lx);
    std::vector<skb_retained_mode> modes;
    modes.reserve(half_width(w.nx) * w.ny);
    for (uint64_t li = 0; li < w.ny; ++li) {
        const int64_t l = list_mode(li, w.ny);
        for (uint64_t ki = 0; ki < w.nx; ++ki) {
            const int64_t k = list_mode(ki, w.nx);
            if (!primary(k, l, w.nx, w.ny) || nyquist(k, l, w.nx, w.ny)) continue;
            const double r = radial(2.0 * kPi * static_cast<double>(k) / w.lx, 2.0 * kPi * static_cast<double>(l) / w.ly);
            if (w.antialias && r > 2.0 * kmax / 3.0) continue;
            skb_retained_mode m{};
            m.k = k;
            m.l = l;
            m.radial = r;
            if (k >= 0) {
                m.stored_kx = static_cast<uint64_t>(k);
                m.stored_ky = wrap(l, w.ny);
            } else {
                m.stored_kx = static_cast<uint64_t>(-k);
                m.stored_ky = wrap(-l, w.ny);
                m.conjugates_stored_value = 1;
            }
            modes.push_back(m);
        }
    }
    std::stable_sort(modes.begin(), modes.end(), [](const skb_retained_mode& a, const skb_retained_mode& b) {
        if (a.radial != b.radial) return a.radial < b.radial;
        if (a.k != b.k) return a.k < b.k;
        return a.l < b.l;
    });
    return modes;
}

std::vector<skb_mode_group> k2_groups(const skb_retained_mode* modes, uint64_t n) {
    if (n == 0 || !modes) fail(SKB_INVALID_ARGUMENT, "Squared-wavenumber grouping requires retained modes.");
    std::vector<skb_mode_group> groups;
    std::unordered_set<uint64_t> closed;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t key = squared_key(modes[i]);
        if (groups.empty() || groups.back().squared_key != key) {
            if (closed.count(key))
                fail(SKB_INVALID_ARGUMENT, "Equal squared-wavenumber modes are not contiguous in the retained order.");
            if (!groups.empty()) closed.insert(groups.back().squared_key);
            groups.push_back({key, i, 1});
        } else {
            ++groups.back().mode_count;
        }
    }
    return groups;
}

std::vector<skb_mode_group> k2_binned_groups(const skb_retained_mode* modes, uint64_t n, uint64_t bins) {
    if (n == 0 || !modes) fail(SKB_INVALID_ARGUMENT, "binned grouping requires retained modes");
    if (bins == 0) fail(SKB_INVALID_ARGUMENT, "bin count must be positive");
    uint64_t kmax = 0;
    for (uint64_t i = 0; i < n; ++i) kmax = std::max(kmax, squared_key(modes[i]));
    std::vector<skb_mode_group> groups;
    std::unordered_set<uint64_t> seen;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t bin = bins * squared_key(modes[i]) / (kmax + 1);
        if (groups.empty() || groups.back().squared_key != bin) {
            if (!seen.insert(bin).second) fail(SKB_INVALID_ARGUMENT, "K^2 bins are not contiguous in the retained order");
            groups.push_back({bin, i, 1});
        } else {
            ++groups.back().mode_count;
        }
    }
    return groups;
}

void orthonormal_vertical(uint64_t nz, uint64_t nj, double* fwd, double* inv) {
    if (nz == 0 || nj == 0 || nj > nz) fail(SKB_INVALID_ARGUMENT, "Vertical fixture requires 1 <= Nj <= Nz.");
    for (uint64_t j = 0; j < nj; ++j) {
        const double norm = j == 0 ? std::sqrt(1.0 / static_cast<double>(nz)) : std::sqrt(2.0 / static_cast<double>(nz));
        for (uint64_t z = 0; z < nz; ++z) {
            const double v = norm * std::cos(kPi * (static_cast<double>(z) + 0.5) * static_cast<double>(j) / static_cast<double>(nz));
            fwd[j * nz + z] = v;
            inv[z * nj + j] = v;
        }
    }
}

}  // namespace skb

using namespace skb;

extern "C" {

size_t skb_last_error(char* buffer, size_t capacity) {
    const std::string& m = g_last_error;
    if (buffer && capacity) {
        const size_t n = std::min(capacity - 1, m.size());
        std::memcpy(buffer, m.data(), n);
        buffer[n] = '\0';
    }
    return m.size();
}

const char* skb_version(void) { return "skb-hip 0.1 gfx950 (hand-written HIP; Stockham LDS FFT, MFMA grouped GEMM)"; }

int skb_retained_modes(const skb_workload* workload, skb_retained_mode* modes, uint64_t capacity, uint64_t* count) {
    return guarded([&] {
        const auto all = retained_modes(*workload);
        if (count) *count = all.size();
        if (modes) {
            if (capacity < all.size()) fail(SKB_INVALID_ARGUMENT, "retained-mode buffer too small");
            std::memcpy(modes, all.data(), all.size() * sizeof(skb_retained_mode));
        }
    });
}

int skb_k2_groups(const skb_retained_mode* modes, uint64_t mode_count, skb_mode_group* groups, uint64_t capacity,
                  uint64_t* count) {
    return guarded([&] {
        const auto g = k2_groups(modes, mode_count);
        if (count) *count = g.size();
        if (groups) {
            if (capacity < g.size()) fail(SKB_INVALID_ARGUMENT, "group buffer too small");
            std::memcpy(groups, g.data(), g.size() * sizeof(skb_mode_group));
        }
    });
}

int skb_k2_binned_groups(const skb_retained_mode* modes, uint64_t mode_count, uint64_t bins, skb_mode_group* groups,
                         uint64_t capacity, uint64_t* count) {
    return guarded([&] {
        const auto g = k2_binned_groups(modes, mode_count, bins);
        if (count) *count = g.size();
        if (groups) {
            if (capacity < g.size()) fail(SKB_INVALID_ARGUMENT, "group buffer too small");
            std::memcpy(groups, g.data(), g.size() * sizeof(skb_mode_group));
        }
    });
}

uint64_t skb_mode_order_hash(const skb_retained_mode* modes, uint64_t count) {
    Fnv f;
    for (uint64_t i = 0; i < count; ++i) {
        f.add(static_cast<uint64_t>(modes[i].k));
        f.add(static_cast<uint64_t>(modes[i].l));
    }
    return f.h;
}

uint64_t skb_group_order_hash(const skb_mode_group* groups, uint64_t count) {
    Fnv f;
    for (uint64_t g = 0; g < count; ++g) {
        f.add(groups[g].squared_key);
        f.add(groups[g].first_mode);
        f.add(groups[g].mode_count);
    }
    return f.h;
}

uint64_t skb_wvm_order_hash(const skb_workload* w) {
    // Visits WVM storage in memory order (z fastest, field, kx, ky) and hashes
    // the logical (kx, signed ky, z, field) at each slot (src/model.cpp:471-487).
    Fnv f;
    for (uint64_t ky = 0; ky < w->ny; ++ky) {
        const int64_t sky = ky <= w->ny / 2 ? static_cast<int64_t>(ky) : static_cast<int64_t>(ky) - static_cast<int64_t>(w->ny);
        for (uint64_t kx = 0; kx < half_width(w->nx); ++kx)
            for (uint64_t field = 0; field < w->fields; ++field)
                for (uint64_t z = 0; z < w->nz; ++z) {
                    f.add(kx);
                    f.add(static_cast<uint64_t>(sky));
                    f.add(z);
                    f.add(field);
                }
    }
    return f.h;
}

int skb_orthonormal_vertical(uint64_t nz, uint64_t nj, double* forward, double* inverse) {
    return guarded([&] { orthonormal_vertical(nz, nj, forward, inverse); });
}

int skb_k2_vertical_family(uint64_t nz, uint64_t nj, const skb_mode_group* groups, uint64_t group_count,
                           double* forward, double* inverse) {
    return guarded([&] {
        check_groups(groups, group_count);
        std::vector<double> bf(nj * nz), bi(nj * nz);
        orthonormal_vertical(nz, nj, bf.data(), bi.data());
        const uint64_t me = nj * nz;
        for (uint64_t g = 0; g < group_count; ++g) {
            // Rows (j, j+1) of the base DCT-II rotated by (pi/2)((key mod 4093)+1)/4094;
            // a trailing odd row is negated for odd keys (src/model.cpp:540-563).
            const uint64_t key = groups[g].squared_key;
            const double angle = 0.5 * kPi * (static_cast<double>((key % 4093u) + 1) / 4094.0);
            const double c = std::cos(angle), s = std::sin(angle);
            double* F = forward + g * me;
            double* I = inverse + g * me;
            for (uint64_t j = 0; j < nj; j += 2)
                for (uint64_t z = 0; z < nz; ++z) {
                    const double a = bf[j * nz + z];
                    double first = a;
                    if (j + 1 < nj) {
                        const double b = bf[(j + 1) * nz + z];
                        first = c * a + s * b;
                        const double second = -s * a + c * b;
                        F[(j + 1) * nz + z] = second;
                        I[z * nj + j + 1] = second;
                    } else if (key & 1u) {
                        first = -a;
                    }
                    F[j * nz + z] = first;
                    I[z * nj + j] = first;
                }
        }
    });
}

int skb_qr_gaussian_family(uint64_t n, uint64_t group_count, uint64_t seed, double* forward, double* inverse) {
    return guarded([&] {
        if (n == 0) fail(SKB_INVALID_ARGUMENT, "matrix order must be positive");
        std::vector<double> A(n * n), Q(n * n), v(n);
        for (uint64_t g = 0; g < group_count; ++g) {
            const auto gauss = gaussians(seed + g, n * n);
            std::copy(gauss.begin(), gauss.end(), A.begin());  // row-major A[i*n+j]
            // Householder QR; reflectors applied to the identity give Q.
            std::fill(Q.begin(), Q.end(), 0.0);
            for (uint64_t i = 0; i < n; ++i) Q[i * n + i] = 1.0;
            std::vector<std::vector<double>> reflectors;
            for (uint64_t k = 0; k + 1 < n; ++k) {
                double norm = 0.0;
                for (uint64_t i = k; i < n; ++i) norm += A[i * n + k] * A[i * n + k];
                norm = std::sqrt(norm);
                std::vector<double> h(n, 0.0);
                if (norm == 0.0) {
                    reflectors.push_back(h);
                    continue;
                }
                const double alpha = A[k * n + k] > 0 ? -norm : norm;
                for (uint64_t i = k; i < n; ++i) h[i] = A[i * n + k];
                h[k] -= alpha;
                double hn = 0.0;
                for (uint64_t i = k; i < n; ++i) hn += h[i] * h[i];
                hn = std::sqrt(hn);
                if (hn == 0.0) {
                    reflectors.push_back(std::vector<double>(n, 0.0));
                    continue;
                }
                for (uint64_t i = k; i < n; ++i) h[i] /= hn;
                for (uint64_t j = 0; j < n; ++j) {
                    double d = 0.0;
                    for (uint64_t i = k; i < n; ++i) d += h[i] * A[i * n + j];
                    for (uint64_t i = k; i < n; ++i) A[i * n + j] -= 2.0 * h[i] * d;
                }
                reflectors.push_back(std::move(h));
            }
            // Q = H_0 H_1 ... H_{n-2}: apply in reverse to the identity.
            for (uint64_t r = reflectors.size(); r-- > 0;) {
                const auto& h = reflectors[r];
                for (uint64_t j = 0; j < n; ++j) {
                    double d = 0.0;
                    for (uint64_t i = 0; i < n; ++i) d += h[i] * Q[i * n + j];
                    if (d == 0.0) continue;
                    for (uint64_t i = 0; i < n; ++i) Q[i * n + j] -= 2.0 * h[i] * d;
                }
            }
            // Unique factor: R with a positive diagonal.
            for (uint64_t j = 0; j < n; ++j) {
                const double sign = A[j * n + j] < 0 ? -1.0 : 1.0;
                for (uint64_t i = 0; i < n; ++i) Q[i * n + j] *= sign;
            }
            double* F = forward + g * n * n;
            double* I = inverse + g * n * n;
            for (uint64_t i = 0; i < n; ++i)
                for (uint64_t j = 0; j < n; ++j) {
                    F[i * n + j] = Q[j * n + i];  // Q^T
                    I[i * n + j] = Q[i * n + j];
                }
        }
    });
}

int skb_make_fixture(const skb_workload* w, int kind, uint64_t seed, double* real) {
    return guarded([&] {
        validate_workload(w);
        if (kind < 0 || kind > 5) fail(SKB_INVALID_ARGUMENT, "unknown fixture kind");
        const uint64_t nx = w->nx, ny = w->ny, nz = w->nz;
        std::fill_n(real, nx * ny * nz * w->fields, 0.0);
        SignedUnit rng(seed);
        for (uint64_t field = 0; field < w->fields; ++field)
            for (uint64_t z = 0; z < nz; ++z) {
                const uint64_t plane = z + nz * field;
                const double amp = 1.0 + 0.125 * static_cast<double>(plane);
                double* P = real + nx * ny * plane;
                if (kind == 0) {
                    P[(1 + plane) % nx + nx * ((2 + 3 * plane) % ny)] = amp;
                    continue;
                }
                for (uint64_t y = 0; y < ny; ++y)
                    for (uint64_t x = 0; x < nx; ++x) {
                        double v = 0.0;
                        switch (kind) {
                            case 1:
                                v = amp * std::cos(2.0 * kPi * (static_cast<double>(x) / static_cast<double>(nx) +
                                                                2.0 * static_cast<double>(y) / static_cast<double>(ny)) +
                                                   0.17 * static_cast<double>(plane));
                                break;
                            case 5:
                                v = amp * std::cos(2.0 * kPi * (5.0 * static_cast<double>(x) / static_cast<double>(nx) +
                                                                17.0 * static_cast<double>(y) / static_cast<double>(ny)) +
                                                   0.17 * static_cast<double>(plane));
                                break;
                            case 2: v = rng.next(); break;
                            case 3: v = amp; break;
                            case 4: v = amp * (((x + y) % 2 == 0) ? 1.0 : -1.0); break;
                        }
                        P[x + nx * y] = v;
                    }
            }
    });
}

int skb_spectral_fixture(const skb_workload* w, const skb_retained_mode* modes, uint64_t mode_count, uint64_t seed,
                         double exponent, double* out) {
    return guarded([&] {
        validate_workload(w);
        if (!modes || mode_count == 0) fail(SKB_INVALID_ARGUMENT, "spectral fixture requires retained modes");
        const uint64_t P = w->nz * w->fields;
        const auto g = gaussians(seed, 2 * P * mode_count);
        for (uint64_t m = 0; m < mode_count; ++m) {
            const double k2 = static_cast<double>(squared_key(modes[m]));
            const double amp = std::pow(1.0 + k2, exponent);
            const bool dc = modes[m].k == 0 && modes[m].l == 0;
            for (uint64_t p = 0; p < P; ++p) {
                const uint64_t i = p + P * m;
                out[2 * i] = g[2 * i] * amp;
                out[2 * i + 1] = dc ? 0.0 : g[2 * i + 1] * amp;
            }
        }
    });
}

// syntheticModalWorkWeights (src/model.cpp:246-267): the real diagonal
// modal weight of (mode, j, field), at the modal index j + Nj*(field +
// F*mode); |k| and |l| are taken in unsigned arithmetic.
int skb_synthetic_modal_weights(const skb_workload* w, const skb_retained_mode* modes, uint64_t mode_count,
                                double* weights) {
    return guarded([&] {
        validate_workload(w);
        if (!modes || !weights) fail(SKB_INVALID_ARGUMENT, "modal weights need modes and an output array");
        const uint64_t nj = w->nz < 2 ? 0 : 2 * (w->nz - 1) / 3, F = w->fields;
        const auto magnitude = [](int64_t v) {
            return v < 0 ? uint64_t{0} - static_cast<uint64_t>(v) : static_cast<uint64_t>(v);
        };
        double* out = weights;
        for (uint64_t m = 0; m < mode_count; ++m) {
            const uint64_t base = 17 * magnitude(modes[m].k) + 29 * magnitude(modes[m].l);
            for (uint64_t f = 0; f < F; ++f)
                for (uint64_t j = 0; j < nj; ++j)
                    *out++ = 0.75 + 0.5 * static_cast<double>((base + 7 * j + 11 * f) % 97) / 96.0;
        }
    });
}

// fillFull / fillCompact of the constant-stratification vertical benchmark
// (src/constant_stratification.cpp:237-269).
int skb_type1_fill(uint64_t rows, uint64_t channels, uint64_t nz, const skb_retained_mode* modes, uint64_t nx_half,
                   double* data) {
    return guarded([&] {
        if (!data || rows == 0 || channels == 0 || nz == 0) fail(SKB_INVALID_ARGUMENT, "bad fill arguments");
        for (uint64_t r = 0; r < rows; ++r) {
            uint64_t row = r;
            bool conj = false;
            if (modes) {
                row = modes[r].stored_kx + nx_half * modes[r].stored_ky;
                conj = modes[r].conjugates_stored_value != 0;
            }
            for (uint64_t c = 0; c < channels; ++c)
                for (uint64_t z = 0; z < nz; ++z) {
                    const double a = 0.017 * static_cast<double>(1 + row % 97) + 0.113 * static_cast<double>(1 + c) +
                                     0.071 * static_cast<double>(1 + z);
                    const uint64_t at = 2 * (z + nz * (c + channels * r));
                    data[at] = std::sin(a);
                    data[at + 1] = conj ? -std::cos(1.7 * a) : std::cos(1.7 * a);
                }
        }
    });
}

}  // extern "C"
