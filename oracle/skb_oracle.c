/*
 * skb_oracle.c -- CPU restatement of skbench's retained-mode spectral-operator
 * path, used ONLY as the checker by the tests, the smoke check and the CPU
 * baseline leg of bench.py.  Nothing in the HIP product links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the skbench source tree).  Plain C99, FP64 throughout,
 * compiled with -ffp-contract=off so that no multiply-add is fused behind the
 * reference's back (the reference forces the same with `volatile` squares in
 * src/model.cpp:67-71).
 *
 * Pinning (what checks each function against something other than itself):
 *   - or_retained_modes / or_mode_order_hash: published mode counts and
 *     retainedModeOrder hashes (workload.permutationHashes in results/published)
 *   - or_k2_groups / or_group_hash: smoke hash fnv1a64:69843c1f93cead00
 *     (tests/skbench_tests.cpp:1998) and published verticalGroupOrderHash values
 *   - or_wvm_order_hash: published wvmFullSpectrumOrder hashes
 *   - or_direct_r2c / or_direct_c2r: closed forms (impulse phase ramp, DC bin,
 *     Nyquist bin, two-bin sinusoid) and agreement with an independent
 *     FFT (scipy pocketfft) in tests/test_oracle.py
 *   - or_gather / or_embed: exact permutation round trip (== 0.0)
 *   - vertical fixtures: row orthonormality and the pair-rotation invariants
 *   - or_vertical_*: brute-force probe columns and modal round trip
 *   - or_pointwise_advection: the 37-point known-answer test
 *     (tests/skbench_tests.cpp:1136-1168)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.141592653589793238462643383279502884

typedef struct {
    double re, im;
} or_complex;

/* ---- sizes (src/model.cpp:112-127) ------------------------------------ */
static uint64_t nx_half(uint64_t nx) { return nx / 2 + 1; }

uint64_t or_retained_vertical_modes(uint64_t nz) {
    if (nz < 2) return 0;
    return 2 * (nz - 1) / 3; /* src/model.cpp:124-127 */
}

/* ---- index maps (src/model.cpp:219-244) ------------------------------- */
static uint64_t real_index(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t x, uint64_t y, uint64_t z,
                           uint64_t field) {
    return x + nx * (y + ny * (z + nz * field));
}

static uint64_t wvm_index(uint64_t nx, uint64_t nz, uint64_t fields, uint64_t kx, uint64_t ky, uint64_t z,
                          uint64_t field) {
    return z + nz * field + nz * fields * (kx + nx_half(nx) * ky);
}

static uint64_t retained_index(uint64_t nz, uint64_t fields, uint64_t mode, uint64_t z, uint64_t field) {
    return z + nz * field + nz * fields * mode;
}

static uint64_t modal_index(uint64_t nj, uint64_t fields, uint64_t mode, uint64_t j, uint64_t field) {
    return j + nj * field + nj * fields * mode;
}

/* ---- DFT mode lists (src/model.cpp:34-65) ----------------------------- */
static int64_t dft_mode(uint64_t index, uint64_t count) {
    /* Position `index` of the list 0..ceil(n/2)-1, -(n/2)..-1 (src/model.cpp:34-45). */
    const uint64_t positive = (count + 1) / 2;
    if (index < positive) return (int64_t)index;
    return -(int64_t)(count / 2 - (index - positive));
}

static int self_conjugate(int64_t mode, uint64_t count) {
    return mode == 0 || (count % 2 == 0 && mode == -(int64_t)(count / 2));
}

static int primary_mode(int64_t k, int64_t l, uint64_t nx, uint64_t ny) {
    const int ks = self_conjugate(k, nx);
    const int ls = self_conjugate(l, ny);
    return l > 0 || (ls && (k > 0 || ks));
}

static int nyquist_mode(int64_t k, int64_t l, uint64_t nx, uint64_t ny) {
    return (nx % 2 == 0 && k == -(int64_t)(nx / 2)) || (ny % 2 == 0 && l == -(int64_t)(ny / 2));
}

static double radial_magnitude(double a, double b) {
    /* src/model.cpp:67-71: each square is rounded on its own. */
    volatile double a2 = a * a;
    volatile double b2 = b * b;
    return sqrt(a2 + b2);
}

typedef struct {
    int64_t k, l;
    uint64_t stored_kx, stored_ky;
    int32_t conj;
    int32_t pad;
    double radial;
} or_mode;

static int mode_less(const or_mode* a, const or_mode* b) {
    if (a->radial != b->radial) return a->radial < b->radial;
    if (a->k != b->k) return a->k < b->k;
    return a->l < b->l;
}

/* Stable merge sort keyed on (radial, k, l) (src/model.cpp:161-165). */
static void merge_sort(or_mode* a, or_mode* tmp, uint64_t n) {
    if (n < 2) return;
    const uint64_t h = n / 2;
    merge_sort(a, tmp, h);
    merge_sort(a + h, tmp, n - h);
    uint64_t i = 0, j = h, o = 0;
    while (i < h && j < n) {
        if (mode_less(&a[j], &a[i])) tmp[o++] = a[j++];
        else tmp[o++] = a[i++];
    }
    while (i < h) tmp[o++] = a[i++];
    while (j < n) tmp[o++] = a[j++];
    memcpy(a, tmp, n * sizeof(or_mode));
}

/* retainedHorizontalModes (src/model.cpp:129-167).  Returns the number of
 * modes, writes at most `capacity` records when out != NULL. */
int64_t or_retained_modes(uint64_t nx, uint64_t ny, double lx, double ly, int antialias, or_mode* out,
                          uint64_t capacity) {
    if (nx < 2 || ny < 2 || !(lx > 0.0) || !(ly > 0.0)) return -1;
    const double maximum_k = 2.0 * OR_PI * ((double)(nx / 2) / lx);
    or_mode* modes = (or_mode*)malloc(sizeof(or_mode) * (nx_half(nx) * ny + 1));
    if (!modes) return -2;
    uint64_t count = 0;
    for (uint64_t li = 0; li < ny; ++li) {
        const int64_t l = dft_mode(li, ny);
        for (uint64_t ki = 0; ki < nx; ++ki) {
            const int64_t k = dft_mode(ki, nx);
            if (!primary_mode(k, l, nx, ny) || nyquist_mode(k, l, nx, ny)) continue;
            const double pk = 2.0 * OR_PI * (double)k / lx;
            const double pl = 2.0 * OR_PI * (double)l / ly;
            const double radial = radial_magnitude(pk, pl);
            if (antialias && radial > 2.0 * maximum_k / 3.0) continue;
            or_mode m;
            memset(&m, 0, sizeof m);
            m.k = k;
            m.l = l;
            m.radial = radial;
            if (k >= 0) {
                m.stored_kx = (uint64_t)k;
                m.stored_ky = l >= 0 ? (uint64_t)l : (uint64_t)((int64_t)ny + l);
            } else {
                m.stored_kx = (uint64_t)(-k);
                const int64_t nl = -l;
                m.stored_ky = nl >= 0 ? (uint64_t)nl : (uint64_t)((int64_t)ny + nl);
                m.conj = 1;
            }
            modes[count++] = m;
        }
    }
    or_mode* tmp = (or_mode*)malloc(sizeof(or_mode) * (count + 1));
    if (!tmp) {
        free(modes);
        return -2;
    }
    merge_sort(modes, tmp, count);
    free(tmp);
    if (out) memcpy(out, modes, sizeof(or_mode) * (count < capacity ? count : capacity));
    free(modes);
    return (int64_t)count;
}

/* ---- FNV-1a 64 (src/model.cpp:86-104) --------------------------------- */
static void fnv_u64(uint64_t* h, uint64_t v) {
    for (int b = 0; b < 8; ++b) {
        *h ^= (v >> (8 * b)) & 0xffu;
        *h *= UINT64_C(1099511628211);
    }
}
#define FNV_OFFSET UINT64_C(14695981039346656037)

/* modeOrderHash (src/model.cpp:462-469) */
uint64_t or_mode_order_hash(const or_mode* modes, uint64_t n) {
    uint64_t h = FNV_OFFSET;
    for (uint64_t i = 0; i < n; ++i) {
        fnv_u64(&h, (uint64_t)modes[i].k);
        fnv_u64(&h, (uint64_t)modes[i].l);
    }
    return h;
}

/* wvmSpectrumOrderHash (src/model.cpp:471-487) */
uint64_t or_wvm_order_hash(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields) {
    uint64_t h = FNV_OFFSET;
    for (uint64_t ky = 0; ky < ny; ++ky) {
        const int64_t sky = ky <= ny / 2 ? (int64_t)ky : (int64_t)ky - (int64_t)ny;
        for (uint64_t kx = 0; kx < nx_half(nx); ++kx)
            for (uint64_t f = 0; f < fields; ++f)
                for (uint64_t z = 0; z < nz; ++z) {
                    fnv_u64(&h, kx);
                    fnv_u64(&h, (uint64_t)sky);
                    fnv_u64(&h, z);
                    fnv_u64(&h, f);
                }
    }
    return h;
}

/* ---- K^2 grouping (src/model.cpp:169-217) ----------------------------- */
static uint64_t abs_u64(int64_t v) { return v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v; }

uint64_t or_squared_key(int64_t k, int64_t l) {
    const uint64_t a = abs_u64(k), b = abs_u64(l);
    return a * a + b * b;
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Returns the group count, or -1 when equal keys are not contiguous (the
 * reference throws std::invalid_argument there, src/model.cpp:196-199).
 * key/first/count must each hold at least n entries. */
int64_t or_k2_groups(const or_mode* modes, uint64_t n, uint64_t* key, uint64_t* first, uint64_t* count) {
    if (n == 0) return -1;
    int64_t groups = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t k2 = or_squared_key(modes[i].k, modes[i].l);
        if (groups == 0 || key[groups - 1] != k2) {
            key[groups] = k2;
            first[groups] = i;
            count[groups] = 1;
            ++groups;
        } else {
            ++count[groups - 1];
        }
    }
    /* a completed key must never start a second run */
    uint64_t* sorted = (uint64_t*)malloc(sizeof(uint64_t) * (uint64_t)groups);
    if (!sorted) return -2;
    memcpy(sorted, key, sizeof(uint64_t) * (uint64_t)groups);
    qsort(sorted, (size_t)groups, sizeof(uint64_t), cmp_u64);
    for (int64_t g = 1; g < groups; ++g)
        if (sorted[g] == sorted[g - 1]) {
            free(sorted);
            return -1;
        }
    free(sorted);
    return groups;
}

/* verticalModeGroupHash (src/model.cpp:209-217) */
uint64_t or_group_hash(const uint64_t* key, const uint64_t* first, const uint64_t* count, uint64_t n) {
    uint64_t h = FNV_OFFSET;
    for (uint64_t g = 0; g < n; ++g) {
        fnv_u64(&h, key[g]);
        fnv_u64(&h, first[g]);
        fnv_u64(&h, count[g]);
    }
    return h;
}

/* ---- vertical fixtures (src/model.cpp:489-565) ------------------------ */
void or_orthonormal_vertical(uint64_t nz, uint64_t nj, double* forward /* nj*nz */, double* inverse /* nz*nj */) {
    for (uint64_t j = 0; j < nj; ++j) {
        const double norm = j == 0 ? sqrt(1.0 / (double)nz) : sqrt(2.0 / (double)nz);
        for (uint64_t z = 0; z < nz; ++z) {
            const double v = norm * cos(OR_PI * ((double)z + 0.5) * (double)j / (double)nz);
            forward[j * nz + z] = v;
            inverse[z * nj + j] = v;
        }
    }
}

/* squaredWavenumberVerticalFixture: one pair-rotated copy of the truncated
 * DCT-II per K^2 group (src/model.cpp:523-565). */
void or_k2_vertical_family(uint64_t nz, uint64_t nj, uint64_t ngroups, const uint64_t* keys, double* forward,
                           double* inverse) {
    double* bf = (double*)malloc(sizeof(double) * nj * nz);
    double* bi = (double*)malloc(sizeof(double) * nj * nz);
    or_orthonormal_vertical(nz, nj, bf, bi);
    const uint64_t me = nj * nz;
    for (uint64_t g = 0; g < ngroups; ++g) {
        const uint64_t key = keys[g];
        const double phase = (double)((key % UINT64_C(4093)) + 1) / 4094.0;
        const double c = cos(0.5 * OR_PI * phase);
        const double s = sin(0.5 * OR_PI * phase);
        const uint64_t off = g * me;
        for (uint64_t j = 0; j < nj; j += 2) {
            for (uint64_t z = 0; z < nz; ++z) {
                const double a = bf[j * nz + z];
                double ar = a;
                if (j + 1 < nj) {
                    const double b = bf[(j + 1) * nz + z];
                    ar = c * a + s * b;
                    const double br = -s * a + c * b;
                    forward[off + (j + 1) * nz + z] = br;
                    inverse[off + z * nj + j + 1] = br;
                } else if (key & 1u) {
                    ar = -a;
                }
                forward[off + j * nz + z] = ar;
                inverse[off + z * nj + j] = ar;
            }
        }
    }
    free(bf);
    free(bi);
}

/* ---- portable generator (src/benchmark.cpp:187-196) -------------------
 * xorshift64 followed by a multiplicative scramble, 53-bit mantissa mapped to
 * [-1, 1).  Used for the "random" real fixture instead of the reference's
 * std::normal_distribution, whose values are library-defined. */
double or_signed_unit(uint64_t* state) {
    *state ^= *state >> 12;
    *state ^= *state << 25;
    *state ^= *state >> 27;
    const uint64_t v = *state * UINT64_C(2685821657736338717);
    const uint64_t mant = v >> 11;
    return 2.0 * (double)mant / (double)(UINT64_C(1) << 53) - 1.0;
}

static uint64_t seed_state(uint64_t seed) { return seed == 0 ? UINT64_C(0x9e3779b97f4a7c15) : seed; }

/* ---- real fixtures (src/model.cpp:618-661) ----------------------------
 * kind: 0 impulse, 1 sinusoid (k,l)=(1,2), 2 random (uniform[-1,1) from the
 * portable generator), 3 dc, 4 nyquist, 5 sinusoid (k,l)=(5,17). */
int or_make_fixture(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields, int kind, uint64_t seed,
                    double* out) {
    const uint64_t total = nx * ny * nz * fields;
    memset(out, 0, sizeof(double) * total);
    uint64_t state = seed_state(seed);
    for (uint64_t f = 0; f < fields; ++f)
        for (uint64_t z = 0; z < nz; ++z) {
            const uint64_t plane = z + nz * f;
            const double amp = 1.0 + 0.125 * (double)plane;
            if (kind == 0) {
                const uint64_t x = (1 + plane) % nx;
                const uint64_t y = (2 + 3 * plane) % ny;
                out[real_index(nx, ny, nz, x, y, z, f)] = amp;
                continue;
            }
            for (uint64_t y = 0; y < ny; ++y)
                for (uint64_t x = 0; x < nx; ++x) {
                    double v = 0.0;
                    if (kind == 1) {
                        const double ph = 0.17 * (double)plane;
                        v = amp * cos(2.0 * OR_PI * ((double)x / (double)nx + 2.0 * (double)y / (double)ny) + ph);
                    } else if (kind == 5) {
                        const double ph = 0.17 * (double)plane;
                        v = amp * cos(2.0 * OR_PI * (5.0 * (double)x / (double)nx + 17.0 * (double)y / (double)ny) + ph);
                    } else if (kind == 2) {
                        v = or_signed_unit(&state);
                    } else if (kind == 3) {
                        v = amp;
                    } else if (kind == 4) {
                        v = amp * (((x + y) % 2 == 0) ? 1.0 : -1.0);
                    } else {
                        return -1;
                    }
                    out[real_index(nx, ny, nz, x, y, z, f)] = v;
                }
        }
    return 0;
}

/* ---- direct DFT oracle (src/model.cpp:663-716) ------------------------ */
void or_direct_r2c(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields, const double* in, or_complex* wvm) {
    for (uint64_t f = 0; f < fields; ++f)
        for (uint64_t z = 0; z < nz; ++z)
            for (uint64_t ky = 0; ky < ny; ++ky)
                for (uint64_t kx = 0; kx < nx_half(nx); ++kx) {
                    double sr = 0.0, si = 0.0;
                    for (uint64_t y = 0; y < ny; ++y)
                        for (uint64_t x = 0; x < nx; ++x) {
                            const double ang = -2.0 * OR_PI *
                                ((double)(kx * x) / (double)nx + (double)(ky * y) / (double)ny);
                            const double v = in[real_index(nx, ny, nz, x, y, z, f)];
                            sr += v * cos(ang);
                            si += v * sin(ang);
                        }
                    or_complex c = {sr, si};
                    wvm[wvm_index(nx, nz, fields, kx, ky, z, f)] = c;
                }
}

void or_direct_c2r(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields, const or_complex* wvm, double* out) {
    for (uint64_t f = 0; f < fields; ++f)
        for (uint64_t z = 0; z < nz; ++z)
            for (uint64_t y = 0; y < ny; ++y)
                for (uint64_t x = 0; x < nx; ++x) {
                    double sr = 0.0;
                    for (uint64_t ky = 0; ky < ny; ++ky)
                        for (uint64_t kx = 0; kx < nx; ++kx) {
                            uint64_t skx = kx, sky = ky;
                            int cj = 0;
                            if (kx > nx / 2) {
                                skx = nx - kx;
                                sky = (ny - ky) % ny;
                                cj = 1;
                            }
                            or_complex c = wvm[wvm_index(nx, nz, fields, skx, sky, z, f)];
                            if (cj) c.im = -c.im;
                            const double ang = 2.0 * OR_PI *
                                ((double)(kx * x) / (double)nx + (double)(ky * y) / (double)ny);
                            /* real part of c * e^{i ang} */
                            sr += c.re * cos(ang) - c.im * sin(ang);
                        }
                    out[real_index(nx, ny, nz, x, y, z, f)] = sr;
                }
}

/* directRetained (src/benchmark.cpp:135-156): DFT evaluated at the retained
 * (k,l) directly, an oracle that bypasses the half-spectrum storage. */
void or_direct_retained(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields, const or_mode* modes,
                        uint64_t nmodes, const double* in, or_complex* retained) {
    for (uint64_t m = 0; m < nmodes; ++m)
        for (uint64_t f = 0; f < fields; ++f)
            for (uint64_t z = 0; z < nz; ++z) {
                double sr = 0.0, si = 0.0;
                for (uint64_t y = 0; y < ny; ++y)
                    for (uint64_t x = 0; x < nx; ++x) {
                        const double ang = -2.0 * OR_PI *
                            ((double)modes[m].k * (double)x / (double)nx + (double)modes[m].l * (double)y / (double)ny);
                        const double v = in[real_index(nx, ny, nz, x, y, z, f)];
                        sr += v * cos(ang);
                        si += v * sin(ang);
                    }
                or_complex c = {sr, si};
                retained[retained_index(nz, fields, m, z, f)] = c;
            }
}

/* ---- retained gather / embed (src/model.cpp:310-339) ------------------ */
void or_gather_retained(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields, const or_mode* modes,
                        uint64_t nmodes, const or_complex* full, or_complex* retained) {
    (void)ny;
    for (uint64_t m = 0; m < nmodes; ++m)
        for (uint64_t f = 0; f < fields; ++f)
            for (uint64_t z = 0; z < nz; ++z) {
                or_complex v = full[wvm_index(nx, nz, fields, modes[m].stored_kx, modes[m].stored_ky, z, f)];
                if (modes[m].conj) v.im = -v.im;
                retained[retained_index(nz, fields, m, z, f)] = v;
            }
}

void or_embed_retained(uint64_t nx, uint64_t ny, uint64_t nz, uint64_t fields, const or_mode* modes,
                       uint64_t nmodes, const or_complex* retained, or_complex* full) {
    memset(full, 0, sizeof(or_complex) * nx_half(nx) * ny * nz * fields);
    for (uint64_t m = 0; m < nmodes; ++m)
        for (uint64_t f = 0; f < fields; ++f)
            for (uint64_t z = 0; z < nz; ++z) {
                or_complex v = retained[retained_index(nz, fields, m, z, f)];
                if (modes[m].conj) v.im = -v.im;
                full[wvm_index(nx, nz, fields, modes[m].stored_kx, modes[m].stored_ky, z, f)] = v;
                if (modes[m].stored_kx == 0 && modes[m].stored_ky != 0 && 2 * modes[m].stored_ky != ny) {
                    const uint64_t cky = (ny - modes[m].stored_ky) % ny;
                    or_complex c = {v.re, -v.im};
                    full[wvm_index(nx, nz, fields, 0, cky, z, f)] = c;
                }
            }
}

/* ---- grouped vertical products (src/model.cpp:567-605; grouped form
 * src/benchmark.cpp:229-273) --------------------------------------------
 * Group g covers modes [first[g], first[g]+count[g]) and owns the matrix at
 * forward + g*nj*nz (row-major nj x nz) / inverse + g*nz*nj (nz x nj). */
void or_vertical_forward(uint64_t nz, uint64_t nj, uint64_t fields, uint64_t ngroups, const uint64_t* first,
                         const uint64_t* count, const double* forward, const or_complex* physical,
                         or_complex* modal) {
    for (uint64_t g = 0; g < ngroups; ++g) {
        const double* a = forward + g * nj * nz;
        for (uint64_t m = first[g]; m < first[g] + count[g]; ++m)
            for (uint64_t f = 0; f < fields; ++f)
                for (uint64_t j = 0; j < nj; ++j) {
                    double sr = 0.0, si = 0.0;
                    for (uint64_t z = 0; z < nz; ++z) {
                        const or_complex v = physical[retained_index(nz, fields, m, z, f)];
                        sr += a[j * nz + z] * v.re;
                        si += a[j * nz + z] * v.im;
                    }
                    or_complex c = {sr, si};
                    modal[modal_index(nj, fields, m, j, f)] = c;
                }
    }
}

void or_vertical_inverse(uint64_t nz, uint64_t nj, uint64_t fields, uint64_t ngroups, const uint64_t* first,
                         const uint64_t* count, const double* inverse, const or_complex* modal,
                         or_complex* physical) {
    for (uint64_t g = 0; g < ngroups; ++g) {
        const double* a = inverse + g * nz * nj;
        for (uint64_t m = first[g]; m < first[g] + count[g]; ++m)
            for (uint64_t f = 0; f < fields; ++f)
                for (uint64_t z = 0; z < nz; ++z) {
                    double sr = 0.0, si = 0.0;
                    for (uint64_t j = 0; j < nj; ++j) {
                        const or_complex v = modal[modal_index(nj, fields, m, j, f)];
                        sr += a[z * nj + j] * v.re;
                        si += a[z * nj + j] * v.im;
                    }
                    or_complex c = {sr, si};
                    physical[retained_index(nz, fields, m, z, f)] = c;
                }
    }
}

/* ---- pointwise advection (src/pointwise_advection.cpp:152-166) --------
 * target = -scale * (u*qx + v*qy + w*qz); shared = [u|v|w], derivative =
 * [qx|qy|qz], each block `volume` long. */
void or_pointwise_advection(uint64_t volume, double scale, const double* shared, const double* derivative,
                            double* target) {
    for (uint64_t i = 0; i < volume; ++i) {
        const double s = shared[i] * derivative[i] + shared[volume + i] * derivative[volume + i] +
            shared[2 * volume + i] * derivative[2 * volume + i];
        target[i] = -scale * s;
    }
}

/* ---- error metrics (src/model.cpp:718-751) ---------------------------- */
double or_max_relative_error_complex(const or_complex* a, const or_complex* b, uint64_t n) {
    double num = 0.0, den = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double d = hypot(a[i].re - b[i].re, a[i].im - b[i].im);
        const double m = hypot(b[i].re, b[i].im);
        if (d > num) num = d;
        if (m > den) den = m;
    }
    return num / (den > 1.0 ? den : 1.0);
}

double or_max_relative_error_real(const double* a, const double* b, uint64_t n, double actual_scale) {
    double num = 0.0, den = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double d = fabs(actual_scale * a[i] - b[i]);
        const double m = fabs(b[i]);
        if (d > num) num = d;
        if (m > den) den = m;
    }
    return num / (den > 1.0 ? den : 1.0);
}

double or_relative_l2_error_complex(const or_complex* a, const or_complex* b, uint64_t n) {
    long double e = 0.0L, r = 0.0L;
    for (uint64_t i = 0; i < n; ++i) {
        const long double dr = (long double)a[i].re - b[i].re;
        const long double di = (long double)a[i].im - b[i].im;
        e += dr * dr + di * di;
        r += (long double)b[i].re * b[i].re + (long double)b[i].im * b[i].im;
    }
    if (r == 0.0L) return sqrt((double)e);
    return sqrt((double)(e / r));
}
