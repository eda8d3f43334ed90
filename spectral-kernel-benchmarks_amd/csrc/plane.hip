// Plane-resident 2-D transforms of the retained-mode operator (FP32) and the
// compact <-> plane-major layout transposes, for gfx950.
//
// The pruned inverse c2r and forward r2c of one N x N plane (N = 256, active
// k_x band kc = 86, the 2/3-rule retained set) run in ONE workgroup of 1024
// threads (16 wavefronts).  The plane's band of kc x N complex values never
// leaves the CU between the column and the row transforms: the column
// results stay in registers, move to rows through LDS, and the exchange
// between the two register passes of every 256-point FFT stays inside a
// wavefront.  HBM sees the retained values once and the real plane once per
// direction; the kc x Ny x P intermediate of the two-kernel transforms is
// gone.
//
// Retained values are read and written in the CHUNKED PLANE-MAJOR layout:
// modes in chunks of 16, mode m of plane p at 16*P*(m / 16) + 16*p + m % 16
// (P planes; the last chunk padded to 16 modes, np = modes rounded up to 16).
// The layout holds SLOT values: the value of the half-spectrum slot a mode is
// stored at, i.e. the conjugate of the caller's value for modes stored
// conjugated (k < 0).  The moves to and from the caller's compact layout
// apply that conjugation (PlaneTables::cmask), so the transforms need none:
// a mode's slot takes the value as it is, a k_x = 0 Hermitian partner its
// conjugate.
// A plane's 16 modes of a chunk are one 128-byte run, and the 16*P values of
// a chunk are the same bytes as the compact layout's (plane + P*mode) 16
// modes, transposed: pm_transpose_kernel moves each chunk through LDS with
// contiguous reads and writes, and the vertical product reads a tile's
// 16 x Nz block of one field as one contiguous run.
//
// Every barrier in these kernels orders LDS only (lds_barrier): the global
// loads and stores of a plane are never read by another wavefront (nor read
// back by the fused operator, whose forward starts from the inverse's row
// registers), so the stores of one row half keep draining while the next
// half is transformed.
//
// Thread roles in a plane workgroup (tid = 64*wave + 16*g + n):
//   columns  column c = 64*r + 4*wave + g (round r), lane n of its 16;
//   rows     row pair (2f, 2f + 1) of a half of 128 rows, f = 4*wave + g,
//            packed as one complex sequence (real rows as Re and Im).
// A 256-point FFT is N1 x N2 = 16 x 16: lane n loads x[16 j + n] (j = 0..15),
// does a 16-point DFT over j, multiplies by W^(n k1), exchanges with the 15
// other lanes of its group through the wavefront's LDS scratch and does the
// second 16-point DFT, ending with X[n + 16 k] (k = 0..15) in lane n.  The
// 16-point DFTs are rdft16_fma: their inner twiddles ride in the butterflies'
// FMAs, 16 fewer operations per DFT than products would take.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "fft_core.hpp"
#include "internal.hpp"
#include "plane.hpp"

namespace skb {

namespace {

constexpr int kPlaneN = 256;
constexpr int kPlaneThreads = 1024;
constexpr int kTwiddlePitch = 18;  // complex entries per lane's twiddle row in LDS (16 used)


// Makes the LDS accesses of one wavefront before it visible to the lanes of the
// same wavefront after it (LDS operations of a wavefront complete in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// In-place exchange between the two register passes of the four 256-point
// FFTs of a wavefront: lane n of group g holds v[k] (k = 0..15) and ends with
// v[m] = (value k = n of lane m).  Real parts, then imaginary parts, through
// the wavefront's scratch of 4 x 272 floats: element (k, m) of group g sits at
// 272 g + 17 k + m, so a lane addresses all its elements from one base with
// immediate offsets, and the writes (fixed k) and reads (fixed m) of a
// half-wavefront (two groups, 272 = 16 mod 32 apart) hit 32 distinct banks.
constexpr int kExchangeFloats = 4 * 272;
__device__ __forceinline__ void wave_exchange16(float* s, cx<float>* v, int g, int n) {
    float* wr = s + 272 * g + n;
    const float* rd = s + 272 * g + 17 * n;
#pragma unroll
    for (int k = 0; k < 16; ++k) wr[17 * k] = v[k].x;
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m].x = rd[m];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) wr[17 * k] = v[k].y;
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m].y = rd[m];
    wave_sync();
}

template <int KC>
struct PlaneGeo {
    static constexpr int N = kPlaneN, H = N / 2;
    static constexpr int ROUNDS = (KC + 63) / 64;  // column rounds of 64 columns
    static constexpr int RP = KC | 1;              // row-buffer pitch (complex, odd: column writes spread over banks)
    static constexpr int YP = N + 1;               // band-buffer pitch (floats, odd)
    static constexpr int KH = (KC + 15) / 16;      // registers k of a row FFT holding kx = n + 16k < KC
    // Registers j of a column FFT whose k_y = 16 j + n (n < 16) lie in the
    // band |k_y| >= KC that a 2/3-rule set never retains: [16 J0, 16 J1 + 15]
    // inside (KC - 1, N - KC + 1).  Plans use the plane path only when their
    // mode set leaves these slots empty (plane_band_empty), so the inverse
    // columns take them as zero and the forward columns skip them.
    static constexpr int J0 = (KC + 15) / 16, J1 = (N - KC + 1 - 16) / 16;
    static constexpr bool empty(int j) { return j >= J0 && j <= J1; }
    static constexpr int SMAX = kPlaneMaxModes;    // retained values staged per plane
    static constexpr int IT = (SMAX / 2 + kPlaneThreads - 1) / kPlaneThreads;  // 16-byte input loads per thread
    static constexpr int Z = SMAX;                 // staged slot holding zero (inverse) / trash (forward)
    // (rounded to 16 bytes: the twiddle rows after the exchange scratch are read 16 bytes at a time)
    static constexpr int R0 = (std::max(std::max(H * RP * 8, KC * YP * 4), (SMAX + 1) * 8) + 15) / 16 * 16;
    static constexpr int WT = 16 * kExchangeFloats * 4;  // one exchange scratch per wavefront
    static constexpr int BYTES = R0 + WT + 16 * kTwiddlePitch * 8;
    static_assert(KC <= 128 && BYTES <= 160 * 1024, "plane workgroup exceeds the CU's LDS");
};

// One 256-point FFT per 16-lane group: lane n holds x[16 j + n] in v[j]; on
// return it holds X[n + 16 k] in v[k].  Lane n's twiddles exp(-2 pi i n k /
// 256), k = 0..15, are one LDS row tw[kTwiddlePitch * n + k], read 16 bytes
// (two twiddles) at a time; rows 144 bytes apart put the 16 lanes of a
// group on distinct banks.  INV uses the conjugate (unnormalised inverse).
template <bool INV>
__device__ __forceinline__ void fft256(cx<float>* v, float* s, const cx<float>* tw, int g, int n) {
    rdft16_fma<INV>(v);
    const bf4* row = reinterpret_cast<const bf4*>(tw + kTwiddlePitch * n);
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const bf4 w = row[h];
        if (h) v[2 * h] = INV ? cmulc(v[2 * h], cx<float>{w.x, w.y}) : cmul(v[2 * h], cx<float>{w.x, w.y});
        v[2 * h + 1] = INV ? cmulc(v[2 * h + 1], cx<float>{w.z, w.w}) : cmul(v[2 * h + 1], cx<float>{w.z, w.w});
    }
    wave_exchange16(s, v, g, n);
    rdft16_fma<INV>(v);
}

// Two independent 256-point FFTs per 16-lane group (the two column rounds of
// one wavefront), interleaved: a's exchange drains while b's first pass
// runs, and a's second pass overlaps b's exchange.  The exchanges share the
// wavefront's scratch; LDS operations of a wavefront complete in order, so
// b's writes follow a's reads.
template <bool INV>
__device__ __forceinline__ void fft256x2(cx<float>* a, cx<float>* b, float* s, const cx<float>* tw, int g, int n) {
    const bf4* row = reinterpret_cast<const bf4*>(tw + kTwiddlePitch * n);
    rdft16_fma<INV>(a);
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const bf4 w = row[h];
        if (h) a[2 * h] = INV ? cmulc(a[2 * h], cx<float>{w.x, w.y}) : cmul(a[2 * h], cx<float>{w.x, w.y});
        a[2 * h + 1] = INV ? cmulc(a[2 * h + 1], cx<float>{w.z, w.w}) : cmul(a[2 * h + 1], cx<float>{w.z, w.w});
    }
    wave_exchange16(s, a, g, n);
    rdft16_fma<INV>(b);
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const bf4 w = row[h];
        if (h) b[2 * h] = INV ? cmulc(b[2 * h], cx<float>{w.x, w.y}) : cmul(b[2 * h], cx<float>{w.x, w.y});
        b[2 * h + 1] = INV ? cmulc(b[2 * h + 1], cx<float>{w.z, w.w}) : cmul(b[2 * h + 1], cx<float>{w.z, w.w});
    }
    wave_exchange16(s, b, g, n);
    rdft16_fma<INV>(a);
    rdft16_fma<INV>(b);
}

// Lane i of each 16-lane row receives the value of lane (16 - i) mod 16
// (row_mirror: lane i <- 15 - i, then row_ror:1: lane i <- i - 1).
__device__ __forceinline__ float from_partner_lane(float v) {
    int x = __float_as_int(v);
    x = __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false);
    x = __builtin_amdgcn_update_dpp(0, x, 0x121, 0xF, 0xF, false);
    return __int_as_float(x);
}

// The code of register k (0..15) of a lane's column: the staged index of the
// slot's value (its mode, or kPlaneMaxModes, a staged zero, where the (kx,
// ky) slot holds no retained value), plus kPartner on the inverse's k_x = 0
// Hermitian partners, which take the conjugate.  Registers in the band no
// 2/3-rule set retains (PlaneGeo::empty) have no code: the others' 16-bit
// codes sit in order (code_slot), two per 32-bit word, in the first six
// words of the lane's 32-byte table entry, so the loads bring in no word
// that goes unused (an unused loaded register would be reused at once and
// stall on its load).
constexpr int kPartner = 0x8000, kCodeIndex = 0x7FFF;
static_assert(kPlaneMaxModes <= kCodeIndex, "staged indices need 15 bits");
template <int KC>
constexpr int code_slot(int k) {
    return k < PlaneGeo<KC>::J0 ? k : k - (PlaneGeo<KC>::J1 - PlaneGeo<KC>::J0 + 1);
}
static_assert(16 - (PlaneGeo<kPlaneKc>::J1 - PlaneGeo<kPlaneKc>::J0 + 1) <= 12, "codes take six words");

template <int KC>
__device__ __forceinline__ int unpack_code(const uint32_t* c, int k) {
    const int sl = code_slot<KC>(k);
    const uint32_t w = c[sl >> 1];
    return static_cast<int>((sl & 1) ? (w >> 16) : (w & 0xFFFFu));
}

__device__ __forceinline__ void load_codes(uint32_t* c, const uint32_t* __restrict__ table, int lane_slot) {
    const bf4* src = reinterpret_cast<const bf4*>(table) + 2 * lane_slot;
    const bf4 a = src[0];
    const bf2 b = *reinterpret_cast<const bf2*>(src + 1);
    c[0] = __float_as_uint(a.x);
    c[1] = __float_as_uint(a.y);
    c[2] = __float_as_uint(a.z);
    c[3] = __float_as_uint(a.w);
    c[4] = __float_as_uint(b.x);
    c[5] = __float_as_uint(b.y);
}

}  // namespace

// ---------------------------------------------------------------------------
// Inverse: plane-major retained values of plane blockIdx.x -> its real plane.
// icodes: per (column c, lane n) 16 codes of (kx = c, ky = 16 j + n), the
// k_x = 0 Hermitian partners included (src/model.cpp:332-335).  The imaginary
// part of the kx = 0 bin is ignored by the row c2r, as FFTW's c2r does.
// ---------------------------------------------------------------------------
// FWD (the fused operator): right after a half's real rows are stored, the
// forward row transforms of the same row pairs run on the registers that hold
// them, so the forward never reads the real plane back; X_a, X_b of both
// halves end in xa, xb.
// Thread indexing of the plane functions: `wbase` = 64 * wavefront index (a
// scalar), plane_lane() the lane, recomputed from the execution mask where it
// is used.  The lane is opaque to the compiler, so the addresses derived from
// it are computed in the phase that uses them instead of being held in
// registers across the persistent operator's loop over planes.
__device__ __forceinline__ int plane_lane() {
    uint32_t z = 0;
    asm volatile("" : "+s"(z));
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, z)));
}

__device__ __forceinline__ int plane_wbase() {
    return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x)) & ~63;
}

// Byte offset of the 16-byte pair `tid` (modes 2 tid, 2 tid + 1) of plane p in
// the chunked plane-major block, and the step to pair tid + 1024 (128 chunks
// on).  tid < 1024, so the chunk index fits the 24-bit multiply.
__device__ __forceinline__ uint32_t pm_pair_offset(int tid, int planes, int p) {
    return (__umul24(static_cast<uint32_t>(tid) >> 3, 16u * static_cast<uint32_t>(planes)) + 16u * p +
            2u * (static_cast<uint32_t>(tid) & 7u)) * 8u;
}
__device__ __forceinline__ uint32_t pm_pair_step(int planes) {
    return (kPlaneThreads / 8) * 16u * 8u * static_cast<uint32_t>(planes);
}

// The plane's retained values as 16-byte pairs of its chunk lines (past an
// odd count a pair reads the chunk's padding), into registers: issued ahead
// of their use so that the plane waits for one memory latency, or none when
// the loads run under the previous plane's work.  live == false reads
// nothing (an empty buffer range: every load returns zero) but still defines
// v, so the previous plane's values need not stay in registers.
template <int KC>
__device__ __forceinline__ void load_plane_values(int wbase, bf4 (&v)[PlaneGeo<KC>::IT], const cx<float>* pm, int modes,
                                                  int np, int planes, int p, bool live = true) {
    const int tid = wbase + plane_lane();
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(pm, live ? static_cast<uint32_t>(np) * planes * 8u : 0u);
    // pair i = tid + 1024 k sits in chunk i / 8 at 8 (16 P (i / 8) + 16 p + 2 (i % 8)); a pair
    // past the last one reads the last chunk's padding or, in a later chunk,
    // past the buffer (zero), and is never staged
    const uint32_t at = pm_pair_offset(tid, planes, p), step = pm_pair_step(planes);
#pragma unroll
    for (int k = 0; k < PlaneGeo<KC>::IT; ++k) v[k] = raw_buffer_load_f4(rs, static_cast<int>(at + k * step), 0, 0);
}

// The twiddle table into its LDS region (read by every FFT of the workgroup;
// nothing else writes the region).
template <int KC>
__device__ __forceinline__ void stage_twiddles(int wbase, unsigned char* lds, const cx<float>* __restrict__ twg) {
    using G = PlaneGeo<KC>;
    const int tid = wbase + plane_lane();
    cx<float>* tw = reinterpret_cast<cx<float>*>(lds + G::R0 + G::WT);
    if (tid < G::N) tw[kTwiddlePitch * (tid & 15) + (tid >> 4)] = twg[tid];
}

// The 16 packed codes of each of this lane's column rounds (the same for
// every plane), issued ahead of their use like the values.
template <int KC>
__device__ __forceinline__ void load_column_codes(int wbase, uint32_t (&code)[PlaneGeo<KC>::ROUNDS][8],
                                                  const uint32_t* __restrict__ table) {
    const int lane = plane_lane(), wave = wbase >> 6, g = lane >> 4, n = lane & 15;
#pragma unroll
    for (int r = 0; r < PlaneGeo<KC>::ROUNDS; ++r) {
        const int c = 64 * r + 4 * wave + g;
        if (64 * r + 4 * wave < KC) load_codes(code[r], table, (c < KC ? c : KC - 1) * 16 + n);
    }
}

template <int KC, bool FWD>
__device__ __forceinline__ void inverse_plane(int wbase, unsigned char* lds, const bf4 (&v)[PlaneGeo<KC>::IT], float* real,
                                              const uint32_t* __restrict__ icodes, int modes, int p,
                                              cx<float> (&xa)[2][PlaneGeo<KC>::KH],
                                              cx<float> (&xb)[2][PlaneGeo<KC>::KH]) {
    using G = PlaneGeo<KC>;
    constexpr int N = G::N, RP = G::RP;
    cx<float>* r0 = reinterpret_cast<cx<float>*>(lds);  // staged modes, then the row buffer [128][RP]
    const int lane = plane_lane(), tid = wbase + lane, wave = wbase >> 6, g = lane >> 4, n = lane & 15;
    float* wt = reinterpret_cast<float*>(lds + G::R0) + kExchangeFloats * wave;
    cx<float>* tw = reinterpret_cast<cx<float>*>(lds + G::R0 + G::WT);

    // 1. the plane's retained values (loaded by the caller, load_plane_values)
    //    staged in LDS, with the zero slot after them; the caller has staged
    //    the twiddles (stage_twiddles) before this call
    {
        const int pairs = (modes + 1) >> 1;
#pragma unroll
        for (int k = 0; k < G::IT; ++k) {
            const int i = tid + kPlaneThreads * k;
            if (i < pairs) *reinterpret_cast<bf4*>(&r0[2 * i]) = v[k];
        }
        float zero = 0.f;
        asm volatile("" : "+v"(zero));  // made here, not held across the operator's loop
        if (tid == 0) r0[G::Z] = cx<float>{zero, zero};
    }
    const int uw = __builtin_amdgcn_readfirstlane(wave);  // wavefront index, uniform
    uint32_t code[G::ROUNDS][8];
    load_column_codes<KC>(wbase, code, icodes);
    lds_barrier();

    // 2. column inputs from the staged values
    cx<float> y[G::ROUNDS][16];
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
        const int c = 64 * r + 4 * wave + g;
        if (64 * r + 4 * wave < KC) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (G::empty(j)) {
                    y[r][j] = cx<float>{0.f, 0.f};
                    continue;
                }
                const int e = unpack_code<KC>(code[r], j);
                const cx<float> x = r0[e & kCodeIndex];
                // partners live in column 0 alone (wave 0, round 0)
                y[r][j] = (r == 0 && uw == 0 && (e & kPartner)) ? cx<float>{x.x, -x.y} : x;
            }
        }
    }
    lds_barrier();  // the staging area becomes the row buffer

    // 3. column transforms: lane n of column c ends with Y(c, n + 16 k) in
    //    y[r][k]; rows 0..127 (k < 8) go to the row buffer at once, rows
    //    128..255 stay in registers for the second half
    const int f = 4 * wave + g;
    static_assert(G::ROUNDS == 2, "two column rounds");
    if (64 + 4 * wave < KC) fft256x2<true>(y[0], y[1], wt, tw, g, n);
    else fft256<true>(y[0], wt, tw, g, n);
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
        const int c = 64 * r + 4 * wave + g;
        if (64 * r + 4 * wave < KC && c < KC) {
#pragma unroll
            for (int k = 0; k < 8; ++k) r0[(n + 16 * k) * RP + c] = y[r][k];
        }
    }

    // 4. rows, two halves of 128
    const __amdgpu_buffer_rsrc_t out =
        buffer_rsrc(real + static_cast<size_t>(p) * N * N, static_cast<uint32_t>(N * N * 4));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h) {
            lds_barrier();  // every row load of the first half is done
#pragma unroll
            for (int r = 0; r < G::ROUNDS; ++r) {
                const int c = 64 * r + 4 * wave + g;
                if (64 * r + 4 * wave < KC && c < KC) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) r0[(n + 16 * k) * RP + c] = y[r][8 + k];
                }
            }
        }
        lds_barrier();
        const cx<float>* ra = r0 + 2 * f * RP;
        const cx<float>* rb = ra + RP;
        cx<float> z[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            z[j] = cx<float>{0.f, 0.f};
            if (16 * j < KC) {
                // X_a[kx] + i X_b[kx] at kx = 16 j + n
                const int kx = 16 * j + n;
                const bool ok = kx < KC;
                cx<float> a = ra[ok ? kx : 0], b = rb[ok ? kx : 0];
                if (j == 0 && n == 0) a.y = b.y = 0.f;
                z[j] = ok ? cx<float>{a.x - b.y, a.y + b.x} : cx<float>{0.f, 0.f};
            } else if (j >= 8 && N - 16 * j - 15 < KC) {
                // conj X_a[N - kx] + i conj X_b[N - kx] at position 16 j + n
                const int kx = N - 16 * j - n;
                const bool ok = kx < KC;
                const cx<float> a = ra[ok ? kx : 0], b = rb[ok ? kx : 0];
                z[j] = ok ? cx<float>{a.x + b.y, b.x - a.y} : cx<float>{0.f, 0.f};
            }
        }
        fft256<true>(z, wt, tw, g, n);
        // lane n holds x = n + 16 k of row 2f (real parts) and 2f + 1 (imaginary)
        const uint32_t row = static_cast<uint32_t>((128 * h + 2 * f) * N + n) * 4u;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            buffer_store_real<float>(z[k].x, out, row + 64u * k);
            buffer_store_real<float>(z[k].y, out, row + N * 4u + 64u * k);
        }
        if constexpr (FWD) forward_row_pair<KC>(z, wt, tw, g, n, xa[h], xb[h]);
    }
}

template <int KC>
__global__ __launch_bounds__(kPlaneThreads) void plane_inverse_kernel(const cx<float>* __restrict__ pm,
                                                                     float* __restrict__ real,
                                                                     const cx<float>* __restrict__ twg,
                                                                     const uint32_t* __restrict__ icodes, int modes,
                                                                     int np, int planes) {
    __shared__ __align__(16) unsigned char lds[PlaneGeo<KC>::BYTES];
    cx<float> xa[2][PlaneGeo<KC>::KH], xb[2][PlaneGeo<KC>::KH];
    const int p = static_cast<int>(blockIdx.x);
    const int wbase = plane_wbase();
    bf4 v[PlaneGeo<KC>::IT];
    load_plane_values<KC>(wbase, v, pm, modes, np, planes, p);
    stage_twiddles<KC>(wbase, lds, twg);
    inverse_plane<KC, false>(wbase, lds, v, real, icodes, modes, p, xa, xb);
}

// ---------------------------------------------------------------------------
// Forward: real plane blockIdx.x -> its plane-major retained values.
// fcodes: per (column c, lane n) 16 codes of (kx = c, ky = n + 16 k).
// ---------------------------------------------------------------------------
// Row FFT of the packed row pair z (Re = row 2f, Im = row 2f + 1 of a half)
// and the Hermitian split into X_a, X_b at kx = n + 16 k < KC.
template <int KC>
__device__ __forceinline__ void forward_row_pair(cx<float>* z, float* wt, const cx<float>* tw, int g, int n,
                                                 cx<float>* xa, cx<float>* xb) {
    fft256<false>(z, wt, tw, g, n);
    // Z[n + 16 k]; its partner Z[N - n - 16 k] is register 15 - k of lane
    // (16 - n) mod 16, or register (16 - k) mod 16 of lane 0 itself
#pragma unroll
    for (int k = 0; k < PlaneGeo<KC>::KH; ++k) {
        const cx<float> pz = z[15 - k];
        cx<float> zn{from_partner_lane(pz.x), from_partner_lane(pz.y)};
        if (n == 0) zn = z[(16 - k) & 15];
        xa[k] = cx<float>{0.5f * (z[k].x + zn.x), 0.5f * (z[k].y - zn.y)};
        xb[k] = cx<float>{0.5f * (z[k].y + zn.y), 0.5f * (zn.x - z[k].x)};
    }
}

// The forward after its row transforms: the band (rows 2f, 2f + 1 of both
// halves at kx = n + 16 k, held by every lane) moves to columns through LDS,
// the column transforms run, and the retained outputs go to plane p of pm.
// The caller's LDS use of the band region [0, R0) must be over (barrier).
// `ahead` runs before the column transforms and `behind` once the outputs
// are staged, before the copy-out: the persistent operator kernel issues the
// next plane's value loads and its column codes there.
template <int KC, typename Ahead, typename Behind>
__device__ __forceinline__ void forward_columns(int wbase, unsigned char* lds, const cx<float> (&xa)[2][PlaneGeo<KC>::KH],
                                                const cx<float> (&xb)[2][PlaneGeo<KC>::KH], cx<float>* pm,
                                                const uint32_t* __restrict__ fcodes, int modes, int np, int planes,
                                                int p, Ahead&& ahead, Behind&& behind) {
    using G = PlaneGeo<KC>;
    constexpr int YP = G::YP, KH = G::KH;
    float* yb = reinterpret_cast<float*>(lds);           // band [KC][YP], real or imaginary parts
    cx<float>* st = reinterpret_cast<cx<float>*>(lds);  // then the output staging [modes]
    const int lane = plane_lane(), tid = wbase + lane, wave = wbase >> 6, g = lane >> 4, n = lane & 15;
    float* wt = reinterpret_cast<float*>(lds + G::R0) + kExchangeFloats * wave;
    const cx<float>* tw = reinterpret_cast<const cx<float>*>(lds + G::R0 + G::WT);
    const int f = 4 * wave + g;

    // 2. the band through LDS, real parts then imaginary parts; the column
    //    lanes gather their inputs after each.  The output codes are loaded
    //    between the parts (once the real parts are out of registers), so
    //    that they have arrived by the staging below.
    uint32_t code[G::ROUNDS][8];
    cx<float> cv[G::ROUNDS][16];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        if (part) {
            load_column_codes<KC>(wbase, code, fcodes);
            lds_barrier();  // every column lane has read the real parts
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                const int kx = n + 16 * k;
                if (kx < KC) {
                    float* col = yb + kx * YP + 128 * h + 2 * f;
                    col[0] = part ? xa[h][k].y : xa[h][k].x;
                    col[1] = part ? xb[h][k].y : xb[h][k].x;
                }
            }
        lds_barrier();
#pragma unroll
        for (int r = 0; r < G::ROUNDS; ++r) {
            const int c = 64 * r + 4 * wave + g;
            if (64 * r + 4 * wave < KC) {
                const float* col = yb + (c < KC ? c : KC - 1) * YP + n;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    if (part) cv[r][j].y = col[16 * j];
                    else cv[r][j].x = col[16 * j];
                }
            }
        }
    }
    lds_barrier();  // the band is consumed; the region becomes the output staging
    ahead();  // after the code loads: the staging below waits for those alone

    // 3. column transforms and the retained outputs
#pragma unroll
    for (int r = 0; r < G::ROUNDS; ++r) {
        const int c = 64 * r + 4 * wave + g;
        if (64 * r + 4 * wave < KC) {
            fft256<false>(cv[r], wt, tw, g, n);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                if (G::empty(k)) continue;  // no retained k_y there: these outputs are never computed
                // empty slots (and the columns past KC) write the trash slot
                st[c < KC ? unpack_code<KC>(code[r], k) : G::Z] = cv[r][k];
            }
        }
    }
    behind();
    lds_barrier();

    // 4. copy out (16-byte pairs of one chunk; the pair past an odd count
    //    writes the chunk's padding)
    {
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(pm, static_cast<uint32_t>(np) * planes * 8u);
        const int pairs = (modes + 1) >> 1;
        const uint32_t at = pm_pair_offset(tid, planes, p), step = pm_pair_step(planes);
#pragma unroll
        for (int k = 0; k < G::IT; ++k) {
            const int i = tid + kPlaneThreads * k;
            if (i < pairs)
                raw_buffer_store_f4(*reinterpret_cast<const bf4*>(&st[2 * i]), rs, static_cast<int>(at + k * step), 0, 0);
        }
    }
}

template <int KC>
__device__ __forceinline__ void forward_plane(int wbase, unsigned char* lds, const float* real, cx<float>* pm,
                                              const cx<float>* __restrict__ twg, const uint32_t* __restrict__ fcodes,
                                              int modes, int np, int planes, int p) {
    using G = PlaneGeo<KC>;
    constexpr int N = G::N, KH = G::KH;
    const int lane = plane_lane(), tid = wbase + lane, wave = wbase >> 6, g = lane >> 4, n = lane & 15;
    float* wt = reinterpret_cast<float*>(lds + G::R0) + kExchangeFloats * wave;
    cx<float>* tw = reinterpret_cast<cx<float>*>(lds + G::R0 + G::WT);
    const __amdgpu_buffer_rsrc_t in =
        buffer_rsrc(real + static_cast<size_t>(p) * N * N, static_cast<uint32_t>(N * N * 4));
    const int f = 4 * wave + g;
    if (tid < N) tw[kTwiddlePitch * (tid & 15) + (tid >> 4)] = twg[tid];

    // 1. rows: X_a, X_b (rows 2f, 2f + 1 of each half) for kx = n + 16 k < KC
    cx<float> xa[2][KH], xb[2][KH];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t row = static_cast<uint32_t>((128 * h + 2 * f) * N + n) * 4u;
        cx<float> z[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            z[j].x = raw_buffer_load_f1(in, static_cast<int>(row + 64u * j), 0, 0);
            z[j].y = raw_buffer_load_f1(in, static_cast<int>(row + N * 4u + 64u * j), 0, 0);
        }
        if (h == 0) lds_barrier();  // the twiddles are staged
        forward_row_pair<KC>(z, wt, tw, g, n, xa[h], xb[h]);
    }
    lds_barrier();  // (the band region held nothing else; kept for the caller contract)
    forward_columns<KC>(wbase, lds, xa, xb, pm, fcodes, modes, np, planes, p, [] {}, [] {});
}

template <int KC>
__global__ __launch_bounds__(kPlaneThreads) void plane_forward_kernel(const float* __restrict__ real,
                                                                     cx<float>* __restrict__ pm,
                                                                     const cx<float>* __restrict__ twg,
                                                                     const uint32_t* __restrict__ fcodes, int modes,
                                                                     int np, int planes) {
    __shared__ __align__(16) unsigned char lds[PlaneGeo<KC>::BYTES];
    forward_plane<KC>(plane_wbase(), lds, real, pm, twg, fcodes, modes, np, planes, static_cast<int>(blockIdx.x));
}

// ---------------------------------------------------------------------------
// The retained operator's two transforms of a plane in one workgroup:
// plane-major values -> inverse -> real plane (written to memory, the
// physical-space result) and, from the same registers, the forward row
// transforms -> forward columns -> plane-major values (may alias the input:
// a plane's values are read in full before they are written, by the one
// workgroup that owns the plane).  The planes of `count` evaluations
// (evaluation e's arrays pm_stride complex and real_stride real elements
// further on) are q = e * planes + p; workgroup b takes q = b, b + grid, ...
// in turn, one workgroup per CU (the LDS of one).  The next plane's values
// are loaded into registers while the current plane's outputs are copied
// out, so a plane does not wait for its input; the twiddles are staged once.
// ---------------------------------------------------------------------------
template <int KC>
__global__ __launch_bounds__(kPlaneThreads) void plane_operator_kernel(const cx<float>* pm_in,
                                                                      cx<float>* pm_out,
                                                                      float* __restrict__ real,
                                                                      const cx<float>* __restrict__ twg,
                                                                      const uint32_t* __restrict__ icodes,
                                                                      const uint32_t* __restrict__ fcodes, int modes,
                                                                      int np, int planes, size_t pm_stride,
                                                                      size_t real_stride, int total) {
    __shared__ __align__(16) unsigned char lds[PlaneGeo<KC>::BYTES];
    const int step = static_cast<int>(gridDim.x);
    int q = static_cast<int>(blockIdx.x);
    if (q >= total) return;
    const int wbase = plane_wbase();
    bf4 v[PlaneGeo<KC>::IT];
    load_plane_values<KC>(wbase, v, pm_in + pm_stride * (q / planes), modes, np, planes, q % planes);
    stage_twiddles<KC>(wbase, lds, twg);
    for (; q < total; q += step) {
        const int e = q / planes, p = q - e * planes;
        if (q != static_cast<int>(blockIdx.x)) lds_barrier();  // the previous plane's copy-out has read the staging
        cx<float> xa[2][PlaneGeo<KC>::KH], xb[2][PlaneGeo<KC>::KH];
        inverse_plane<KC, true>(wbase, lds, v, real + real_stride * e, icodes, modes, p, xa, xb);
        lds_barrier();  // every row-buffer read of the inverse is done before the band overwrites it
        const int qn = q + step;
        forward_columns<KC>(wbase, lds, xa, xb, pm_out + pm_stride * e, fcodes, modes, np, planes, p, [&] {
            const bool live = qn < total;
            const int qv = live ? qn : q;
            load_plane_values<KC>(wbase, v, pm_in + pm_stride * (qv / planes), modes, np, planes, qv % planes, live);
        }, [] {});
    }
}

// ---------------------------------------------------------------------------
// Compact (plane + P*mode) <-> chunked plane-major.  Workgroup b moves chunk
// b: the compact run [16*P*b, 16*P*b + nm*P) (nm = the chunk's modes, 16 but
// in the last chunk) and the chunked run [16*P*b, 16*P*b + 16*P), through an
// LDS tile [P][17].  Both sides are contiguous.  The chunk's padding modes
// are written as zeros.  Modes stored conjugated (bit m of the chunk's cmask
// word) are conjugated on the way: the compact side holds the caller's
// values, the chunked side slot values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ cx<float> conj_if(cx<float> v, uint32_t mask, int m) {
    return ((mask >> m) & 1u) ? cx<float>{v.x, -v.y} : v;
}

__device__ __forceinline__ uint32_t chunk_mask(const uint32_t* cmask, int b) {
    using ConstWord = const __attribute__((address_space(4))) uint32_t;
    return *((ConstWord*)cmask + b);
}

template <bool TO_PM>
__device__ __forceinline__ void pm_transpose_chunk(const cx<float>* __restrict__ in, cx<float>* __restrict__ out,
                                                   const uint32_t* __restrict__ cmask, int modes, int planes) {
    // 16-byte accesses (two complex values) on both sides: a chunk starts at
    // 16 * P * 8 bytes (a multiple of 128); its compact run nm * P is odd only
    // for a last partial chunk (one single tail element).  Loads go out in
    // batches of U per thread before any of them is used, so a chunk waits
    // for about one memory latency (P <= 256: one batch).
    constexpr int U = 8;
    extern __shared__ __align__(16) unsigned char pm_tile_lds[];
    cx<float>* tile = reinterpret_cast<cx<float>*>(pm_tile_lds);
    const int blk = static_cast<int>(blockIdx.x);
    const size_t base = static_cast<size_t>(16) * planes * blk;
    const int nm = modes - 16 * blk < 16 ? modes - 16 * blk : 16;
    const uint32_t mask = chunk_mask(cmask, blk);
    const int tid = static_cast<int>(threadIdx.x);
    const int run = nm * planes;  // compact elements of the chunk
    const int full = run >> 1;    // whole pairs of the compact run
    // compact element e <-> tile slot (plane, mode), conjugated as its mode says
    auto put = [&](int e, cx<float> v) {
        const int m = e / planes;
        tile[(e - m * planes) * 17 + m] = conj_if(v, mask, m);
    };
    auto get = [&](int e) {
        const int m = e / planes;
        return conj_if(tile[(e - m * planes) * 17 + m], mask, m);
    };
    if constexpr (TO_PM) {
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(in + base, static_cast<uint32_t>(run) * 8u);
        for (int b = 0; b < full; b += 256 * U) {
            bf4 q[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e2 = b + tid + 256 * u;
                q[u] = raw_buffer_load_f4(rs, static_cast<int>(e2 < full ? 16u * e2 : kBufferPast), 0, 0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e2 = b + tid + 256 * u;
                if (e2 < full) {
                    put(2 * e2, cx<float>{q[u].x, q[u].y});
                    put(2 * e2 + 1, cx<float>{q[u].z, q[u].w});
                }
            }
        }
        if ((run & 1) && tid == 0) put(run - 1, in[base + run - 1]);
        lds_barrier();
        bf4* dst = reinterpret_cast<bf4*>(out + base);
        for (int e2 = tid; e2 < 8 * planes; e2 += 256) {
            const int pl = e2 >> 3, m = 2 * (e2 & 7);
            const cx<float> a = m < nm ? tile[pl * 17 + m] : cx<float>{0.f, 0.f};
            const cx<float> c = m + 1 < nm ? tile[pl * 17 + m + 1] : cx<float>{0.f, 0.f};
            dst[e2] = bf4{a.x, a.y, c.x, c.y};
        }
    } else {
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(in + base, static_cast<uint32_t>(16 * planes) * 8u);
        for (int b = 0; b < 8 * planes; b += 256 * U) {
            bf4 q[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e2 = b + tid + 256 * u;
                q[u] = raw_buffer_load_f4(rs, static_cast<int>(e2 < 8 * planes ? 16u * e2 : kBufferPast), 0, 0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e2 = b + tid + 256 * u;
                if (e2 < 8 * planes) {
                    const int pl = e2 >> 3, m = 2 * (e2 & 7);
                    tile[pl * 17 + m] = cx<float>{q[u].x, q[u].y};
                    tile[pl * 17 + m + 1] = cx<float>{q[u].z, q[u].w};
                }
            }
        }
        lds_barrier();
        bf4* dst = reinterpret_cast<bf4*>(out + base);
        for (int e2 = tid; e2 < full; e2 += 256) {
            const cx<float> a = get(2 * e2), c = get(2 * e2 + 1);
            dst[e2] = bf4{a.x, a.y, c.x, c.y};
        }
        if ((run & 1) && tid == 0) out[base + run - 1] = get(run - 1);
    }
}

template <bool TO_PM>
__global__ __launch_bounds__(256) void pm_transpose_kernel(const cx<float>* __restrict__ in,
                                                           cx<float>* __restrict__ out,
                                                           const uint32_t* __restrict__ cmask, int modes,
                                                           int planes) {
    pm_transpose_chunk<TO_PM>(in, out, cmask, modes, planes);
}

// Compact -> chunked plane-major for the evaluations of a batch: blockIdx.y
// is the evaluation, its compact input src.in[y], its block at out +
// y * out_stride.
__global__ __launch_bounds__(256) void pm_to_pm_batch_kernel(const BatchInputs src, cx<float>* __restrict__ out,
                                                             size_t out_stride, const uint32_t* __restrict__ cmask,
                                                             int modes, int planes) {
    pm_transpose_chunk<true>(static_cast<const cx<float>*>(src.in[blockIdx.y]), out + out_stride * blockIdx.y, cmask,
                             modes, planes);
}

// The same move with 8-byte accesses on the compact side, for compact arrays
// that are not 16-byte aligned.
template <bool TO_PM>
__global__ __launch_bounds__(256) void pm_transpose_scalar_kernel(const cx<float>* __restrict__ in,
                                                                  cx<float>* __restrict__ out,
                                                                  const uint32_t* __restrict__ cmask, int modes,
                                                                  int planes) {
    extern __shared__ __align__(16) unsigned char pm_tile_lds[];
    cx<float>* tile = reinterpret_cast<cx<float>*>(pm_tile_lds);
    const int blk = static_cast<int>(blockIdx.x);
    const size_t base = static_cast<size_t>(16) * planes * blk;
    const int nm = modes - 16 * blk < 16 ? modes - 16 * blk : 16;
    const uint32_t mask = chunk_mask(cmask, blk);
    const int tid = static_cast<int>(threadIdx.x);
    if constexpr (TO_PM) {
        for (int e = tid; e < nm * planes; e += 256) {
            const int m = e / planes, pl = e - m * planes;
            tile[pl * 17 + m] = conj_if(in[base + e], mask, m);
        }
        lds_barrier();
        for (int e = tid; e < 16 * planes; e += 256) {
            const int m = e & 15, pl = e >> 4;
            out[base + e] = m < nm ? tile[pl * 17 + m] : cx<float>{0.f, 0.f};
        }
    } else {
        for (int e = tid; e < 16 * planes; e += 256) {
            const int m = e & 15, pl = e >> 4;
            tile[pl * 17 + m] = in[base + e];
        }
        lds_barrier();
        for (int e = tid; e < nm * planes; e += 256) {
            const int m = e / planes, pl = e - m * planes;
            out[base + e] = conj_if(tile[pl * 17 + m], mask, m);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
bool plane_band_empty(const std::vector<int>& fmap, const std::vector<int>& imap, uint64_t kc) {
    using G = PlaneGeo<kPlaneKc>;
    if (kc != kPlaneKc) return false;
    for (int c = 0; c < kPlaneKc; ++c)
        for (int ky = 16 * G::J0; ky < 16 * G::J1 + 16; ++ky)
            if (fmap[static_cast<size_t>(c) * kPlaneN + ky] >= 0 || imap[static_cast<size_t>(c) * kPlaneN + ky] >= 0)
                return false;
    return true;
}

bool plane_supported(const skb_workload& w, uint64_t kc, uint64_t modes, int precision) {
    const uint64_t planes = w.nz * w.fields;
    return precision == SKB_F32 && w.nx == kPlaneN && w.ny == kPlaneN && kc == kPlaneKc &&
           modes <= static_cast<uint64_t>(kPlaneMaxModes) && modes > 0 && planes > 0 &&
           planes <= static_cast<uint64_t>(kPlaneMaxPlanes) &&
           plane_np(modes) * planes * 8ull < (1ull << 31);
}

uint64_t plane_np(uint64_t modes) { return (modes + 15) / 16 * 16; }

void PlaneTables::build(const skb_workload& w, uint64_t kc_, uint64_t modes_, const std::vector<int>& fmap,
                        const std::vector<int>& imap) {
    if (!plane_supported(w, kc_, modes_, SKB_F32)) fail(SKB_UNSUPPORTED, "plane-resident transforms: unsupported shape");
    modes = static_cast<int>(modes_);
    np = static_cast<int>(plane_np(modes_));
    const int N = kPlaneN, KC = kPlaneKc;
    // per (column c, lane n) a 16-entry slot: the codes of the registers
    // outside the never-retained band at code_slot(k), two per 32-bit word
    std::vector<uint16_t> f(static_cast<size_t>(KC) * 16 * 16, 0), i(f.size(), 0);
    // Slot codes (2*mode + conj from the column lists) -> staged indices.  A
    // slot the forward writes holds its mode's slot value, so the inverse
    // reads it as it is; the inverse's extra slots are the k_x = 0 partners,
    // whose value is the conjugate (kPartner).  Every mode's conj flag goes to
    // cmask (bit m % 16 of word m / 16).
    std::vector<uint32_t> cm(static_cast<size_t>(np / 16), 0u);
    for (int c = 0; c < KC; ++c)
        for (int n = 0; n < 16; ++n)
            for (int k = 0; k < 16; ++k) {
                const int fc = fmap[static_cast<size_t>(c) * N + n + 16 * k];
                const int fi = fmap[static_cast<size_t>(c) * N + 16 * k + n];
                const int ic = imap[static_cast<size_t>(c) * N + 16 * k + n];
                if (fc >= 0 && (fc & 1)) cm[static_cast<size_t>(fc >> 1) / 16] |= 1u << ((fc >> 1) % 16);
                if (fi >= 0 && fi != ic) fail(SKB_LOGIC_ERROR, "plane codes: forward and inverse slots disagree");
                if (PlaneGeo<kPlaneKc>::empty(k)) {
                    if (fc >= 0 || ic >= 0) fail(SKB_LOGIC_ERROR, "plane codes: a retained slot in the empty band");
                    continue;
                }
                const size_t at = (static_cast<size_t>(c) * 16 + n) * 16 + code_slot<kPlaneKc>(k);
                f[at] = static_cast<uint16_t>(fc < 0 ? kPlaneMaxModes : fc >> 1);
                i[at] = static_cast<uint16_t>(ic < 0 ? kPlaneMaxModes : (ic >> 1) | (fi < 0 ? kPartner : 0));
            }
    cmask.upload(cm.data(), cm.size() * sizeof(uint32_t));
    fcodes.upload(f.data(), f.size() * sizeof(uint16_t));
    icodes.upload(i.data(), i.size() * sizeof(uint16_t));
    // twiddles of the 16 x 16 split, W^(n k) at 16 k + n
    std::vector<cx<float>> t(N);
    for (int k = 0; k < 16; ++k)
        for (int n = 0; n < 16; ++n) {
            const double a = 2.0 * M_PI * (n * k) / N;
            t[16 * k + n] = cx<float>{static_cast<float>(std::cos(a)), static_cast<float>(-std::sin(a))};
        }
    tw.upload(t.data(), t.size() * sizeof(t[0]));
    int dev = 0, cus = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    workgroups = std::max(1, cus);
}

void plane_inverse(const PlaneTables& t, const void* pm, void* real, int planes, hipStream_t s) {
    hipLaunchKernelGGL((plane_inverse_kernel<kPlaneKc>), dim3(planes), dim3(kPlaneThreads), 0, s,
                       static_cast<const cx<float>*>(pm), static_cast<float*>(real), t.tw.as<cx<float>>(),
                       t.icodes.as<uint32_t>(), t.modes, t.np, planes);
    check_hip(hipGetLastError(), "plane inverse launch");
}

void plane_forward(const PlaneTables& t, const void* real, void* pm, int planes, hipStream_t s) {
    hipLaunchKernelGGL((plane_forward_kernel<kPlaneKc>), dim3(planes), dim3(kPlaneThreads), 0, s,
                       static_cast<const float*>(real), static_cast<cx<float>*>(pm), t.tw.as<cx<float>>(),
                       t.fcodes.as<uint32_t>(), t.modes, t.np, planes);
    check_hip(hipGetLastError(), "plane forward launch");
}

void plane_operator(const PlaneTables& t, const void* pm_in, void* pm_out, void* real, int planes, hipStream_t s,
                    int count, size_t pm_stride, size_t real_stride) {
    const int total = planes * count;
    const int grid = std::min(total, t.workgroups);
    hipLaunchKernelGGL((plane_operator_kernel<kPlaneKc>), dim3(grid), dim3(kPlaneThreads), 0, s,
                       static_cast<const cx<float>*>(pm_in), static_cast<cx<float>*>(pm_out),
                       static_cast<float*>(real), t.tw.as<cx<float>>(), t.icodes.as<uint32_t>(),
                       t.fcodes.as<uint32_t>(), t.modes, t.np, planes, pm_stride, real_stride, total);
    check_hip(hipGetLastError(), "plane operator launch");
}

void pm_transpose(const PlaneTables& t, bool to_pm, const void* in, void* out, int planes, hipStream_t s) {
    const int modes = t.modes;
    const dim3 grid((modes + 15) / 16);
    const size_t lds = static_cast<size_t>(planes) * 17 * sizeof(cx<float>);
    const auto* i = static_cast<const cx<float>*>(in);
    auto* o = static_cast<cx<float>*>(out);
    const uint32_t* m = t.cmask.as<uint32_t>();
    const bool wide = (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    if (wide && to_pm) hipLaunchKernelGGL(pm_transpose_kernel<true>, grid, dim3(256), lds, s, i, o, m, modes, planes);
    else if (wide) hipLaunchKernelGGL(pm_transpose_kernel<false>, grid, dim3(256), lds, s, i, o, m, modes, planes);
    else if (to_pm) hipLaunchKernelGGL(pm_transpose_scalar_kernel<true>, grid, dim3(256), lds, s, i, o, m, modes, planes);
    else hipLaunchKernelGGL(pm_transpose_scalar_kernel<false>, grid, dim3(256), lds, s, i, o, m, modes, planes);
    check_hip(hipGetLastError(), "plane-major transpose launch");
}

void pm_transpose_batch(const PlaneTables& t, const BatchInputs& in, int count, void* out, size_t out_stride,
                        int planes, hipStream_t s) {
    if (count < 1 || count > kBatchMax) fail(SKB_INVALID_ARGUMENT, "batch of evaluations out of range");
    for (int e = 0; e < count; ++e)
        if (reinterpret_cast<uintptr_t>(in.in[e]) % 16 != 0)
            fail(SKB_INVALID_ARGUMENT, "batched evaluations need 16-byte aligned retained inputs");
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0 || (out_stride * sizeof(cx<float>)) % 16 != 0)
        fail(SKB_LOGIC_ERROR, "batched plane-major blocks must be 16-byte aligned");
    const dim3 grid((t.modes + 15) / 16, count);
    const size_t lds = static_cast<size_t>(planes) * 17 * sizeof(cx<float>);
    hipLaunchKernelGGL(pm_to_pm_batch_kernel, grid, dim3(256), lds, s, in, static_cast<cx<float>*>(out), out_stride,
                       t.cmask.as<uint32_t>(), t.modes, planes);
    check_hip(hipGetLastError(), "plane-major transpose launch");
}

}  // namespace skb
