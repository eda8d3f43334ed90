/*
 * skb_hip.h -- C ABI of the MI355X (gfx950) `hip` provider for skbench's
 * retained-mode spectral-operator path.
 *
 * Every compute entry point works on DEVICE pointers and enqueues on the
 * caller's hipStream_t (passed as void*; NULL = default stream).  Nothing
 * here allocates after plan creation.  Plans are not re-entrant: one
 * in-flight call per plan (the reference providers are single-caller too,
 * include/skbench/skbench.hpp:364-372).
 *
 * Status codes map to the reference's exception types (src/fftw_provider.cpp
 * :468-487, src/vertical_gemm.cpp:431-463):
 *   SKB_INVALID_ARGUMENT -> std::invalid_argument  (bad shape / stride / size)
 *   SKB_RUNTIME_ERROR    -> std::runtime_error     (HIP or device failure)
 *   SKB_LOGIC_ERROR      -> std::logic_error       (wrong entry point / layout)
 *   SKB_BAD_ALLOC        -> std::bad_alloc
 *   SKB_UNSUPPORTED      -> provider capability() "unsupported" record
 * skb_last_error() returns the calling thread's last message.
 *
 * Layouts are the reference's, element for element (src/model.cpp:219-244):
 *   real       x + Nx*(y + Ny*(z + Nz*field))
 *   WVM        z + Nz*field + P*(kx + (Nx/2+1)*ky),   P = Nz*fields
 *   retained   z + Nz*field + P*mode
 *   modal      j + Nj*field + Nj*fields*mode
 * Complex values are interleaved (re, im) pairs of the plan's precision:
 * SKB_F64 -> double[2] (binary-identical to skbench::Complex), SKB_F32 ->
 * float[2].
 */
#ifndef SKB_HIP_H
#define SKB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum skb_status {
    SKB_OK = 0,
    SKB_INVALID_ARGUMENT = 1,
    SKB_RUNTIME_ERROR = 2,
    SKB_LOGIC_ERROR = 3,
    SKB_BAD_ALLOC = 4,
    SKB_UNSUPPORTED = 5
};

enum skb_precision { SKB_F64 = 0, SKB_F32 = 1 };

/* skbench::Workload (include/skbench/skbench.hpp:23-39). */
typedef struct skb_workload {
    uint64_t nx, ny, nz, fields;
    double lx, ly;
    int32_t antialias;
    int32_t reserved;
} skb_workload;

/* skbench::RetainedMode (include/skbench/skbench.hpp:41-48). */
typedef struct skb_retained_mode {
    int64_t k, l;
    uint64_t stored_kx, stored_ky;
    int32_t conjugates_stored_value;
    int32_t reserved;
    double radial;
} skb_retained_mode;

/* skbench::VerticalModeGroup (include/skbench/skbench.hpp:50-56). */
typedef struct skb_mode_group {
    uint64_t squared_key, first_mode, mode_count;
} skb_mode_group;

/* ---- errors / library identity --------------------------------------- */
/* Copies the calling thread's last error message (NUL-terminated). */
size_t skb_last_error(char* buffer, size_t capacity);
/* Library identity string, e.g. "skb-hip gfx950" (FFTWProvider::version,
 * include/skbench/skbench.hpp:415-416). */
const char* skb_version(void);

/* ---- host model (src/model.cpp) -------------------------------------- */
/* retainedHorizontalModes (src/model.cpp:129-167).  Pass modes=NULL to get
 * the count only. */
int skb_retained_modes(const skb_workload* workload, skb_retained_mode* modes, uint64_t capacity,
                       uint64_t* count);
/* squaredWavenumberGroups (src/model.cpp:169-207). */
int skb_k2_groups(const skb_retained_mode* modes, uint64_t mode_count, skb_mode_group* groups,
                  uint64_t capacity, uint64_t* count);
/* Contiguous bins of K^2 = k^2 + l^2 over the retained order:
 * bin = floor(bins * K^2 / (K^2_max + 1)), empty bins dropped.  Builds the
 * coarse grouping of the 32-bin vertical family; no reference counterpart. */
int skb_k2_binned_groups(const skb_retained_mode* modes, uint64_t mode_count, uint64_t bins,
                         skb_mode_group* groups, uint64_t capacity, uint64_t* count);
/* FNV-1a hashes (src/model.cpp:209-217, 462-487). */
uint64_t skb_mode_order_hash(const skb_retained_mode* modes, uint64_t count);
uint64_t skb_group_order_hash(const skb_mode_group* groups, uint64_t count);
uint64_t skb_wvm_order_hash(const skb_workload* workload);
/* orthonormalVerticalFixture (src/model.cpp:489-506): forward nj x nz,
 * inverse nz x nj, row-major. */
int skb_orthonormal_vertical(uint64_t nz, uint64_t nj, double* forward, double* inverse);
/* squaredWavenumberVerticalFixture (src/model.cpp:523-565): one nj x nz
 * (forward) and nz x nj (inverse) matrix per group, contiguous. */
int skb_k2_vertical_family(uint64_t nz, uint64_t nj, const skb_mode_group* groups, uint64_t group_count,
                           double* forward, double* inverse);
/* n x n orthogonal family from the QR factorisation of seeded Gaussian
 * matrices (group g uses seed + g); forward = Q^T, inverse = Q.  No
 * reference counterpart (32-bin 65x65 family). */
int skb_qr_gaussian_family(uint64_t n, uint64_t group_count, uint64_t seed, double* forward, double* inverse);
/* makeFixture (src/model.cpp:618-661).  kind: 0 impulse, 1 sinusoid (1,2),
 * 2 random = uniform[-1,1) from the portable generator of
 * src/benchmark.cpp:187-196, 3 dc, 4 nyquist, 5 sinusoid (5,17). */
int skb_make_fixture(const skb_workload* workload, int kind, uint64_t seed, double* real);
/* Compact retained spectrum: complex Gaussian (Box-Muller over the same
 * generator) times (1 + k^2 + l^2)^exponent, DC imaginary part zero. */
int skb_spectral_fixture(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                         uint64_t seed, double exponent, double* retained_interleaved);

/* ---- full batched 2-D r2c / c2r in WVM order ---------------------------
 * Replaces FFTWProvider (include/skbench/skbench.hpp:364-421;
 * src/fftw_provider.cpp:304-316 forward/inverse, 537-610 guru64 geometry). */
typedef struct skb_fft_plan skb_fft_plan;
int skb_fft_plan_create(const skb_workload* workload, int precision, skb_fft_plan** plan);
int skb_fft_plan_destroy(skb_fft_plan* plan);
/* FFTWProvider::forward: unnormalised r2c, real -> WVM half spectrum. */
int skb_fft_forward(skb_fft_plan* plan, const void* real, void* wvm, void* stream);
/* FFTWProvider::inverse: unnormalised c2r; DESTROYS `wvm` (as FFTW c2r may,
 * src/fftw_provider.cpp:1290-1292). */
int skb_fft_inverse(skb_fft_plan* plan, void* wvm, void* real, void* stream);
/* FFTWInterleavedWvmFieldViewProvider::forwardField / inverseFields
 * (include/skbench/skbench.hpp:423-465; src/fftw_provider.cpp:1175-1230):
 * the full transforms of one field held inside an interleaved WVM array.
 * The plan's workload is the field's (planes = Nz * fields, normally
 * fields = 1); the view holds element (kx, ky, plane) at
 * spectrum[plane + frequency_stride * (kx + (Nx/2+1) * ky)] with
 * frequency_stride >= planes (a field f of an F-field array: spectrum =
 * base + f * Nz, stride F * Nz).  forward_view writes the view's planes only;
 * inverse_view may overwrite the view, as FFTW's c2r does.  A stride below
 * the plane count is SKB_INVALID_ARGUMENT. */
int skb_fft_forward_view(skb_fft_plan* plan, const void* real, void* spectrum, uint64_t frequency_stride,
                         void* stream);
int skb_fft_inverse_view(skb_fft_plan* plan, void* spectrum, uint64_t frequency_stride, void* real, void* stream);
/* FFTWProvider::gatherRetainedOuter / embedRetainedOuter
 * (src/fftw_provider.cpp:318-329; src/model.cpp:310-339). */
int skb_fft_gather_retained(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t mode_count,
                            const void* wvm, void* retained, void* stream);
int skb_fft_embed_retained(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t mode_count,
                           const void* retained, void* wvm, void* stream);
/* gatherRetainedModal (src/model.cpp:341-355): full WVM-order modal array
 * (j + Nj*field + Nj*fields*(kx + (Nx/2+1)*ky)) -> compact modal array
 * (j + Nj*field + Nj*fields*mode), conjugated where the mode conjugates its
 * stored value. */
int skb_fft_gather_retained_modal(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t mode_count,
                                  uint64_t nj, const void* full_modal, void* modal, void* stream);

/* ---- pruned retained-mode transforms -----------------------------------
 * Replaces FFTWPrunedProvider (include/skbench/skbench.hpp:467-520;
 * src/pruned_fftw_provider.cpp:230-241 forward, 265-276 inverse): only the
 * active k_x band is transformed along y and only retained modes are read
 * or written.  Caller inputs are preserved. */
typedef struct skb_pruned_plan skb_pruned_plan;
int skb_pruned_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                           int precision, skb_pruned_plan** plan);
int skb_pruned_plan_destroy(skb_pruned_plan* plan);
/* T_h^+ = P_h F_xy: real planes -> compact retained (z, field, mode). */
int skb_pruned_forward(skb_pruned_plan* plan, const void* real, void* retained, void* stream);
/* T_h^- = F_xy^-1 E_h: compact retained -> real planes (unnormalised). */
int skb_pruned_inverse(skb_pruned_plan* plan, const void* retained, void* real, void* stream);
/* Stage-by-stage form of the two calls above, for per-kernel timing
 * (FFTWPrunedProvider::executeForwardRows / executeForwardColumns +
 * gatherForward / embedInverse + executeInverseColumns / executeInverseRows,
 * include/skbench/skbench.hpp:478-489).  The intermediate lives in the
 * plan's scratch, so a forward_rows must precede forward_columns and an
 * inverse_columns must precede inverse_rows on the same stream. */
/* Split compact form: separate real and imaginary arrays.
 * forward_split / inverse_split: element (plane, mode) at plane + P*mode
 * (FFTWPrunedProvider::forwardSplit / inverseSplit and
 * FFTWStreamingPrunedSplitProvider::forwardSplit / inverseSplit,
 * include/skbench/skbench.hpp:484-492, 563-566;
 * src/pruned_fftw_provider.cpp:1053-1066).  Forward outputs are multiplied
 * by `scale`. */
int skb_pruned_forward_split(skb_pruned_plan* plan, const void* real, void* retained_real, void* retained_imag,
                             double scale, void* stream);
int skb_pruned_inverse_split(skb_pruned_plan* plan, const void* retained_real, const void* retained_imag, void* real,
                             void* stream);
/* One field of a split compact array: element (z, mode) of the field at
 * real[z + mode_stride * mode] and imaginary[...], device pointers in the
 * plan's precision, mode_stride >= Nz (FieldView / ConstFieldView,
 * include/skbench/skbench.hpp:536-546; src/pruned_fftw_provider.cpp:1068-1083,
 * 1176-1196).  field_count must equal the workload's fields (at most 16). */
typedef struct skb_field_view {
    void* real;
    void* imaginary;
    uint64_t mode_stride;
} skb_field_view;
/* FFTWStreamingPrunedSplitProvider::forwardSplitFields / inverseSplitFields
 * (include/skbench/skbench.hpp:567-570; src/pruned_fftw_provider.cpp:787-805). */
int skb_pruned_forward_split_fields(skb_pruned_plan* plan, const void* real, const skb_field_view* fields,
                                    uint64_t field_count, double scale, void* stream);
int skb_pruned_inverse_split_fields(skb_pruned_plan* plan, const skb_field_view* fields, uint64_t field_count,
                                    void* real, void* stream);
int skb_pruned_forward_rows(skb_pruned_plan* plan, const void* real, void* stream);
int skb_pruned_forward_columns(skb_pruned_plan* plan, void* retained, void* stream);
int skb_pruned_inverse_columns(skb_pruned_plan* plan, const void* retained, void* stream);
int skb_pruned_inverse_rows(skb_pruned_plan* plan, void* real, void* stream);
/* Launch shape of one stage (0 forward rows, 1 forward columns, 2 inverse
 * columns, 3 inverse rows): the number of workgroups, 0 = one per work item
 * (the default).  Fewer workgroups than items makes them persistent, each
 * prefetching its next item.  Set by the launch autotuner.  The persistent
 * kernels are separate instantiations and may round differently in the last
 * bits (FMA contraction), so results agree with the default launch to
 * rounding, not bit for bit.  No reference counterpart (FFTW plans choose
 * internally). */
int skb_pruned_plan_set_grid(skb_pruned_plan* plan, int stage, int workgroups);
uint64_t skb_pruned_plan_items(const skb_pruned_plan* plan, int stage);
/* FFTWPrunedProvider::activeKxCount / scratchBytes. */
uint64_t skb_pruned_active_kx(const skb_pruned_plan* plan);
uint64_t skb_pruned_scratch_bytes(const skb_pruned_plan* plan);
/* Transform path of skb_pruned_forward and skb_pruned_inverse.
 * SKB_PATH_AUTO (the default) takes SKB_PATH_PLANE where it is built (FP32,
 * 256 x 256 planes, active k_x band 86: the 2/3-rule set) and
 * SKB_PATH_TWO_PASS elsewhere.  SKB_PATH_PLANE: each plane's 2-D transform
 * runs in one workgroup with the retained band kept on chip; the compact
 * operand passes through a plane-major copy (mode + np*plane) in the plan's
 * scratch.  SKB_PATH_TWO_PASS: row and column kernels through the
 * kc x Ny x P intermediate (the only path of the split, field-view and
 * stage entry points).  The paths agree to rounding.  Setting
 * SKB_PATH_PLANE where it is not built fails with SKB_UNSUPPORTED.  No
 * reference counterpart (FFTW plans choose their algorithm internally). */
enum skb_transform_path { SKB_PATH_AUTO = 0, SKB_PATH_TWO_PASS = 1, SKB_PATH_PLANE = 2 };
int skb_pruned_plan_set_path(skb_pruned_plan* plan, int path);
/* The path forward / inverse take: SKB_PATH_PLANE or SKB_PATH_TWO_PASS. */
int skb_pruned_plan_path(const skb_pruned_plan* plan);

/* ---- retained-mode operator ---------------------------------------------
 * One evaluation of the antialiased retained-mode operator enqueued from C++
 * in one call: the pruned inverse c2r of the compact retained spectrum
 * (z, field, mode) into real planes, the forward r2c truncated back to the
 * retained modes, and the K^2-grouped vertical forward product into the
 * modal layout j + Nj*(field + F*mode).  The reference's harnesses time the
 * same composition as separate provider calls: FFTWPrunedProvider::inverse
 * and forward (include/skbench/skbench.hpp:467-520;
 * src/pruned_fftw_provider.cpp:230-276) and VerticalGemmProvider::
 * executeForward (include/skbench/skbench.hpp:152-218;
 * src/vertical_gemm.cpp:534-596).  The real planes are written and read in
 * the plan's own buffer (skb_retained_operator_real).  groups, forward and
 * inverse: the vertical family as for skb_vertical_plan_create.  The input
 * is preserved; the plan owns every intermediate, allocated at creation. */
typedef struct skb_retained_operator skb_retained_operator;
int skb_retained_operator_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                 uint64_t nj, const skb_mode_group* groups, uint64_t group_count,
                                 const double* forward, const double* inverse, int precision,
                                 skb_retained_operator** op);
int skb_retained_operator_destroy(skb_retained_operator* op);
int skb_retained_operator_execute(skb_retained_operator* op, const void* retained, void* modal, void* stream);
/* One stage of an evaluation, for per-kernel timing: SKB_OP_STAGE_LAYOUT
 * (the plane path's plane-major copy of `retained`; nothing on the two-pass
 * path), SKB_OP_STAGE_INVERSE (into the real planes), SKB_OP_STAGE_FORWARD
 * (real planes to the plan's retained intermediate) and SKB_OP_STAGE_VERTICAL
 * (into `modal`).  Each reads what the stage before it left in the plan's
 * buffers, so LAYOUT, INVERSE, FORWARD, VERTICAL in order on one stream
 * equal one execute.  SKB_OP_STAGE_TRANSFORMS runs INVERSE then FORWARD as
 * execute does (one kernel on the fused plane path). */
enum skb_operator_stage {
    SKB_OP_STAGE_LAYOUT = 0,
    SKB_OP_STAGE_INVERSE = 1,
    SKB_OP_STAGE_FORWARD = 2,
    SKB_OP_STAGE_VERTICAL = 3,
    SKB_OP_STAGE_TRANSFORMS = 4
};
int skb_retained_operator_stage(skb_retained_operator* op, int stage, const void* retained, void* modal,
                                void* stream);
/* Transform path (enum skb_transform_path), as for the pruned plan; on the
 * plane path the forward transform hands the vertical product its
 * plane-major output directly when the vertical plan reads that layout
 * (square FP32 split-BF16 families, Nz = 33, 65, 129). */
int skb_retained_operator_set_path(skb_retained_operator* op, int path);
int skb_retained_operator_path(const skb_retained_operator* op);
/* Plane path: execute runs each plane's inverse and forward transform in one
 * workgroup (1, the default): the real plane is written to memory and read
 * back by the same workgroup, so the two passes' memory phases overlap.
 * 0 launches the two transforms separately, as the stages do.  Results are
 * identical. */
int skb_retained_operator_set_fused(skb_retained_operator* op, int fused);
int skb_retained_operator_fused(const skb_retained_operator* op);
/* Batched evaluation.  reserve_lanes(n) gives the plan n lanes (n in
 * [1, 8]): streams of its own, each checked at reservation to sit on a
 * hardware queue no other lane uses (lanes_apart), and, past the first lane,
 * a second set of intermediates (real planes and plane-major blocks); it
 * allocates, and n = 1 releases them.  Plane path only (elsewhere it keeps
 * one lane).  execute_batch runs `count` independent evaluations
 * (retained[i] -> modal[i], host arrays of device pointers) alternating over
 * the lanes, forked from `stream` and joined back to it: every output is
 * complete when `stream` passes the call, and each equals the same
 * evaluation run alone bit for bit.  Outputs must not alias each other or an
 * input.  One evaluation's layout copy and vertical product then run on the
 * CUs another evaluation's plane workgroups leave idle.  With one lane the
 * batch runs in order on `stream`.  No reference counterpart (the reference
 * evaluates one operator at a time on the host). */
int skb_retained_operator_reserve_lanes(skb_retained_operator* op, int lanes);
/* Lanes reserved (>= 1); *apart = 1 when every lane has a queue of its own. */
int skb_retained_operator_lanes(const skb_retained_operator* op, int* apart);
/* Stage-batched evaluation.  reserve_batch(n), n in [2, 128], gives the plan
 * the plane-major blocks and real planes of n evaluations (it allocates; 0
 * or 1 releases them).  execute_batch then takes the evaluations in groups
 * of up to n and runs each stage as ONE launch over the group: the layout
 * copies, the plane transforms (the plane workgroups of all evaluations fill
 * every CU) and the vertical products (persistent workgroups over all
 * evaluations' tiles).  It takes precedence over the lanes, runs in order on
 * `stream`, allocates nothing, needs 16-byte aligned inputs, and each output
 * equals the same evaluation run alone bit for bit.  Plane path with a
 * plane-major vertical product (square split-BF16 FP32 families) only;
 * elsewhere reserve_batch fails with SKB_UNSUPPORTED.  No reference
 * counterpart. */
int skb_retained_operator_reserve_batch(skb_retained_operator* op, int evaluations);
/* Evaluations of one stage-batched group (0 when none is reserved). */
int skb_retained_operator_batch(const skb_retained_operator* op);
/* One stage of the stage-batched execute (SKB_OP_STAGE_LAYOUT, _TRANSFORMS
 * or _VERTICAL), one launch per group, for per-kernel timing.  The groups
 * share the plan's intermediates, so the three calls in that order equal
 * execute_batch when count is at most the reserved batch; a larger count
 * launches the same work (every group overwrites the last).  SKB_LOGIC_ERROR
 * when no batch is reserved.  No reference counterpart. */
int skb_retained_operator_execute_batch_stage(skb_retained_operator* op, int stage, int count,
                                              const void* const* retained, void* const* modal, void* stream);
int skb_retained_operator_execute_batch(skb_retained_operator* op, int count, const void* const* retained,
                                        void* const* modal, void* stream);
/* Device pointer to the real planes (x + Nx(y + Ny*plane)) of the last
 * evaluation (after a batch: the last one run on the first lane), and the
 * plan's device allocation in bytes. */
void* skb_retained_operator_real(const skb_retained_operator* op);
uint64_t skb_retained_operator_device_bytes(const skb_retained_operator* op);

/* ---- K^2-grouped vertical GEMM -----------------------------------------
 * Replaces VerticalGemmProvider (include/skbench/skbench.hpp:152-218;
 * src/vertical_gemm.cpp:534-596 grouped zgemm / split dgemm): per group g,
 * modal[:, cols_g] = F_g (nj x nz) * physical[:, cols_g] and
 * physical[:, cols_g] = I_g (nz x nj) * modal[:, cols_g], columns
 * (field + fields*mode) of contiguous z (or j) vectors.  Runs on MFMA
 * (f64 16x16x4 for SKB_F64, f32 16x16x4 for SKB_F32). */
typedef struct skb_vertical_plan skb_vertical_plan;
int skb_vertical_plan_create(const skb_workload* workload, uint64_t nj, const skb_mode_group* groups,
                             uint64_t group_count, const double* forward, const double* inverse, int precision,
                             skb_vertical_plan** plan);
int skb_vertical_plan_destroy(skb_vertical_plan* plan);
int skb_vertical_forward(skb_vertical_plan* plan, const void* physical, void* modal, void* stream);
int skb_vertical_inverse(skb_vertical_plan* plan, const void* modal, void* physical, void* stream);
/* Split layout (VerticalGemmLayout::split, include/skbench/skbench.hpp:125-128;
 * the split dgemm of src/vertical_gemm.cpp:558-596 and its data accessors,
 * include/skbench/skbench.hpp:203-210): real and imaginary parts in separate
 * arrays, each in the interleaved layout's element order.  reserve_split
 * allocates the plan's interleaved staging once (the provider constructor's
 * choice of layout); the split calls without it fail with SKB_LOGIC_ERROR.
 * The matrices are real, so the real and imaginary parts are independent
 * products (executeForwardReal / executeForwardImaginary, src/vertical_gemm.cpp:
 * 771-819): a part whose input and output pointers are both null is not
 * computed and its output array is not touched. */
int skb_vertical_plan_reserve_split(skb_vertical_plan* plan);
int skb_vertical_forward_split(skb_vertical_plan* plan, const void* physical_real, const void* physical_imag,
                               void* modal_real, void* modal_imag, void* stream);
int skb_vertical_inverse_split(skb_vertical_plan* plan, const void* modal_real, const void* modal_imag,
                               void* physical_real, void* physical_imag, void* stream);
/* skb_vertical_forward_split without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops): it needs no
 * skb_vertical_plan_reserve_split and calls no hipMalloc, hipFree or other
 * allocating routine.  Its interleaved staging lives in the library's static
 * device scratch of 64 MiB per GPU (held until the process ends); a plan
 * whose staging, modes*Nz*fields + modes*Nj*fields complex values (the first
 * part rounded up to 256 bytes), exceeds it gets SKB_INVALID_ARGUMENT.  All
 * calls on one GPU share that scratch: calls that overlap on one GPU (on
 * different streams, or from different threads) race. */
int skb_vertical_forward_split_noalloc(skb_vertical_plan* plan, const void* physical_real, const void* physical_imag,
                                       void* modal_real, void* modal_imag, void* stream);
/* VerticalGemmProvider::matrixBytesPerDirection. */
uint64_t skb_vertical_matrix_bytes(const skb_vertical_plan* plan);
/* Launch shape: workgroups per launch, 0 = one per 16-column tile (the
 * default); fewer makes them persistent.  Set by the launch autotuner; no
 * reference counterpart. */
int skb_vertical_plan_set_grid(skb_vertical_plan* plan, int workgroups);
uint64_t skb_vertical_plan_items(const skb_vertical_plan* plan);
/* Products of an FP32 plan's compact GEMMs.  SKB_PRODUCTS_SPLIT_BF16 forms
 * each product on the bf16 matrix cores as the six leading terms of
 * (A0 + A1 + A2)(b0 + b1 + b2), every factor split into three bf16 terms,
 * with FP32 accumulation (the dropped terms are below 2^-24 of the product);
 * SKB_PRODUCTS_FP32 uses the f32 16x16x4 matrix cores.  Split products need
 * K = 32*n + 1 and M = 16*m + r with r <= 4 (Nz = 33, 65, 129 square
 * families); other shapes, FP64 plans and the WVM-direct GEMM always use
 * FP32 (FP64) products.  The default is SKB_PRODUCTS_SPLIT_BF16.  No
 * reference counterpart (the reference's GEMMs are FP64 BLAS). */
enum skb_gemm_products { SKB_PRODUCTS_FP32 = 0, SKB_PRODUCTS_SPLIT_BF16 = 1 };
int skb_vertical_plan_set_products(skb_vertical_plan* plan, int products);
/* The products the plan's forward (direction 0) or inverse (1) GEMM uses. */
int skb_vertical_plan_products(const skb_vertical_plan* plan, int direction);

/* ---- WVM-direct vertical GEMM ------------------------------------------
 * Replaces WvmDirectVerticalGemmProvider (include/skbench/skbench.hpp:220-257;
 * src/vertical_gemm.cpp:935-1273): the same grouped products applied in place
 * on FULL WVM arrays.  Mode m at frequency q = stored_kx + (Nx/2+1)*stored_ky
 * reads the Nz x fields block at spectrum + P*q and writes the Nj x fields
 * block at modal + Nj*fields*q (z / j fastest); for k_x = 0 modes with
 * k_y != 0, 2k_y != Ny the conjugate is also written at (0, Ny - k_y)
 * (src/vertical_gemm.cpp:1081-1131).  Entries of non-retained frequencies are
 * not written: initialise the output once with the initialize calls
 * (initializeModalOutput / initializeSpectrumOutput, which zero the whole
 * array).  nj must equal the workload's retained vertical modes
 * 2(Nz-1)/3 and every mode needs a unique stored frequency
 * (src/vertical_gemm.cpp:965-1034).  The compact and WVM entry points reject
 * a plan of the other kind with SKB_LOGIC_ERROR. */
int skb_vertical_wvm_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                 uint64_t nj, const skb_mode_group* groups, uint64_t group_count,
                                 const double* forward, const double* inverse, int precision,
                                 skb_vertical_plan** plan);
int skb_vertical_wvm_forward(skb_vertical_plan* plan, const void* full_spectrum, void* full_modal, void* stream);
int skb_vertical_wvm_inverse(skb_vertical_plan* plan, const void* full_modal, void* full_spectrum, void* stream);
/* WvmDirectVerticalGemmProvider::modalSpectrumElements: (Nx/2+1)*Ny*Nj*fields. */
uint64_t skb_vertical_wvm_modal_elements(const skb_vertical_plan* plan);
int skb_vertical_wvm_initialize_modal(skb_vertical_plan* plan, void* full_modal, void* stream);
int skb_vertical_wvm_initialize_spectrum(skb_vertical_plan* plan, void* full_spectrum, void* stream);

/* ---- synthetic modal work and the spectral pipeline ---------------------
 * syntheticModalWorkWeights (src/model.cpp:246-267): host, one real weight
 * per modal element at j + Nj*(field + fields*mode):
 * 0.75 + 0.5 * ((17|k| + 29|l| + 7j + 11 field) mod 97) / 96. */
int skb_synthetic_modal_weights(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                double* weights);
/* applySyntheticModalWorkInterleaved / Split (src/model.cpp:269-284) on the
 * device: output[e] = weights[e] * input[e] over `count` complex elements;
 * weights, input and output in `precision`.  In place is allowed. */
int skb_modal_work(int precision, uint64_t count, const void* weights, const void* input, void* output,
                   void* stream);
int skb_modal_work_split(int precision, uint64_t count, const void* weights, const void* input_real,
                         const void* input_imag, void* output_real, void* output_imag, void* stream);
/* applySyntheticModalWorkWvm (src/model.cpp:286-308) on a WVM-direct plan's
 * full modal arrays: the retained entries of the input scaled by the
 * weights (compact modal layout) into the output, and for k_x = 0,
 * k_y != 0, 2k_y != Ny the conjugate written at (0, Ny - k_y).  Out of
 * place; other output entries are left as they are. */
int skb_vertical_wvm_modal_work(skb_vertical_plan* plan, const void* weights, const void* full_modal_input,
                                void* full_modal_output, void* stream);

/* Synthetic antialiased spectral pipeline: the spectral-pipeline kernel of
 * runSpectralPipelineBenchmark (src/benchmark.cpp:2978-3547; timed sequence
 * 3389-3403).  One evaluation maps real planes to real planes:
 * unnormalised r2c -> retained modes -> vertical forward (K^2 family,
 * Nj x Nz per group, Nj = 2(Nz-1)/3) -> diagonal modal weights -> vertical
 * inverse -> zero-padded half spectrum -> unnormalised c2r.
 * policy SKB_PIPELINE_WVM_DIRECT follows the wvm-direct boundary
 * (src/benchmark.cpp:3204-3403): full WVM spectrum and modal arrays.
 * policy SKB_PIPELINE_COMPACT follows the compact split boundaries
 * (plane-major-fused-split, streaming-pruned-compact-split,
 * src/benchmark.cpp:3405-3547): pruned transforms and the compact modal
 * round trip, by default one kernel per tile (forward product, weights and
 * inverse product with the modal block in LDS) where it is instantiated
 * (Nz 5-8, 9-12, 17-20, 33-36, 65-68, 129-132; FP32 also 257-260).
 * weights: host array in the modal layout, or NULL for the synthetic
 * weights above.  Inputs are preserved; real_input may equal real_output. */
enum skb_pipeline_policy { SKB_PIPELINE_WVM_DIRECT = 0, SKB_PIPELINE_COMPACT = 1 };
typedef struct skb_spectral_pipeline skb_spectral_pipeline;
int skb_spectral_pipeline_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                 const skb_mode_group* groups, uint64_t group_count, const double* forward,
                                 const double* inverse, const double* weights, int policy, int precision,
                                 skb_spectral_pipeline** plan);
int skb_spectral_pipeline_destroy(skb_spectral_pipeline* plan);
int skb_spectral_pipeline_execute(skb_spectral_pipeline* plan, const void* real_input, void* real_output,
                                  void* stream);
/* The same evaluation stage by stage, writing the compact intermediates the
 * reference's correctness metrics read (src/benchmark.cpp:3175-3198):
 * retained horizontal coefficients and the inverse vertical output
 * (retained layout), the forward modal and the post-work modal (modal
 * layout).  Separate launches throughout (no fused round trip). */
int skb_spectral_pipeline_stages(skb_spectral_pipeline* plan, const void* real_input, void* real_output,
                                 void* retained, void* forward_modal, void* post_modal, void* inverse_retained,
                                 void* stream);
/* Compact policy: 1 = one-kernel modal round trip, 0 = three launches
 * through a modal array (allocated on first use).  No reference
 * counterpart. */
int skb_spectral_pipeline_set_fused(skb_spectral_pipeline* plan, int fused);
int skb_spectral_pipeline_fused(const skb_spectral_pipeline* plan);
/* Device memory the plan holds (arrays, matrices, scratch). */
uint64_t skb_spectral_pipeline_device_bytes(const skb_spectral_pipeline* plan);

/* ---- constant-stratification type-I vertical transforms -----------------
 * Replaces the six FFTW guru64 REDFT00/RODFT00 plans and the forward
 * normalisation of the constant-stratification vertical benchmark
 * (src/constant_stratification.cpp:95-147 R2RPlan, 196-235
 * normalizeForwardDct / normalizeForwardDst, 271-292
 * executeCompleteVerticalGraph).  Arrays are complex [rows][channels][nz],
 * z adjacent (element (row, channel, z) at z + nz*(channel + channels*row));
 * a transform applies, in place, to channels [first, first + count) of every
 * row, real and imaginary parts alike.  Kinds:
 *   SKB_TYPE1_DCT          REDFT00 of length nz (unnormalised)
 *   SKB_TYPE1_DST          RODFT00 of the nz - 2 interior points; the two
 *                          endpoints are left as they are
 *   SKB_TYPE1_DCT_FORWARD  DCT-I, then x 1/(nz-1) and the last point x 0.5
 *   SKB_TYPE1_DST_FORWARD  DST-I, then x 1/(nz-1), endpoints set to zero
 * Each is an nz x nz real matrix on the matrix cores (nz <= 132 FP64, 260
 * FP32). */
enum skb_type1_kind { SKB_TYPE1_DCT = 0, SKB_TYPE1_DST = 1, SKB_TYPE1_DCT_FORWARD = 2, SKB_TYPE1_DST_FORWARD = 3 };
typedef struct skb_type1_plan skb_type1_plan;
/* The transform of `kind` as a row-major nz x nz matrix (host). */
int skb_type1_matrix(uint64_t nz, int kind, double* matrix);
int skb_type1_plan_create(uint64_t rows, uint64_t nz, int precision, skb_type1_plan** plan);
int skb_type1_plan_destroy(skb_type1_plan* plan);
int skb_type1_execute(skb_type1_plan* plan, int kind, uint64_t channels, uint64_t first_channel, uint64_t count,
                      void* data, void* stream);
/* executeCompleteVerticalGraph: on a three-channel arena, 3 x (DCT of
 * channels 0-1, DST of channel 2) and 2 x (DST of channels 0-1, DCT of
 * channel 2); on a one-channel arena, 2 x forward DCT and 2 x forward DST
 * (15 inverse and 4 forward complex channel transforms). */
int skb_type1_vertical_schedule(skb_type1_plan* plan, void* arena3, void* arena1, void* stream);
/* The benchmark's deterministic fill (src/constant_stratification.cpp:237-269):
 * value(row, channel, z) = (sin a, cos 1.7a), a = 0.017(1 + row mod 97) +
 * 0.113(1 + channel) + 0.071(1 + z).  modes == NULL fills `rows` full-half
 * rows; otherwise row m holds the retained mode m's stored frequency
 * stored_kx + (Nx/2+1) stored_ky (nx_half = Nx/2+1), conjugated where the
 * mode conjugates its stored value.  Interleaved FP64 output. */
int skb_type1_fill(uint64_t rows, uint64_t channels, uint64_t nz, const skb_retained_mode* modes, uint64_t nx_half,
                   double* data);

/* ---- antialiased nonlinear flux graph ------------------------------------
 * Replaces the production-lifetime flux graph of runProductionLifetimeFluxBenchmark
 * (src/benchmark.cpp:3957-4234; serial oracle 4182-4234).  workload->fields
 * is the number of targets T; the input modal array has 3 + 3T fields
 * [U, V, W, q0x, q0y, q0z, ...] and the output modal array T fields, both in
 * the modal layout j + Nj*field + Nj*fields*mode.  One evaluation: vertical
 * inverse (K^2 family, Nj x Nz per group) -> pruned horizontal inverse ->
 * per target -(u*qx + v*qy + w*qz)/(Nx*Ny)^2 -> raw pruned forward ->
 * vertical forward.  The grid must be square (src/benchmark.cpp:3981-3984).
 * By default FP32 plans with Nx >= 64 run the physical-space stage as one
 * fused row pass (inverse rows, products and forward rows in registers);
 * skb_flux_set_fused(plan, 0), FP64 plans and smaller grids use separate
 * kernels with the real fields in HBM. */
typedef struct skb_flux_plan skb_flux_plan;
int skb_flux_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                         uint64_t nj, const skb_mode_group* groups, uint64_t group_count, const double* forward,
                         const double* inverse, int precision, skb_flux_plan** plan);
int skb_flux_plan_destroy(skb_flux_plan* plan);
int skb_flux_set_fused(skb_flux_plan* plan, int fused);
int skb_flux_execute(skb_flux_plan* plan, const void* input_modal, void* output_modal, void* stream);

/* Advection flux of one tracer, the graph above with spectral inputs (no
 * reference counterpart as a single call; it composes FFTWPrunedProvider
 * inverse / forward, a vertical product and PointwiseAdvectionExecutor).
 * workload->fields must be 4: the compact retained input holds u, v, w, q
 * (plane z + Nz*field).  q_x = i*2pi*k/lx q and q_y = i*2pi*l/ly q are formed
 * spectrally, q_z = D q with D the Nz x Nz row-major derivative matrix (on
 * MFMA), and the output is the compact retained field
 * F_xy[-(u q_x + v q_y + w q_z)] / (Nx*Ny)^2 of the unnormalised inverses,
 * truncated to the retained modes. */
typedef struct skb_advection_flux_plan skb_advection_flux_plan;
int skb_advection_flux_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                   const double* dz, int precision, skb_advection_flux_plan** plan);
int skb_advection_flux_plan_destroy(skb_advection_flux_plan* plan);
int skb_advection_flux_set_fused(skb_advection_flux_plan* plan, int fused);
/* Caps every kernel of the plan at `workgroups` workgroups, which then loop
 * over the work items (0 = one workgroup per work item, the default).  A
 * serving loop caps its large requests so that kernels of small requests on
 * other streams find free CUs.  Results are unchanged.  No reference
 * counterpart (the reference runs one workload per process). */
int skb_advection_flux_set_workgroups(skb_advection_flux_plan* plan, int workgroups);
int skb_advection_flux_execute(skb_advection_flux_plan* plan, const void* spectral, void* flux, void* stream);
/* Device memory the plan holds (intermediates and scratch), for admission
 * control by an HBM budget. */
uint64_t skb_advection_flux_device_bytes(const skb_advection_flux_plan* plan);
/* Second-order centred d/dz on Nz uniform levels over [0, lz], one-sided
 * second order at both ends (rows sum to zero); Nz x Nz row-major.  No
 * reference counterpart. */
int skb_vertical_derivative_matrix(uint64_t nz, double lz, double* d);

/* ---- batched uneven workload --------------------------------------------
 * Retained-operator evaluations (pruned inverse -> pruned forward -> grouped
 * vertical GEMM) over size classes of uneven grids, recorded once into a
 * hipGraph and replayed with one launch call.  Class c stacks `instances`
 * copies of `workload` as one workload with fields * instances fields
 * (instance i = fields [F*i, F*i + F)), so each stage is one launch over all
 * instances of the class.  No reference counterpart (the reference runs one
 * workload per process). */
typedef struct skb_batch skb_batch;
int skb_batch_create(int precision, skb_batch** batch);
int skb_batch_destroy(skb_batch* batch);
int skb_batch_add_class(skb_batch* batch, const skb_workload* workload, uint64_t instances,
                        const skb_retained_mode* modes, uint64_t mode_count, uint64_t nj,
                        const skb_mode_group* groups, uint64_t group_count, const double* forward,
                        const double* inverse, int* class_index);
/* Binds the class's caller-owned device arrays in the stacked layout: input
 * compact retained (plane + P*mode, P = Nz*F*instances) and output modal
 * (j + Nj*(field + F*instances*mode)).  Binding drops a captured graph. */
int skb_batch_bind(skb_batch* batch, int cls, const void* input, void* output);
/* One evaluation of every class, launched eagerly. */
int skb_batch_run(skb_batch* batch, void* stream);
/* Records one evaluation into a hipGraph (stream must not be the default
 * stream); skb_batch_launch replays it. */
int skb_batch_capture(skb_batch* batch, void* stream);
int skb_batch_launch(skb_batch* batch, void* stream);
/* Times every candidate of every stage of every class (reps launches each)
 * and keeps the fastest: each stage's launch shape and, for the vertical
 * product of FP32 square families, split-BF16 or FP32 products; writes a
 * JSON report.  Invalidates a captured graph.  Results agree with the
 * defaults to the FP32 tolerance (see skb_pruned_plan_set_grid,
 * skb_vertical_plan_set_products). */
int skb_batch_autotune(skb_batch* batch, int reps, void* stream, char* report, size_t capacity);
/* The batch's current choices as a spec string
 * ("c<class>.s<stage>=<grid>;...;c<class>.p=<products>;"), and a spec
 * applied (all entries checked before any is applied; drops a captured
 * graph): a kept tuning is reused without timing.  No reference
 * counterpart (the reference's worker-count tuner keeps a manifest,
 * tools/run_portable_machine_tuning.py:147, 1452-1497). */
int skb_batch_tuning(const skb_batch* batch, char* spec, size_t capacity);
int skb_batch_set_tuning(skb_batch* batch, const char* spec);

/* ---- allocation accounting ---------------------------------------------------
 * Device allocations (count, bytes) the library has made in this process,
 * for a steady-state check that repeated executions allocate nothing
 * (docs/benchmark-contract.md:98, 196; tests/allocation_tracker.cpp:42-89).
 * Allocation happens at plan creation, on the first call with a new
 * retained-mode list (gather / embed tables), and when an entry point that
 * documents it is first used (skb_flux_set_fused(plan, 0), the unfused
 * pipeline's modal array, skb_vertical_plan_reserve_split). */
uint64_t skb_device_allocations(void);
uint64_t skb_device_allocated_bytes(void);

/* ---- split <-> interleaved complex arrays ---------------------------------
 * The conversions under the reference's split-layout adapters
 * (FFTWProvider::forwardSplit / inverseSplit / gatherRetainedToSplitOuter /
 * embedRetainedFromSplitOuter / scaleRetainedSplitOuter /
 * gatherRetainedSplitOuter / embedRetainedSplitOuter,
 * include/skbench/skbench.hpp:374-398; src/fftw_provider.cpp:331-395,
 * 701-830): interleave writes out[i] = (real[i], imaginary[i]); deinterleave
 * writes real[i] = scale * in[i].re and imaginary[i] = scale * in[i].im;
 * scale_split multiplies both arrays in place.  count = complex elements. */
int skb_interleave(int precision, uint64_t count, const void* real, const void* imaginary, void* out, void* stream);
int skb_deinterleave(int precision, uint64_t count, const void* in, void* real, void* imaginary, double scale,
                     void* stream);
int skb_scale_split(int precision, uint64_t count, void* real, void* imaginary, double scale, void* stream);

/* ---- pointwise advection -----------------------------------------------
 * Replaces PointwiseAdvectionExecutor::execute (src/pointwise_advection.hpp
 * :18-43; src/pointwise_advection.cpp:152-166):
 * target = -scale * (u*qx + v*qy + w*qz) with shared = [u|v|w] and
 * derivative = [qx|qy|qz], each block `volume` elements. */
int skb_pointwise_advection(uint64_t volume, double scale, const void* shared, const void* derivative,
                            void* target, int precision, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SKB_HIP_H */
