"""ctypes binding of libskb_hip.so (the C ABI declared in include/skb_hip.h).

Device buffers are torch tensors (torch is only the allocator and the stream
provider here); every compute call goes through the C ABI into the
hand-written gfx950 kernels.  If the library is missing this module raises on
import -- there is no CPU fallback.

Class names mirror the reference providers they replace:
  FftPlan       <- skbench::FFTWProvider      (include/skbench/skbench.hpp:364-421)
  PrunedPlan    <- skbench::FFTWPrunedProvider (include/skbench/skbench.hpp:467-520)
  VerticalPlan  <- skbench::VerticalGemmProvider (include/skbench/skbench.hpp:152-218)
  OperatorPlan  <- the inverse -> forward -> vertical composition the
                   reference's harness times as separate provider calls
  pointwise_advection <- PointwiseAdvectionExecutor::execute (src/pointwise_advection.hpp:18-43)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "build", "libskb_hip.so")

SKB_F64, SKB_F32 = 0, 1
STATUS = {0: "ok", 1: "invalid_argument", 2: "runtime_error", 3: "logic_error", 4: "bad_alloc", 5: "unsupported"}


class SkbError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("%s: %s" % (STATUS.get(status, status), message))
        self.status = status


class InvalidArgument(SkbError, ValueError):
    pass


class LogicError(SkbError):
    """SKB_LOGIC_ERROR: wrong entry point for the plan (std::logic_error)."""


class Unsupported(SkbError):
    pass


class Workload_(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint64),
        ("ny", ctypes.c_uint64),
        ("nz", ctypes.c_uint64),
        ("fields", ctypes.c_uint64),
        ("lx", ctypes.c_double),
        ("ly", ctypes.c_double),
        ("antialias", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


MODE_DTYPE = np.dtype(
    [
        ("k", "<i8"),
        ("l", "<i8"),
        ("stored_kx", "<u8"),
        ("stored_ky", "<u8"),
        ("conj", "<i4"),
        ("reserved", "<i4"),
        ("radial", "<f8"),
    ]
)
class FieldView_(ctypes.Structure):
    """skb_field_view: one field of a split compact array (device pointers)."""
    _fields_ = [("real", ctypes.c_void_p), ("imaginary", ctypes.c_void_p), ("mode_stride", ctypes.c_uint64)]


GROUP_DTYPE = np.dtype([("squared_key", "<u8"), ("first_mode", "<u8"), ("mode_count", "<u8")])

FIXTURES = {"impulse": 0, "sinusoid": 1, "random": 2, "dc": 3, "nyquist": 4, "sinusoid-5-17": 5}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libskb_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    # torch carries its own HIP runtime (SONAME libamdhip64.so.7, loaded by
    # file name).  Loading torch first lets the library's NEEDED entry bind to
    # that same runtime, so the process has ONE HIP runtime and device
    # pointers and streams from torch are valid in the library.
    import torch  # noqa: F401

    L = ctypes.CDLL(LIB_PATH)
    vp, u64, i32, dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    W = ctypes.POINTER(Workload_)
    pp = ctypes.POINTER(ctypes.c_void_p)
    sigs = {
        "skb_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
        "skb_version": (ctypes.c_char_p, []),
        "skb_retained_modes": (i32, [W, vp, u64, ctypes.POINTER(u64)]),
        "skb_k2_groups": (i32, [vp, u64, vp, u64, ctypes.POINTER(u64)]),
        "skb_k2_binned_groups": (i32, [vp, u64, u64, vp, u64, ctypes.POINTER(u64)]),
        "skb_mode_order_hash": (u64, [vp, u64]),
        "skb_group_order_hash": (u64, [vp, u64]),
        "skb_wvm_order_hash": (u64, [W]),
        "skb_orthonormal_vertical": (i32, [u64, u64, vp, vp]),
        "skb_k2_vertical_family": (i32, [u64, u64, vp, u64, vp, vp]),
        "skb_qr_gaussian_family": (i32, [u64, u64, u64, vp, vp]),
        "skb_make_fixture": (i32, [W, i32, u64, vp]),
        "skb_spectral_fixture": (i32, [W, vp, u64, u64, dbl, vp]),
        "skb_fft_plan_create": (i32, [W, i32, pp]),
        "skb_fft_plan_destroy": (i32, [vp]),
        "skb_fft_forward": (i32, [vp, vp, vp, vp]),
        "skb_fft_inverse": (i32, [vp, vp, vp, vp]),
        "skb_fft_gather_retained": (i32, [vp, vp, u64, vp, vp, vp]),
        "skb_fft_embed_retained": (i32, [vp, vp, u64, vp, vp, vp]),
        "skb_pruned_plan_create": (i32, [W, vp, u64, i32, pp]),
        "skb_pruned_plan_destroy": (i32, [vp]),
        "skb_pruned_forward": (i32, [vp, vp, vp, vp]),
        "skb_pruned_inverse": (i32, [vp, vp, vp, vp]),
        "skb_pruned_forward_split": (i32, [vp, vp, vp, vp, dbl, vp]),
        "skb_pruned_inverse_split": (i32, [vp, vp, vp, vp, vp]),
        "skb_pruned_forward_split_fields": (i32, [vp, vp, vp, u64, dbl, vp]),
        "skb_pruned_inverse_split_fields": (i32, [vp, vp, u64, vp, vp]),
        "skb_pruned_forward_rows": (i32, [vp, vp, vp]),
        "skb_pruned_forward_columns": (i32, [vp, vp, vp]),
        "skb_pruned_inverse_columns": (i32, [vp, vp, vp]),
        "skb_pruned_inverse_rows": (i32, [vp, vp, vp]),
        "skb_pruned_active_kx": (u64, [vp]),
        "skb_pruned_plan_set_grid": (i32, [vp, i32, i32]),
        "skb_pruned_plan_items": (u64, [vp, i32]),
        "skb_vertical_plan_set_grid": (i32, [vp, i32]),
        "skb_vertical_plan_items": (u64, [vp]),
        "skb_vertical_plan_set_products": (i32, [vp, i32]),
        "skb_vertical_plan_products": (i32, [vp, i32]),
        "skb_pruned_scratch_bytes": (u64, [vp]),
        "skb_vertical_plan_create": (i32, [W, u64, vp, u64, vp, vp, i32, pp]),
        "skb_vertical_plan_destroy": (i32, [vp]),
        "skb_vertical_forward": (i32, [vp, vp, vp, vp]),
        "skb_vertical_inverse": (i32, [vp, vp, vp, vp]),
        "skb_vertical_matrix_bytes": (u64, [vp]),
        "skb_pointwise_advection": (i32, [u64, dbl, vp, vp, vp, i32, vp]),
        "skb_vertical_wvm_plan_create": (i32, [W, vp, u64, u64, vp, u64, vp, vp, i32, pp]),
        "skb_flux_plan_create": (i32, [W, vp, u64, u64, vp, u64, vp, vp, i32, pp]),
        "skb_flux_plan_destroy": (i32, [vp]),
        "skb_flux_set_fused": (i32, [vp, i32]),
        "skb_flux_execute": (i32, [vp, vp, vp, vp]),
        "skb_advection_flux_plan_create": (i32, [W, vp, u64, vp, i32, pp]),
        "skb_advection_flux_plan_destroy": (i32, [vp]),
        "skb_advection_flux_set_fused": (i32, [vp, i32]),
        "skb_advection_flux_set_workgroups": (i32, [vp, i32]),
        "skb_advection_flux_execute": (i32, [vp, vp, vp, vp]),
        "skb_vertical_derivative_matrix": (i32, [u64, dbl, vp]),
        "skb_advection_flux_device_bytes": (u64, [vp]),
        "skb_batch_create": (i32, [i32, pp]),
        "skb_batch_destroy": (i32, [vp]),
        "skb_batch_add_class": (i32, [vp, W, u64, vp, u64, u64, vp, u64, vp, vp, ctypes.POINTER(i32)]),
        "skb_batch_bind": (i32, [vp, i32, vp, vp]),
        "skb_batch_run": (i32, [vp, vp]),
        "skb_batch_capture": (i32, [vp, vp]),
        "skb_batch_launch": (i32, [vp, vp]),
        "skb_batch_autotune": (i32, [vp, i32, vp, ctypes.c_char_p, ctypes.c_size_t]),
        "skb_batch_tuning": (i32, [vp, ctypes.c_char_p, ctypes.c_size_t]),
        "skb_batch_set_tuning": (i32, [vp, ctypes.c_char_p]),
        "skb_vertical_wvm_forward": (i32, [vp, vp, vp, vp]),
        "skb_vertical_wvm_inverse": (i32, [vp, vp, vp, vp]),
        "skb_vertical_wvm_modal_elements": (u64, [vp]),
        "skb_vertical_wvm_initialize_modal": (i32, [vp, vp, vp]),
        "skb_vertical_wvm_initialize_spectrum": (i32, [vp, vp, vp]),
        "skb_fft_gather_retained_modal": (i32, [vp, vp, u64, u64, vp, vp, vp]),
        "skb_synthetic_modal_weights": (i32, [W, vp, u64, vp]),
        "skb_modal_work": (i32, [i32, u64, vp, vp, vp, vp]),
        "skb_modal_work_split": (i32, [i32, u64, vp, vp, vp, vp, vp, vp]),
        "skb_vertical_wvm_modal_work": (i32, [vp, vp, vp, vp, vp]),
        "skb_spectral_pipeline_create": (i32, [W, vp, u64, vp, u64, vp, vp, vp, i32, i32, pp]),
        "skb_spectral_pipeline_destroy": (i32, [vp]),
        "skb_spectral_pipeline_execute": (i32, [vp, vp, vp, vp]),
        "skb_spectral_pipeline_stages": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "skb_spectral_pipeline_set_fused": (i32, [vp, i32]),
        "skb_spectral_pipeline_fused": (i32, [vp]),
        "skb_spectral_pipeline_device_bytes": (u64, [vp]),
        "skb_type1_matrix": (i32, [u64, i32, vp]),
        "skb_type1_plan_create": (i32, [u64, u64, i32, pp]),
        "skb_type1_plan_destroy": (i32, [vp]),
        "skb_type1_execute": (i32, [vp, i32, u64, u64, u64, vp, vp]),
        "skb_type1_vertical_schedule": (i32, [vp, vp, vp, vp]),
        "skb_type1_fill": (i32, [u64, u64, u64, vp, u64, vp]),
        "skb_pruned_plan_set_path": (i32, [vp, i32]),
        "skb_pruned_plan_path": (i32, [vp]),
        "skb_retained_operator_create": (i32, [W, vp, u64, u64, vp, u64, vp, vp, i32, pp]),
        "skb_retained_operator_destroy": (i32, [vp]),
        "skb_retained_operator_execute": (i32, [vp, vp, vp, vp]),
        "skb_retained_operator_stage": (i32, [vp, i32, vp, vp, vp]),
        "skb_retained_operator_reserve_lanes": (i32, [vp, i32]),
        "skb_retained_operator_lanes": (i32, [vp, ctypes.POINTER(i32)]),
        "skb_retained_operator_execute_batch": (i32, [vp, i32, vp, vp, vp]),
        "skb_retained_operator_reserve_batch": (i32, [vp, i32]),
        "skb_retained_operator_batch": (i32, [vp]),
        "skb_retained_operator_execute_batch_stage": (i32, [vp, i32, i32, vp, vp, vp]),
        "skb_retained_operator_set_path": (i32, [vp, i32]),
        "skb_retained_operator_path": (i32, [vp]),
        "skb_retained_operator_real": (vp, [vp]),
        "skb_retained_operator_device_bytes": (u64, [vp]),
        "skb_retained_operator_set_fused": (i32, [vp, i32]),
        "skb_retained_operator_fused": (i32, [vp]),
        "skb_interleave": (i32, [i32, u64, vp, vp, vp, vp]),
        "skb_deinterleave": (i32, [i32, u64, vp, vp, vp, dbl, vp]),
        "skb_scale_split": (i32, [i32, u64, vp, vp, dbl, vp]),
        "skb_vertical_plan_reserve_split": (i32, [vp]),
        "skb_vertical_forward_split": (i32, [vp, vp, vp, vp, vp, vp]),
        "skb_vertical_inverse_split": (i32, [vp, vp, vp, vp, vp, vp]),
        "skb_vertical_forward_split_noalloc": (i32, [vp, vp, vp, vp, vp, vp]),
        "skb_fft_forward_view": (i32, [vp, vp, vp, u64, vp]),
        "skb_fft_inverse_view": (i32, [vp, vp, u64, vp, vp]),
        "skb_device_allocations": (u64, []),
        "skb_device_allocated_bytes": (u64, []),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()
EXPORTED = [
    "skb_last_error", "skb_version", "skb_retained_modes", "skb_k2_groups", "skb_k2_binned_groups",
    "skb_mode_order_hash", "skb_group_order_hash", "skb_wvm_order_hash", "skb_orthonormal_vertical",
    "skb_k2_vertical_family", "skb_qr_gaussian_family", "skb_make_fixture", "skb_spectral_fixture",
    "skb_fft_plan_create", "skb_fft_plan_destroy", "skb_fft_forward", "skb_fft_inverse",
    "skb_fft_gather_retained", "skb_fft_embed_retained", "skb_pruned_plan_create", "skb_pruned_plan_destroy",
    "skb_pruned_forward", "skb_pruned_inverse", "skb_pruned_active_kx", "skb_pruned_scratch_bytes",
    "skb_pruned_forward_rows", "skb_pruned_forward_columns", "skb_pruned_inverse_columns", "skb_pruned_inverse_rows",
    "skb_vertical_plan_create", "skb_vertical_plan_destroy", "skb_vertical_forward", "skb_vertical_inverse",
    "skb_vertical_matrix_bytes", "skb_pointwise_advection", "skb_pruned_plan_set_grid", "skb_pruned_plan_items",
    "skb_vertical_plan_set_products", "skb_vertical_plan_products",
    "skb_vertical_plan_set_grid", "skb_vertical_plan_items", "skb_pruned_forward_split", "skb_pruned_inverse_split",
    "skb_pruned_forward_split_fields", "skb_pruned_inverse_split_fields", "skb_vertical_wvm_plan_create",
    "skb_vertical_wvm_forward", "skb_vertical_wvm_inverse", "skb_vertical_wvm_modal_elements",
    "skb_vertical_wvm_initialize_modal", "skb_vertical_wvm_initialize_spectrum", "skb_flux_plan_create",
    "skb_flux_plan_destroy", "skb_flux_set_fused", "skb_flux_execute", "skb_advection_flux_plan_create",
    "skb_advection_flux_plan_destroy", "skb_advection_flux_set_fused", "skb_advection_flux_set_workgroups",
    "skb_advection_flux_execute",
    "skb_vertical_derivative_matrix", "skb_batch_create", "skb_batch_destroy", "skb_batch_add_class",
    "skb_batch_bind", "skb_batch_run", "skb_batch_capture", "skb_batch_launch", "skb_batch_autotune",
    "skb_batch_tuning", "skb_batch_set_tuning",
    "skb_advection_flux_device_bytes", "skb_fft_gather_retained_modal", "skb_synthetic_modal_weights",
    "skb_modal_work", "skb_modal_work_split", "skb_vertical_wvm_modal_work", "skb_spectral_pipeline_create",
    "skb_spectral_pipeline_destroy", "skb_spectral_pipeline_execute", "skb_spectral_pipeline_stages",
    "skb_spectral_pipeline_set_fused", "skb_spectral_pipeline_fused", "skb_spectral_pipeline_device_bytes",
    "skb_type1_matrix", "skb_type1_plan_create", "skb_type1_plan_destroy", "skb_type1_execute",
    "skb_type1_vertical_schedule", "skb_type1_fill", "skb_pruned_plan_set_path", "skb_pruned_plan_path",
    "skb_retained_operator_create", "skb_retained_operator_destroy", "skb_retained_operator_execute",
    "skb_retained_operator_stage", "skb_retained_operator_set_path", "skb_retained_operator_path",
    "skb_retained_operator_reserve_lanes", "skb_retained_operator_lanes", "skb_retained_operator_execute_batch",
    "skb_retained_operator_reserve_batch", "skb_retained_operator_batch", "skb_retained_operator_execute_batch_stage",
    "skb_retained_operator_real", "skb_retained_operator_device_bytes", "skb_retained_operator_set_fused",
    "skb_retained_operator_fused", "skb_interleave", "skb_deinterleave", "skb_scale_split",
    "skb_vertical_plan_reserve_split", "skb_vertical_forward_split", "skb_vertical_inverse_split",
    "skb_vertical_forward_split_noalloc", "skb_fft_forward_view", "skb_fft_inverse_view", "skb_device_allocations",
    "skb_device_allocated_bytes",
]


def device_allocations() -> tuple:
    """(count, bytes) of the device allocations the library has made."""
    return int(lib.skb_device_allocations()), int(lib.skb_device_allocated_bytes())


def last_error() -> str:
    buf = ctypes.create_string_buffer(4096)
    lib.skb_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int):
    if status == 0:
        return
    msg = last_error()
    if status == 1:
        raise InvalidArgument(status, msg)
    if status == 3:
        raise LogicError(status, msg)
    if status == 5:
        raise Unsupported(status, msg)
    raise SkbError(status, msg)


def version() -> str:
    return lib.skb_version().decode()


@dataclass(frozen=True)
class Workload:
    """skbench::Workload (include/skbench/skbench.hpp:23-39)."""

    nx: int
    ny: int
    nz: int
    fields: int
    lx: float = 1.0
    ly: float = 1.0
    antialias: bool = True

    @property
    def planes(self):
        return self.nz * self.fields

    @property
    def nx_half(self):
        return self.nx // 2 + 1

    @property
    def real_elements(self):
        return self.nx * self.ny * self.planes

    @property
    def spectrum_elements(self):
        return self.nx_half * self.ny * self.planes

    @property
    def nj(self):
        return 0 if self.nz < 2 else 2 * (self.nz - 1) // 3

    def c(self) -> Workload_:
        return Workload_(self.nx, self.ny, self.nz, self.fields, self.lx, self.ly, int(self.antialias), 0)


def _np(a):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


# ---- host model ---------------------------------------------------------------
SKB_PIPELINE_WVM_DIRECT = 0
SKB_PIPELINE_COMPACT = 1
SKB_PRODUCTS_FP32 = 0
SKB_PRODUCTS_SPLIT_BF16 = 1
SKB_PATH_AUTO, SKB_PATH_TWO_PASS, SKB_PATH_PLANE = 0, 1, 2
SKB_OP_STAGE_LAYOUT, SKB_OP_STAGE_INVERSE, SKB_OP_STAGE_FORWARD, SKB_OP_STAGE_VERTICAL = 0, 1, 2, 3
SKB_OP_STAGE_TRANSFORMS = 4


def synthetic_modal_weights(w: "Workload", modes: np.ndarray) -> np.ndarray:
    """syntheticModalWorkWeights (src/model.cpp:246-267), modal layout."""
    out = np.zeros(len(modes) * w.fields * w.nj)
    wc = w.c()
    check(lib.skb_synthetic_modal_weights(ctypes.byref(wc), _np(modes), len(modes), _np(out)))
    return out


def retained_modes(w: Workload) -> np.ndarray:
    wc = w.c()
    n = ctypes.c_uint64()
    check(lib.skb_retained_modes(ctypes.byref(wc), None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, MODE_DTYPE)
    check(lib.skb_retained_modes(ctypes.byref(wc), _np(out), n.value, ctypes.byref(n)))
    return out


def k2_groups(modes: np.ndarray) -> np.ndarray:
    out = np.zeros(len(modes), GROUP_DTYPE)
    n = ctypes.c_uint64()
    check(lib.skb_k2_groups(_np(modes), len(modes), _np(out), len(out), ctypes.byref(n)))
    return out[: n.value].copy()


def k2_binned_groups(modes: np.ndarray, bins: int = 32) -> np.ndarray:
    out = np.zeros(len(modes), GROUP_DTYPE)
    n = ctypes.c_uint64()
    check(lib.skb_k2_binned_groups(_np(modes), len(modes), bins, _np(out), len(out), ctypes.byref(n)))
    return out[: n.value].copy()


def mode_order_hash(modes) -> str:
    return "fnv1a64:%016x" % lib.skb_mode_order_hash(_np(modes), len(modes))


def group_order_hash(groups) -> str:
    return "fnv1a64:%016x" % lib.skb_group_order_hash(_np(groups), len(groups))


def wvm_order_hash(w: Workload) -> str:
    wc = w.c()
    return "fnv1a64:%016x" % lib.skb_wvm_order_hash(ctypes.byref(wc))


def orthonormal_vertical(nz: int, nj: int):
    f = np.zeros(nj * nz)
    i = np.zeros(nz * nj)
    check(lib.skb_orthonormal_vertical(nz, nj, _np(f), _np(i)))
    return f, i


def k2_vertical_family(nz: int, nj: int, groups: np.ndarray):
    f = np.zeros(len(groups) * nj * nz)
    i = np.zeros(len(groups) * nz * nj)
    check(lib.skb_k2_vertical_family(nz, nj, _np(groups), len(groups), _np(f), _np(i)))
    return f, i


def qr_gaussian_family(n: int, groups: int, seed: int):
    f = np.zeros(groups * n * n)
    i = np.zeros(groups * n * n)
    check(lib.skb_qr_gaussian_family(n, groups, seed, _np(f), _np(i)))
    return f, i


def make_fixture(w: Workload, kind: str, seed: int) -> np.ndarray:
    out = np.zeros(w.real_elements)
    wc = w.c()
    check(lib.skb_make_fixture(ctypes.byref(wc), FIXTURES[kind], seed, _np(out)))
    return out


def spectral_fixture(w: Workload, modes: np.ndarray, seed: int, exponent: float = -1.0) -> np.ndarray:
    out = np.zeros(2 * len(modes) * w.planes)
    wc = w.c()
    check(lib.skb_spectral_fixture(ctypes.byref(wc), _np(modes), len(modes), seed, exponent, _np(out)))
    return out.view(np.complex128)


# ---- device helpers -------------------------------------------------------------
def _torch():
    import torch

    return torch


def real_dtype(precision):
    t = _torch()
    return t.float64 if precision == SKB_F64 else t.float32


def complex_dtype(precision):
    t = _torch()
    return t.complex128 if precision == SKB_F64 else t.complex64


def _dev(t, numel: int, dtype, name: str):
    """Validate a device tensor on the host before any launch."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if t.device.type != "cuda":
        raise ValueError("%s must be a device tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s has dtype %s, expected %s" % (name, t.dtype, dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if t.numel() != numel:
        raise ValueError("%s has %d elements, expected %d" % (name, t.numel(), numel))
    return ctypes.c_void_p(t.data_ptr())


class _DevicePointer:
    """A device allocation owned by a plan, shown to torch through the
    __cuda_array_interface__ protocol (no copy; `owner` keeps it alive)."""

    def __init__(self, ptr: int, numel: int, typestr: str, owner):
        self.owner = owner
        self.__cuda_array_interface__ = {"shape": (numel,), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


def _device_view(ptr, numel: int, dtype, owner):
    torch = _torch()
    if not ptr:
        raise LogicError(3, "null device pointer")
    typestr = {torch.float32: "<f4", torch.float64: "<f8"}[dtype]
    return torch.as_tensor(_DevicePointer(int(ptr), numel, typestr, owner), device="cuda")


def _dev_opt(t, numel: int, dtype, name: str):
    """_dev, or a null pointer for None (a skipped split component)."""
    return ctypes.c_void_p(0) if t is None else _dev(t, numel, dtype, name)


def interleave(precision: int, real, imaginary, out, stream=None):
    """out[i] = real[i] + 1j * imaginary[i] on the device (skb_interleave)."""
    n = out.numel()
    rd = real_dtype(precision)
    check(lib.skb_interleave(precision, n, _dev(real, n, rd, "real"), _dev(imaginary, n, rd, "imaginary"),
                             _dev(out, n, complex_dtype(precision), "out"), _stream(stream)))


def deinterleave(precision: int, inp, real, imaginary, scale: float = 1.0, stream=None):
    """real[i], imaginary[i] = scale * inp[i] parts (skb_deinterleave; None skips a part)."""
    n = inp.numel()
    rd = real_dtype(precision)
    check(lib.skb_deinterleave(precision, n, _dev(inp, n, complex_dtype(precision), "in"),
                               _dev_opt(real, n, rd, "real"), _dev_opt(imaginary, n, rd, "imaginary"), scale,
                               _stream(stream)))


def scale_split(precision: int, real, imaginary, scale: float, stream=None):
    n = real.numel()
    rd = real_dtype(precision)
    check(lib.skb_scale_split(precision, n, _dev(real, n, rd, "real"), _dev(imaginary, n, rd, "imaginary"), scale,
                              _stream(stream)))


def _stream(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Bound:
    """One C-ABI call with its arguments validated and converted once: calling
    the object launches it again on the same buffers and stream, with no
    per-call checks on the host (the serving loop's form of a stage).  The
    tensors are kept alive with the object; the last entry of `keep` is the
    plan that owns the handle, and a call after that plan was closed raises
    instead of passing a freed handle to the library."""

    __slots__ = ("fn", "args", "keep", "owner")

    def __init__(self, fn, args, keep):
        self.fn, self.args, self.keep = fn, args, keep
        self.owner = keep[-1]

    def __call__(self):
        if self.owner.h is None:
            raise LogicError(3, "the plan of this bound call was closed")
        status = self.fn(*self.args)
        if status:
            check(status)


class FftPlan:
    """Full batched 2-D r2c / c2r in WVM order (FFTWProvider)."""

    def __init__(self, w: Workload, precision: int = SKB_F64):
        self.w, self.precision = w, precision
        self._wc = w.c()
        h = ctypes.c_void_p()
        check(lib.skb_fft_plan_create(ctypes.byref(self._wc), precision, ctypes.byref(h)))
        self.h = h

    def forward(self, real, wvm, stream=None):
        w = self.w
        check(lib.skb_fft_forward(self.h, _dev(real, w.real_elements, real_dtype(self.precision), "real"),
                                  _dev(wvm, w.spectrum_elements, complex_dtype(self.precision), "wvm"), _stream(stream)))

    def inverse(self, wvm, real, stream=None):
        """Destroys `wvm` (FFTW c2r semantics)."""
        w = self.w
        check(lib.skb_fft_inverse(self.h, _dev(wvm, w.spectrum_elements, complex_dtype(self.precision), "wvm"),
                                  _dev(real, w.real_elements, real_dtype(self.precision), "real"), _stream(stream)))

    def _view(self, base, offset: int, stride: int):
        """Pointer to a WVM field view inside `base` (a device complex array):
        element (kx, ky, plane) at base[offset + plane + stride * (kx + nx_half * ky)]."""
        w = self.w
        cd = complex_dtype(self.precision)
        _dev(base, base.numel(), cd, "spectrum")
        last = offset + w.planes - 1 + stride * (w.nx_half * w.ny - 1)
        if offset < 0 or stride < w.planes or last >= base.numel():
            raise InvalidArgument(1, "field view outside its array")
        return ctypes.c_void_p(base.data_ptr() + offset * base.element_size())

    def forward_view(self, real, base, offset: int, stride: int, stream=None):
        """FFTWInterleavedWvmFieldViewProvider::forwardField into a view of `base`."""
        w = self.w
        check(lib.skb_fft_forward_view(self.h, _dev(real, w.real_elements, real_dtype(self.precision), "real"),
                                       self._view(base, offset, stride), stride, _stream(stream)))

    def inverse_view(self, base, offset: int, stride: int, real, stream=None):
        """FFTWInterleavedWvmFieldViewProvider::inverseFields for one view (may overwrite it)."""
        w = self.w
        check(lib.skb_fft_inverse_view(self.h, self._view(base, offset, stride), stride,
                                       _dev(real, w.real_elements, real_dtype(self.precision), "real"),
                                       _stream(stream)))

    def gather_retained(self, modes, wvm, retained, stream=None):
        w = self.w
        check(lib.skb_fft_gather_retained(
            self.h, _np(modes), len(modes), _dev(wvm, w.spectrum_elements, complex_dtype(self.precision), "wvm"),
            _dev(retained, len(modes) * w.planes, complex_dtype(self.precision), "retained"), _stream(stream)))

    def gather_retained_modal(self, modes, full_modal, modal, stream=None):
        """gatherRetainedModal: full WVM-order modal -> compact modal."""
        w, cd = self.w, complex_dtype(self.precision)
        nj = w.nj
        check(lib.skb_fft_gather_retained_modal(
            self.h, _np(modes), len(modes), nj, _dev(full_modal, w.nx_half * w.ny * w.fields * nj, cd, "full_modal"),
            _dev(modal, len(modes) * w.fields * nj, cd, "modal"), _stream(stream)))

    def embed_retained(self, modes, retained, wvm, stream=None):
        w = self.w
        check(lib.skb_fft_embed_retained(
            self.h, _np(modes), len(modes),
            _dev(retained, len(modes) * w.planes, complex_dtype(self.precision), "retained"),
            _dev(wvm, w.spectrum_elements, complex_dtype(self.precision), "wvm"), _stream(stream)))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_fft_plan_destroy(self.h)
            self.h = None

    __del__ = close


class PrunedPlan:
    """Pruned retained-mode transforms (FFTWPrunedProvider)."""

    def __init__(self, w: Workload, modes: np.ndarray, precision: int = SKB_F64):
        self.w, self.precision, self.modes = w, precision, modes
        self._wc = w.c()
        h = ctypes.c_void_p()
        check(lib.skb_pruned_plan_create(ctypes.byref(self._wc), _np(modes), len(modes), precision, ctypes.byref(h)))
        self.h = h

    @property
    def retained_elements(self):
        return len(self.modes) * self.w.planes

    def forward(self, real, retained, stream=None):
        check(lib.skb_pruned_forward(self.h, _dev(real, self.w.real_elements, real_dtype(self.precision), "real"),
                                     _dev(retained, self.retained_elements, complex_dtype(self.precision), "retained"),
                                     _stream(stream)))

    def inverse(self, retained, real, stream=None):
        check(lib.skb_pruned_inverse(self.h,
                                     _dev(retained, self.retained_elements, complex_dtype(self.precision), "retained"),
                                     _dev(real, self.w.real_elements, real_dtype(self.precision), "real"),
                                     _stream(stream)))

    # split compact form (FFTWPrunedProvider / FFTWStreamingPrunedSplitProvider)
    def forward_split(self, real, retained_real, retained_imag, scale=1.0, stream=None):
        rd = real_dtype(self.precision)
        check(lib.skb_pruned_forward_split(
            self.h, _dev(real, self.w.real_elements, rd, "real"),
            _dev(retained_real, self.retained_elements, rd, "retained_real"),
            _dev(retained_imag, self.retained_elements, rd, "retained_imag"), float(scale), _stream(stream)))

    def inverse_split(self, retained_real, retained_imag, real, stream=None):
        rd = real_dtype(self.precision)
        check(lib.skb_pruned_inverse_split(
            self.h, _dev(retained_real, self.retained_elements, rd, "retained_real"),
            _dev(retained_imag, self.retained_elements, rd, "retained_imag"),
            _dev(real, self.w.real_elements, rd, "real"), _stream(stream)))

    def _views(self, views):
        """[(real tensor, imag tensor, mode_stride)] -> skb_field_view array.
        Each tensor must cover element (Nz - 1) + mode_stride * (modes - 1)."""
        rd = real_dtype(self.precision)
        arr = (FieldView_ * max(1, len(views)))()
        for f, (re, im, stride) in enumerate(views):
            stride = int(stride)
            if stride < self.w.nz:
                raise InvalidArgument(1, "field view mode stride below Nz")
            need = self.w.nz + stride * (len(self.modes) - 1)
            for t, nm in ((re, "field real"), (im, "field imaginary")):
                torch = _torch()
                if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != rd:
                    raise TypeError("%s must be a %s device tensor" % (nm, rd))
                if not t.is_contiguous() or t.numel() < need:
                    raise ValueError("%s must be contiguous with >= %d elements" % (nm, need))
            arr[f] = FieldView_(re.data_ptr(), im.data_ptr(), stride)
        return arr

    def forward_split_fields(self, real, views, scale=1.0, stream=None):
        arr = self._views(views)
        check(lib.skb_pruned_forward_split_fields(
            self.h, _dev(real, self.w.real_elements, real_dtype(self.precision), "real"), arr, len(views),
            float(scale), _stream(stream)))

    def inverse_split_fields(self, views, real, stream=None):
        arr = self._views(views)
        check(lib.skb_pruned_inverse_split_fields(
            self.h, arr, len(views), _dev(real, self.w.real_elements, real_dtype(self.precision), "real"),
            _stream(stream)))

    # stage-by-stage form (intermediate in the plan's scratch)
    def forward_rows(self, real, stream=None):
        check(lib.skb_pruned_forward_rows(
            self.h, _dev(real, self.w.real_elements, real_dtype(self.precision), "real"), _stream(stream)))

    def forward_columns(self, retained, stream=None):
        check(lib.skb_pruned_forward_columns(
            self.h, _dev(retained, self.retained_elements, complex_dtype(self.precision), "retained"), _stream(stream)))

    def inverse_columns(self, retained, stream=None):
        check(lib.skb_pruned_inverse_columns(
            self.h, _dev(retained, self.retained_elements, complex_dtype(self.precision), "retained"), _stream(stream)))

    def inverse_rows(self, real, stream=None):
        check(lib.skb_pruned_inverse_rows(
            self.h, _dev(real, self.w.real_elements, real_dtype(self.precision), "real"), _stream(stream)))

    STAGES = ("forward_rows", "forward_columns", "inverse_columns", "inverse_rows")

    def bind_stage(self, name: str, tensor, stream=None) -> Bound:
        """A prepared launch of one stage (see Bound) on `tensor`: the real
        planes for the row stages, the compact retained array for the column
        stages."""
        if name not in self.STAGES:
            raise InvalidArgument(1, "unknown stage %r" % name)
        if name.endswith("rows"):
            ptr = _dev(tensor, self.w.real_elements, real_dtype(self.precision), "real")
        else:
            ptr = _dev(tensor, self.retained_elements, complex_dtype(self.precision), "retained")
        return Bound(getattr(lib, "skb_pruned_" + name), (self.h, ptr, _stream(stream)), (tensor, self))

    def bind_forward(self, real, retained, stream=None) -> Bound:
        """A prepared forward operator call (rows, then columns; see Bound)."""
        args = (self.h, _dev(real, self.w.real_elements, real_dtype(self.precision), "real"),
                _dev(retained, self.retained_elements, complex_dtype(self.precision), "retained"), _stream(stream))
        return Bound(lib.skb_pruned_forward, args, (real, retained, self))

    def bind_inverse(self, retained, real, stream=None) -> Bound:
        """A prepared inverse operator call (columns, then rows; see Bound)."""
        args = (self.h, _dev(retained, self.retained_elements, complex_dtype(self.precision), "retained"),
                _dev(real, self.w.real_elements, real_dtype(self.precision), "real"), _stream(stream))
        return Bound(lib.skb_pruned_inverse, args, (retained, real, self))

    def set_grid(self, stage: int, workgroups: int):
        """Launch shape of one stage (0 = one workgroup per work item)."""
        check(lib.skb_pruned_plan_set_grid(self.h, stage, workgroups))

    def set_path(self, path: int):
        """SKB_PATH_AUTO (default), SKB_PATH_TWO_PASS or SKB_PATH_PLANE."""
        check(lib.skb_pruned_plan_set_path(self.h, path))

    def path(self) -> int:
        """The path forward / inverse take: SKB_PATH_PLANE or SKB_PATH_TWO_PASS."""
        return int(lib.skb_pruned_plan_path(self.h))

    def items(self, stage: int) -> int:
        return int(lib.skb_pruned_plan_items(self.h, stage))

    def active_kx(self) -> int:
        return int(lib.skb_pruned_active_kx(self.h))

    def scratch_bytes(self) -> int:
        return int(lib.skb_pruned_scratch_bytes(self.h))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_pruned_plan_destroy(self.h)
            self.h = None

    __del__ = close


class VerticalPlan:
    """K^2-grouped vertical GEMM on MFMA (VerticalGemmProvider)."""

    def __init__(self, w: Workload, nj: int, groups: np.ndarray, forward: np.ndarray, inverse: np.ndarray,
                 precision: int = SKB_F64):
        self.w, self.nj, self.precision = w, nj, precision
        self.modes = int(groups["first_mode"][-1] + groups["mode_count"][-1])
        self._wc = w.c()
        forward = np.ascontiguousarray(forward, np.float64)
        inverse = np.ascontiguousarray(inverse, np.float64)
        if forward.size != len(groups) * nj * w.nz or inverse.size != forward.size:
            raise ValueError("matrix family size does not match groups x nj x nz")
        h = ctypes.c_void_p()
        check(lib.skb_vertical_plan_create(ctypes.byref(self._wc), nj, _np(groups), len(groups), _np(forward),
                                           _np(inverse), precision, ctypes.byref(h)))
        self.h = h

    def forward(self, physical, modal, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_vertical_forward(self.h, _dev(physical, self.modes * self.w.planes, cd, "physical"),
                                       _dev(modal, self.modes * self.w.fields * self.nj, cd, "modal"), _stream(stream)))

    def bind_forward(self, physical, modal, stream=None) -> Bound:
        """A prepared forward launch (see Bound)."""
        cd = complex_dtype(self.precision)
        args = (self.h, _dev(physical, self.modes * self.w.planes, cd, "physical"),
                _dev(modal, self.modes * self.w.fields * self.nj, cd, "modal"), _stream(stream))
        return Bound(lib.skb_vertical_forward, args, (physical, modal, self))

    def inverse(self, modal, physical, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_vertical_inverse(self.h, _dev(modal, self.modes * self.w.fields * self.nj, cd, "modal"),
                                       _dev(physical, self.modes * self.w.planes, cd, "physical"), _stream(stream)))

    def reserve_split(self):
        """Allocate the interleaved staging of the split-layout calls (once)."""
        check(lib.skb_vertical_plan_reserve_split(self.h))

    def _split(self, fn, n_in, n_out, in_re, in_im, out_re, out_im, stream):
        rd = real_dtype(self.precision)
        check(fn(self.h, _dev_opt(in_re, n_in, rd, "input real"), _dev_opt(in_im, n_in, rd, "input imaginary"),
                 _dev_opt(out_re, n_out, rd, "output real"), _dev_opt(out_im, n_out, rd, "output imaginary"),
                 _stream(stream)))

    def forward_split(self, physical_real, physical_imag, modal_real, modal_imag, stream=None):
        """Split-layout forward product; a part whose input and output are None is left out."""
        self._split(lib.skb_vertical_forward_split, self.modes * self.w.planes, self.modes * self.w.fields * self.nj,
                    physical_real, physical_imag, modal_real, modal_imag, stream)

    def forward_split_noalloc(self, physical_real, physical_imag, modal_real, modal_imag, stream=None):
        """forward_split with its staging in the library's static 64 MiB device scratch: no
        reserve_split, no allocation; overlapping calls on one GPU race."""
        self._split(lib.skb_vertical_forward_split_noalloc, self.modes * self.w.planes,
                    self.modes * self.w.fields * self.nj, physical_real, physical_imag, modal_real, modal_imag, stream)

    def inverse_split(self, modal_real, modal_imag, physical_real, physical_imag, stream=None):
        self._split(lib.skb_vertical_inverse_split, self.modes * self.w.fields * self.nj, self.modes * self.w.planes,
                    modal_real, modal_imag, physical_real, physical_imag, stream)

    def matrix_bytes(self) -> int:
        return int(lib.skb_vertical_matrix_bytes(self.h))

    def set_grid(self, workgroups: int):
        """Workgroups per launch (0 = one per 16-column tile)."""
        check(lib.skb_vertical_plan_set_grid(self.h, workgroups))

    def items(self) -> int:
        return int(lib.skb_vertical_plan_items(self.h))

    def set_products(self, products: int):
        """SKB_PRODUCTS_SPLIT_BF16 (default) or SKB_PRODUCTS_FP32 for FP32 plans."""
        check(lib.skb_vertical_plan_set_products(self.h, products))

    def products(self, direction: int = 0) -> int:
        """Products the forward (0) or inverse (1) GEMM uses."""
        return int(lib.skb_vertical_plan_products(self.h, direction))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_vertical_plan_destroy(self.h)
            self.h = None

    __del__ = close


class OperatorPlan:
    """One retained-operator evaluation per call (skb_retained_operator):
    pruned inverse c2r of the compact retained spectrum into real planes,
    forward r2c back to the retained modes, K^2-grouped vertical forward
    product into the modal layout j + Nj*(field + F*mode)."""

    STAGES = (SKB_OP_STAGE_LAYOUT, SKB_OP_STAGE_INVERSE, SKB_OP_STAGE_FORWARD, SKB_OP_STAGE_VERTICAL)

    def __init__(self, w: Workload, modes: np.ndarray, nj: int, groups: np.ndarray, forward: np.ndarray,
                 inverse: np.ndarray, precision: int = SKB_F32):
        self.w, self.nj, self.precision, self.modes = w, nj, precision, modes
        self._wc = w.c()
        forward = np.ascontiguousarray(forward, np.float64)
        inverse = np.ascontiguousarray(inverse, np.float64)
        if forward.size != len(groups) * nj * w.nz or inverse.size != forward.size:
            raise ValueError("matrix family size does not match groups x nj x nz")
        h = ctypes.c_void_p()
        check(lib.skb_retained_operator_create(ctypes.byref(self._wc), _np(modes), len(modes), nj, _np(groups),
                                               len(groups), _np(forward), _np(inverse), precision, ctypes.byref(h)))
        self.h = h

    @property
    def retained_elements(self):
        return len(self.modes) * self.w.planes

    @property
    def modal_elements(self):
        return len(self.modes) * self.w.fields * self.nj

    def _args(self, retained, modal):
        cd = complex_dtype(self.precision)
        return (_dev(retained, self.retained_elements, cd, "retained"),
                _dev(modal, self.modal_elements, cd, "modal"))

    def execute(self, retained, modal, stream=None):
        r, m = self._args(retained, modal)
        check(lib.skb_retained_operator_execute(self.h, r, m, _stream(stream)))

    def bind_execute(self, retained, modal, stream=None) -> Bound:
        """A prepared evaluation (see Bound): one C-ABI call per evaluation."""
        r, m = self._args(retained, modal)
        return Bound(lib.skb_retained_operator_execute, (self.h, r, m, _stream(stream)), (retained, modal, self))

    def stage(self, stage: int, retained, modal, stream=None):
        r, m = self._args(retained, modal)
        check(lib.skb_retained_operator_stage(self.h, stage, r, m, _stream(stream)))

    def bind_stage(self, stage: int, retained, modal, stream=None) -> Bound:
        r, m = self._args(retained, modal)
        return Bound(lib.skb_retained_operator_stage, (self.h, stage, r, m, _stream(stream)), (retained, modal, self))

    def set_path(self, path: int):
        check(lib.skb_retained_operator_set_path(self.h, path))

    def path(self) -> int:
        return int(lib.skb_retained_operator_path(self.h))

    def set_fused(self, fused: bool):
        check(lib.skb_retained_operator_set_fused(self.h, 1 if fused else 0))

    def fused(self) -> bool:
        return bool(lib.skb_retained_operator_fused(self.h))

    def reserve_batch(self, evaluations: int):
        """Stage-batched execute_batch over groups of up to `evaluations` (0 or 1 releases)."""
        check(lib.skb_retained_operator_reserve_batch(self.h, int(evaluations)))

    def batch(self) -> int:
        return int(lib.skb_retained_operator_batch(self.h))

    def reserve_lanes(self, lanes: int):
        """Lanes of the batched execute (own streams and intermediates; 1 releases them)."""
        check(lib.skb_retained_operator_reserve_lanes(self.h, int(lanes)))

    def lanes(self):
        """(lanes reserved, every lane on a hardware queue of its own)"""
        apart = ctypes.c_int32(0)
        n = int(lib.skb_retained_operator_lanes(self.h, ctypes.byref(apart)))
        return n, bool(apart.value)

    def _batch_args(self, retained, modal):
        if len(retained) != len(modal):
            raise ValueError("one modal output per retained input")
        pairs = [self._args(r, m) for r, m in zip(retained, modal)]
        n = len(pairs)
        ra = (ctypes.c_void_p * max(1, n))(*[p[0] for p in pairs])
        ma = (ctypes.c_void_p * max(1, n))(*[p[1] for p in pairs])
        outs = {m.data_ptr() for m in modal}
        if len(outs) != n or outs & {r.data_ptr() for r in retained}:
            raise ValueError("batch outputs must not alias each other or an input")
        return n, ra, ma

    def execute_batch(self, retained, modal, stream=None):
        """Independent evaluations retained[i] -> modal[i]: stage-batched groups (reserve_batch) or lanes."""
        n, ra, ma = self._batch_args(retained, modal)
        check(lib.skb_retained_operator_execute_batch(self.h, n, ra, ma, _stream(stream)))

    def bind_execute_batch_stage(self, stage: int, retained, modal, stream=None) -> Bound:
        """One stage of the stage-batched execute (one launch per group), prepared."""
        n, ra, ma = self._batch_args(retained, modal)
        return Bound(lib.skb_retained_operator_execute_batch_stage, (self.h, int(stage), n, ra, ma, _stream(stream)),
                     (list(retained), list(modal), ra, ma, self))

    def bind_execute_batch(self, retained, modal, stream=None) -> Bound:
        n, ra, ma = self._batch_args(retained, modal)
        return Bound(lib.skb_retained_operator_execute_batch, (self.h, n, ra, ma, _stream(stream)),
                     (list(retained), list(modal), ra, ma, self))

    def real(self):
        """The operator's real planes (x + Nx(y + Ny*plane)) as a torch tensor view."""
        ptr = lib.skb_retained_operator_real(self.h)
        return _device_view(ptr, self.w.real_elements, real_dtype(self.precision), self)

    def device_bytes(self) -> int:
        return int(lib.skb_retained_operator_device_bytes(self.h))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_retained_operator_destroy(self.h)
            self.h = None

    __del__ = close


class WvmVerticalPlan:
    """Vertical GEMM applied in place on full WVM arrays
    (WvmDirectVerticalGemmProvider): retained modes only, k_x = 0 partners
    written as conjugates; other entries are left as initialised."""

    def __init__(self, w: Workload, modes: np.ndarray, nj: int, groups: np.ndarray, forward: np.ndarray,
                 inverse: np.ndarray, precision: int = SKB_F64):
        self.w, self.nj, self.precision = w, nj, precision
        self.modes = len(modes)
        self._wc = w.c()
        forward = np.ascontiguousarray(forward, np.float64)
        inverse = np.ascontiguousarray(inverse, np.float64)
        if forward.size != len(groups) * nj * w.nz or inverse.size != forward.size:
            raise ValueError("matrix family size does not match groups x nj x nz")
        h = ctypes.c_void_p()
        check(lib.skb_vertical_wvm_plan_create(ctypes.byref(self._wc), _np(modes), len(modes), nj, _np(groups),
                                               len(groups), _np(forward), _np(inverse), precision, ctypes.byref(h)))
        self.h = h

    @property
    def modal_elements(self) -> int:
        return int(lib.skb_vertical_wvm_modal_elements(self.h))

    def forward(self, full_spectrum, full_modal, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_vertical_wvm_forward(self.h, _dev(full_spectrum, self.w.spectrum_elements, cd, "spectrum"),
                                           _dev(full_modal, self.modal_elements, cd, "modal"), _stream(stream)))

    def inverse(self, full_modal, full_spectrum, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_vertical_wvm_inverse(self.h, _dev(full_modal, self.modal_elements, cd, "modal"),
                                           _dev(full_spectrum, self.w.spectrum_elements, cd, "spectrum"),
                                           _stream(stream)))

    def initialize_modal(self, full_modal, stream=None):
        check(lib.skb_vertical_wvm_initialize_modal(
            self.h, _dev(full_modal, self.modal_elements, complex_dtype(self.precision), "modal"), _stream(stream)))

    def initialize_spectrum(self, full_spectrum, stream=None):
        check(lib.skb_vertical_wvm_initialize_spectrum(
            self.h, _dev(full_spectrum, self.w.spectrum_elements, complex_dtype(self.precision), "spectrum"),
            _stream(stream)))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_vertical_plan_destroy(self.h)
            self.h = None

    __del__ = close


class FluxPlan:
    """Production-lifetime flux graph (src/benchmark.cpp:3957-4234): 3 + 3T
    modal input fields [U, V, W, q0x, q0y, q0z, ...] -> T modal outputs, with
    T = w.fields.  `fused` selects the single-pass physical-space stage."""

    def __init__(self, w: Workload, modes: np.ndarray, nj: int, groups: np.ndarray, forward: np.ndarray,
                 inverse: np.ndarray, precision: int = SKB_F64):
        self.w, self.nj, self.precision = w, nj, precision
        self.modes = len(modes)
        self.targets = w.fields
        self._wc = w.c()
        forward = np.ascontiguousarray(forward, np.float64)
        inverse = np.ascontiguousarray(inverse, np.float64)
        if forward.size != len(groups) * nj * w.nz or inverse.size != forward.size:
            raise ValueError("matrix family size does not match groups x nj x nz")
        h = ctypes.c_void_p()
        check(lib.skb_flux_plan_create(ctypes.byref(self._wc), _np(modes), len(modes), nj, _np(groups), len(groups),
                                       _np(forward), _np(inverse), precision, ctypes.byref(h)))
        self.h = h

    @property
    def input_elements(self) -> int:
        return self.modes * self.nj * (3 + 3 * self.targets)

    @property
    def output_elements(self) -> int:
        return self.modes * self.nj * self.targets

    def set_fused(self, fused: bool):
        check(lib.skb_flux_set_fused(self.h, 1 if fused else 0))

    def execute(self, input_modal, output_modal, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_flux_execute(self.h, _dev(input_modal, self.input_elements, cd, "input_modal"),
                                   _dev(output_modal, self.output_elements, cd, "output_modal"), _stream(stream)))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_flux_plan_destroy(self.h)
            self.h = None

    __del__ = close


def vertical_derivative_matrix(nz: int, lz: float = 1.0) -> np.ndarray:
    d = np.zeros(nz * nz)
    check(lib.skb_vertical_derivative_matrix(nz, float(lz), _np(d)))
    return d


class AdvectionFluxPlan:
    """Advection flux of one tracer: compact retained (u, v, w, q) ->
    compact retained -(u q_x + v q_y + w q_z) / (Nx*Ny)^2 (w.fields == 4)."""

    def __init__(self, w: Workload, modes: np.ndarray, dz: np.ndarray, precision: int = SKB_F64):
        self.w, self.precision, self.modes = w, precision, len(modes)
        self._wc = w.c()
        dz = np.ascontiguousarray(dz, np.float64)
        if dz.size != w.nz * w.nz:
            raise ValueError("dz must be Nz x Nz")
        h = ctypes.c_void_p()
        check(lib.skb_advection_flux_plan_create(ctypes.byref(self._wc), _np(modes), len(modes), _np(dz), precision,
                                                 ctypes.byref(h)))
        self.h = h

    def set_fused(self, fused: bool):
        check(lib.skb_advection_flux_set_fused(self.h, 1 if fused else 0))

    def set_workgroups(self, workgroups: int):
        """Caps every kernel of the plan at `workgroups` workgroups (0: one per
        work item); results are unchanged."""
        check(lib.skb_advection_flux_set_workgroups(self.h, int(workgroups)))

    def device_bytes(self) -> int:
        return int(lib.skb_advection_flux_device_bytes(self.h))

    def execute(self, spectral, flux, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_advection_flux_execute(self.h, _dev(spectral, self.modes * self.w.planes, cd, "spectral"),
                                             _dev(flux, self.modes * self.w.nz, cd, "flux"), _stream(stream)))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_advection_flux_plan_destroy(self.h)
            self.h = None

    __del__ = close


def modal_work(weights, inp, out, precision: int, stream=None):
    """applySyntheticModalWorkInterleaved on the device: out = weights * inp."""
    n = weights.numel()
    check(lib.skb_modal_work(precision, n, _dev(weights, n, real_dtype(precision), "weights"),
                             _dev(inp, n, complex_dtype(precision), "input"),
                             _dev(out, n, complex_dtype(precision), "output"), _stream(stream)))


def modal_work_split(weights, in_re, in_im, out_re, out_im, precision: int, stream=None):
    """applySyntheticModalWorkSplit on the device."""
    n = weights.numel()
    rd = real_dtype(precision)
    check(lib.skb_modal_work_split(precision, n, _dev(weights, n, rd, "weights"), _dev(in_re, n, rd, "input_real"),
                                   _dev(in_im, n, rd, "input_imag"), _dev(out_re, n, rd, "output_real"),
                                   _dev(out_im, n, rd, "output_imag"), _stream(stream)))


def wvm_modal_work(plan: "WvmVerticalPlan", weights, full_in, full_out, stream=None):
    """applySyntheticModalWorkWvm on a WVM-direct plan's full modal arrays."""
    n = plan.modes * plan.w.fields * plan.nj
    cd = complex_dtype(plan.precision)
    check(lib.skb_vertical_wvm_modal_work(plan.h, _dev(weights, n, real_dtype(plan.precision), "weights"),
                                          _dev(full_in, plan.modal_elements, cd, "full_modal_input"),
                                          _dev(full_out, plan.modal_elements, cd, "full_modal_output"),
                                          _stream(stream)))


class SpectralPipeline:
    """Synthetic antialiased spectral pipeline (src/benchmark.cpp:2978-3547):
    real planes -> r2c -> retained -> vertical forward -> modal weights ->
    vertical inverse -> c2r.  policy SKB_PIPELINE_WVM_DIRECT or
    SKB_PIPELINE_COMPACT; weights None = the synthetic weights."""

    def __init__(self, w: Workload, modes: np.ndarray, groups: np.ndarray, forward: np.ndarray, inverse: np.ndarray,
                 policy: int = SKB_PIPELINE_COMPACT, precision: int = SKB_F64, weights: np.ndarray = None):
        self.w, self.policy, self.precision = w, policy, precision
        self.modes, self.nj = len(modes), w.nj
        self._wc = w.c()
        forward = np.ascontiguousarray(forward, np.float64)
        inverse = np.ascontiguousarray(inverse, np.float64)
        if forward.size != len(groups) * self.nj * w.nz or inverse.size != forward.size:
            raise ValueError("matrix family size does not match groups x nj x nz")
        wp = None
        if weights is not None:
            weights = np.ascontiguousarray(weights, np.float64)
            if weights.size != self.modes * w.fields * self.nj:
                raise ValueError("weights must hold one value per modal element")
            wp = _np(weights)
        h = ctypes.c_void_p()
        check(lib.skb_spectral_pipeline_create(ctypes.byref(self._wc), _np(modes), len(modes), _np(groups),
                                               len(groups), _np(forward), _np(inverse), wp, policy, precision,
                                               ctypes.byref(h)))
        self.h = h

    @property
    def fused(self) -> bool:
        return bool(lib.skb_spectral_pipeline_fused(self.h))

    def set_fused(self, fused: bool):
        check(lib.skb_spectral_pipeline_set_fused(self.h, 1 if fused else 0))

    def device_bytes(self) -> int:
        return int(lib.skb_spectral_pipeline_device_bytes(self.h))

    def execute(self, real_in, real_out, stream=None):
        rd = real_dtype(self.precision)
        n = self.w.real_elements
        check(lib.skb_spectral_pipeline_execute(self.h, _dev(real_in, n, rd, "real_input"),
                                                _dev(real_out, n, rd, "real_output"), _stream(stream)))

    def stages(self, real_in, real_out, retained, forward_modal, post_modal, inverse_retained, stream=None):
        rd, cd = real_dtype(self.precision), complex_dtype(self.precision)
        n, nr, nm = self.w.real_elements, self.modes * self.w.planes, self.modes * self.w.fields * self.nj
        check(lib.skb_spectral_pipeline_stages(
            self.h, _dev(real_in, n, rd, "real_input"), _dev(real_out, n, rd, "real_output"),
            _dev(retained, nr, cd, "retained"), _dev(forward_modal, nm, cd, "forward_modal"),
            _dev(post_modal, nm, cd, "post_modal"), _dev(inverse_retained, nr, cd, "inverse_retained"),
            _stream(stream)))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_spectral_pipeline_destroy(self.h)
            self.h = None

    __del__ = close


SKB_TYPE1_DCT, SKB_TYPE1_DST, SKB_TYPE1_DCT_FORWARD, SKB_TYPE1_DST_FORWARD = 0, 1, 2, 3


def type1_matrix(nz: int, kind: int) -> np.ndarray:
    """The type-I transform of `kind` as a row-major nz x nz matrix."""
    m = np.zeros(nz * nz)
    check(lib.skb_type1_matrix(nz, kind, _np(m)))
    return m.reshape(nz, nz)


def type1_fill(rows: int, channels: int, nz: int, modes: np.ndarray = None, nx_half: int = 0) -> np.ndarray:
    """The constant-stratification benchmark's deterministic fill, complex
    (rows, channels, nz)."""
    out = np.zeros(rows * channels * nz * 2)
    check(lib.skb_type1_fill(rows, channels, nz, None if modes is None else _np(modes), nx_half, _np(out)))
    return out.view(np.complex128).reshape(rows, channels, nz)


class Type1Plan:
    """Constant-stratification type-I vertical transforms on [rows][channels][nz]
    complex device arrays, in place."""

    def __init__(self, rows: int, nz: int, precision: int = SKB_F64):
        self.rows, self.nz, self.precision = rows, nz, precision
        h = ctypes.c_void_p()
        check(lib.skb_type1_plan_create(rows, nz, precision, ctypes.byref(h)))
        self.h = h

    def execute(self, kind: int, data, channels: int, first: int = 0, count: int = None, stream=None):
        count = channels - first if count is None else count
        check(lib.skb_type1_execute(self.h, kind, channels, first, count,
                                    _dev(data, self.rows * channels * self.nz, complex_dtype(self.precision), "data"),
                                    _stream(stream)))

    def schedule(self, arena3, arena1, stream=None):
        cd = complex_dtype(self.precision)
        check(lib.skb_type1_vertical_schedule(self.h, _dev(arena3, self.rows * 3 * self.nz, cd, "arena3"),
                                              _dev(arena1, self.rows * self.nz, cd, "arena1"), _stream(stream)))

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_type1_plan_destroy(self.h)
            self.h = None

    __del__ = close


class Batch:
    """Retained-operator evaluations over size classes of uneven grids,
    replayed from one hipGraph.  Class arrays use the stacked layout of a
    workload with fields * instances fields."""

    def __init__(self, precision: int = SKB_F32):
        self.precision = precision
        h = ctypes.c_void_p()
        check(lib.skb_batch_create(precision, ctypes.byref(h)))
        self.h = h
        self.classes = []
        self._keep = []

    def add_class(self, w: Workload, instances: int, modes: np.ndarray, nj: int, groups: np.ndarray,
                  forward: np.ndarray, inverse: np.ndarray) -> int:
        wc = w.c()
        forward = np.ascontiguousarray(forward, np.float64)
        inverse = np.ascontiguousarray(inverse, np.float64)
        idx = ctypes.c_int()
        check(lib.skb_batch_add_class(self.h, ctypes.byref(wc), instances, _np(modes), len(modes), nj, _np(groups),
                                      len(groups), _np(forward), _np(inverse), ctypes.byref(idx)))
        self.classes.append({"w": w, "instances": instances, "modes": len(modes), "nj": nj})
        return idx.value

    def input_elements(self, cls: int) -> int:
        c = self.classes[cls]
        return c["modes"] * c["w"].planes * c["instances"]

    def output_elements(self, cls: int) -> int:
        c = self.classes[cls]
        return c["modes"] * c["w"].fields * c["instances"] * c["nj"]

    def bind(self, cls: int, input_tensor, output_tensor):
        cd = complex_dtype(self.precision)
        check(lib.skb_batch_bind(self.h, cls, _dev(input_tensor, self.input_elements(cls), cd, "input"),
                                 _dev(output_tensor, self.output_elements(cls), cd, "output")))
        self._keep.append((input_tensor, output_tensor))

    def run(self, stream=None):
        check(lib.skb_batch_run(self.h, _stream(stream)))

    def capture(self, stream):
        check(lib.skb_batch_capture(self.h, _stream(stream)))

    def launch(self, stream):
        check(lib.skb_batch_launch(self.h, _stream(stream)))

    def tuning(self) -> str:
        """The current choices (launch shape per stage, products per class) as a spec string."""
        buf = ctypes.create_string_buffer(1 << 14)
        check(lib.skb_batch_tuning(self.h, buf, len(buf)))
        return buf.value.decode()

    def set_tuning(self, spec: str):
        check(lib.skb_batch_set_tuning(self.h, spec.encode()))

    def signature(self) -> dict:
        """What a kept tuning is valid for: the device, the library and every class's shape."""
        device = ""
        try:
            import torch

            device = torch.cuda.get_device_name(0)
        except Exception:  # noqa: BLE001 - the name is informative only
            pass
        return {"device": device, "library": version(), "precision": self.precision,
                "classes": [[c["w"].nx, c["w"].ny, c["w"].nz, c["w"].fields, c["instances"], c["modes"], c["nj"]]
                            for c in self.classes]}

    def autotune(self, reps: int = 20, stream=None, cache=None):
        """Times the candidates of every stage of every class and keeps the
        fastest (skb_batch_autotune); returns the per-class report.  With
        `cache` (a JSON file path), a tuning kept there for this batch's
        signature is applied without timing (the report then carries
        "cached": true), and a new tuning is added to the file, which keeps
        one entry per signature (written to a temporary file, then renamed)."""
        import json
        import os

        key = json.dumps(self.signature(), sort_keys=True)
        manifest = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                manifest = json.load(f)
            entry = manifest.get(key)
            if entry:
                self.set_tuning(entry["spec"])
                return {"cached": True, "spec": entry["spec"], "classes": entry["report"]}
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib.skb_batch_autotune(self.h, reps, _stream(stream), buf, len(buf)))
        report = json.loads(buf.value.decode())
        spec = self.tuning()
        if cache:
            manifest[key] = {"spec": spec, "report": report, "reps": reps}
            tmp = "%s.tmp%d" % (cache, os.getpid())
            with open(tmp, "w") as f:
                json.dump(manifest, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return {"cached": False, "spec": spec, "classes": report}

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.skb_batch_destroy(self.h)
            self.h = None

    __del__ = close


def pointwise_advection(volume: int, scale: float, shared, derivative, target, precision=SKB_F64, stream=None):
    rd = real_dtype(precision)
    check(lib.skb_pointwise_advection(volume, scale, _dev(shared, 3 * volume, rd, "shared"),
                                      _dev(derivative, 3 * volume, rd, "derivative"), _dev(target, volume, rd, "target"),
                                      precision, _stream(stream)))
