"""CPU oracle for the retained-mode spectral operator -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module; the HIP product never does.  It wraps the C
restatement in ``skb_oracle.c`` (built to ``oracle/build/liboracle.so``) and
adds the FP64 full-size paths with scipy's pocketfft, which is an FFT
implementation independent of both the C direct DFT and the HIP kernels.

Reference citations are to the skbench source tree.

Pinning summary (details in tests/test_oracle.py):
  * retained_modes/mode_order_hash/k2_groups/group_hash/wvm_order_hash:
    published golden hashes and counts (tests/golden/published_workloads.json).
  * direct_r2c/direct_c2r: closed forms + agreement with pocketfft.
  * fft_r2c_wvm/fft_c2r_wvm: agreement with the direct DFT on small grids,
    round trip c2r(r2c(x)) == Nx*Ny*x.
  * retained_operator: closed form NxNy * vertical(x) on Hermitian-consistent
    inputs, and a direct-DFT composition on small grids.
  * k2_binned_groups / qr_gaussian_family: no reference counterpart (the
    32-bin QR-Gaussian family is the benchmark's own variant).  Pinned by QR
    uniqueness: Q^T Q = I and Q^T G upper triangular with positive diagonal.
  * spectral_gaussian: no reference counterpart; "parity unpinned" beyond its
    statistical and Hermitian properties (it only generates inputs).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

MODE_DTYPE = np.dtype(
    [
        ("k", "<i8"),
        ("l", "<i8"),
        ("stored_kx", "<u8"),
        ("stored_ky", "<u8"),
        ("conj", "<i4"),
        ("pad", "<i4"),
        ("radial", "<f8"),
    ]
)

FIXTURES = {"impulse": 0, "sinusoid": 1, "random": 2, "dc": 3, "nyquist": 4, "sinusoid-5-17": 5}


def build() -> str:
    """Compile the C oracle with gcc (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, i64, dbl, vp = ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        sig = {
            "or_retained_vertical_modes": (u64, [u64]),
            "or_retained_modes": (i64, [u64, u64, dbl, dbl, ctypes.c_int, vp, u64]),
            "or_mode_order_hash": (u64, [vp, u64]),
            "or_wvm_order_hash": (u64, [u64, u64, u64, u64]),
            "or_squared_key": (u64, [i64, i64]),
            "or_k2_groups": (i64, [vp, u64, vp, vp, vp]),
            "or_group_hash": (u64, [vp, vp, vp, u64]),
            "or_orthonormal_vertical": (None, [u64, u64, vp, vp]),
            "or_k2_vertical_family": (None, [u64, u64, u64, vp, vp, vp]),
            "or_signed_unit": (dbl, [vp]),
            "or_make_fixture": (ctypes.c_int, [u64, u64, u64, u64, ctypes.c_int, u64, vp]),
            "or_direct_r2c": (None, [u64, u64, u64, u64, vp, vp]),
            "or_direct_c2r": (None, [u64, u64, u64, u64, vp, vp]),
            "or_direct_retained": (None, [u64, u64, u64, u64, vp, u64, vp, vp]),
            "or_gather_retained": (None, [u64, u64, u64, u64, vp, u64, vp, vp]),
            "or_embed_retained": (None, [u64, u64, u64, u64, vp, u64, vp, vp]),
            "or_vertical_forward": (None, [u64, u64, u64, u64, vp, vp, vp, vp, vp]),
            "or_vertical_inverse": (None, [u64, u64, u64, u64, vp, vp, vp, vp, vp]),
            "or_pointwise_advection": (None, [u64, dbl, vp, vp, vp]),
            "or_max_relative_error_complex": (dbl, [vp, vp, u64]),
            "or_max_relative_error_real": (dbl, [vp, vp, u64, dbl]),
            "or_relative_l2_error_complex": (dbl, [vp, vp, u64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


@dataclass(frozen=True)
class Workload:
    """skbench Workload (include/skbench/skbench.hpp:23-39)."""

    nx: int
    ny: int
    nz: int
    fields: int
    lx: float = 1.0
    ly: float = 1.0
    antialias: bool = True

    @property
    def planes(self) -> int:
        return self.nz * self.fields

    @property
    def nx_half(self) -> int:
        return self.nx // 2 + 1

    @property
    def real_elements(self) -> int:
        return self.nx * self.ny * self.planes

    @property
    def spectrum_elements(self) -> int:
        return self.nx_half * self.ny * self.planes

    @property
    def nj(self) -> int:
        return int(lib().or_retained_vertical_modes(self.nz))


def fnv(h: int) -> str:
    return "fnv1a64:%016x" % h


# ---- modes, groups, hashes ------------------------------------------------
def retained_modes(w: Workload) -> np.ndarray:
    L = lib()
    n = L.or_retained_modes(w.nx, w.ny, w.lx, w.ly, int(w.antialias), None, 0)
    if n < 0:
        raise ValueError("invalid workload")
    out = np.zeros(n, dtype=MODE_DTYPE)
    L.or_retained_modes(w.nx, w.ny, w.lx, w.ly, int(w.antialias), _p(out), n)
    return out


def mode_order_hash(modes: np.ndarray) -> str:
    return fnv(lib().or_mode_order_hash(_p(modes), len(modes)))


def wvm_order_hash(w: Workload) -> str:
    return fnv(lib().or_wvm_order_hash(w.nx, w.ny, w.nz, w.fields))


def k2_groups(modes: np.ndarray):
    n = len(modes)
    key = np.zeros(n, np.uint64)
    first = np.zeros(n, np.uint64)
    count = np.zeros(n, np.uint64)
    g = lib().or_k2_groups(_p(modes), n, _p(key), _p(first), _p(count))
    if g < 0:
        raise ValueError("equal squared-wavenumber modes are not contiguous")
    return key[:g].copy(), first[:g].copy(), count[:g].copy()


def group_hash(key, first, count) -> str:
    key, first, count = (np.ascontiguousarray(a, np.uint64) for a in (key, first, count))
    return fnv(lib().or_group_hash(_p(key), _p(first), _p(count), len(key)))


def squared_keys(modes: np.ndarray) -> np.ndarray:
    k = modes["k"].astype(np.int64)
    l = modes["l"].astype(np.int64)
    return (k * k + l * l).astype(np.uint64)


def k2_binned_groups(modes: np.ndarray, bins: int = 32):
    """Benchmark grouping (no reference counterpart): bin b = floor(bins*K2/(K2max+1)) over the
    retained order, empty bins dropped.  Returns (bin_ids, first, count)."""
    k2 = squared_keys(modes).astype(np.int64)
    b = (bins * k2) // (int(k2.max()) + 1)
    ids, first, count = [], [], []
    for i, bi in enumerate(b):
        if not ids or ids[-1] != bi:
            ids.append(int(bi))
            first.append(i)
            count.append(1)
        else:
            count[-1] += 1
    if len(set(ids)) != len(ids):
        raise ValueError("bins are not contiguous in the retained order")
    return np.array(ids, np.uint64), np.array(first, np.uint64), np.array(count, np.uint64)


# ---- vertical fixtures ----------------------------------------------------
def orthonormal_vertical(nz: int, nj: int):
    f = np.zeros(nj * nz)
    i = np.zeros(nz * nj)
    lib().or_orthonormal_vertical(nz, nj, _p(f), _p(i))
    return f, i


def k2_vertical_family(nz: int, nj: int, keys: np.ndarray):
    keys = np.ascontiguousarray(keys, np.uint64)
    g = len(keys)
    f = np.zeros(g * nj * nz)
    i = np.zeros(g * nz * nj)
    lib().or_k2_vertical_family(nz, nj, g, _p(keys), _p(f), _p(i))
    return f, i


def signed_units(seed: int, count: int) -> np.ndarray:
    state = ctypes.c_uint64(0x9E3779B97F4A7C15 if seed == 0 else seed)
    fn = lib().or_signed_unit
    ptr = ctypes.byref(state)
    return np.array([fn(ptr) for _ in range(count)])


def gaussians(seed: int, count: int) -> np.ndarray:
    """Box-Muller pairs over the portable generator: pair i uses draws 2i and
    2i+1; an odd count drops the last sine."""
    pairs = (count + 1) // 2
    u = signed_units(seed, 2 * pairs)
    u1 = (u[0::2] + 1.0) * 0.5
    u2 = (u[1::2] + 1.0) * 0.5
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    out = np.empty(2 * pairs)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:count]


def qr_gaussian_family(n: int, groups: int, seed: int) -> np.ndarray:
    """Benchmark vertical family (no reference counterpart): per group, Q of the QR factorisation of
    an n x n Gaussian matrix (row-major, drawn from gaussians(seed + group)),
    with R's diagonal made positive.  Returns forward matrices (groups*n*n,
    row-major Q^T so that modal = Q^T phys) -- the inverse is Q."""
    out = np.zeros((groups, n, n))
    for g in range(groups):
        G = gaussians(seed + g, n * n).reshape(n, n)
        Q, R = np.linalg.qr(G)
        s = np.sign(np.diag(R))
        s[s == 0] = 1.0
        Q = Q * s[None, :]
        out[g] = Q.T
    return out.reshape(-1)


# ---- fixtures ---------------------------------------------------------------
def make_fixture(w: Workload, kind: str, seed: int) -> np.ndarray:
    out = np.zeros(w.real_elements)
    if lib().or_make_fixture(w.nx, w.ny, w.nz, w.fields, FIXTURES[kind], seed, _p(out)) != 0:
        raise ValueError(kind)
    return out


def spectral_gaussian(w: Workload, modes: np.ndarray, seed: int, exponent: float = -1.0) -> np.ndarray:
    """Compact retained input (z + Nz*field + P*mode): complex Gaussian times
    (1 + k^2 + l^2)^exponent, DC imaginary part zero.  exponent=-1 gives the
    (1+|k|^2)^-1 spectrum, exponent=-5/6 the |k|^-5/3 one."""
    P = w.planes
    g = gaussians(seed, 2 * P * len(modes)).reshape(len(modes), P, 2)
    amp = (1.0 + squared_keys(modes).astype(np.float64)) ** exponent
    out = (g[..., 0] + 1j * g[..., 1]) * amp[:, None]
    dc = (modes["k"] == 0) & (modes["l"] == 0)
    out[dc] = out[dc].real
    return np.ascontiguousarray(out.reshape(-1))


# ---- direct DFT and permutations (C) ----------------------------------------
def direct_r2c(w: Workload, real: np.ndarray) -> np.ndarray:
    real = np.ascontiguousarray(real, np.float64)
    out = np.zeros(w.spectrum_elements, np.complex128)
    lib().or_direct_r2c(w.nx, w.ny, w.nz, w.fields, _p(real), _p(out))
    return out


def direct_c2r(w: Workload, spec: np.ndarray) -> np.ndarray:
    spec = np.ascontiguousarray(spec, np.complex128)
    out = np.zeros(w.real_elements)
    lib().or_direct_c2r(w.nx, w.ny, w.nz, w.fields, _p(spec), _p(out))
    return out


def direct_retained(w: Workload, modes: np.ndarray, real: np.ndarray) -> np.ndarray:
    real = np.ascontiguousarray(real, np.float64)
    out = np.zeros(len(modes) * w.planes, np.complex128)
    lib().or_direct_retained(w.nx, w.ny, w.nz, w.fields, _p(modes), len(modes), _p(real), _p(out))
    return out


def gather_retained(w: Workload, modes: np.ndarray, full: np.ndarray) -> np.ndarray:
    full = np.ascontiguousarray(full, np.complex128)
    out = np.zeros(len(modes) * w.planes, np.complex128)
    lib().or_gather_retained(w.nx, w.ny, w.nz, w.fields, _p(modes), len(modes), _p(full), _p(out))
    return out


def embed_retained(w: Workload, modes: np.ndarray, retained: np.ndarray) -> np.ndarray:
    retained = np.ascontiguousarray(retained, np.complex128)
    out = np.zeros(w.spectrum_elements, np.complex128)
    lib().or_embed_retained(w.nx, w.ny, w.nz, w.fields, _p(modes), len(modes), _p(retained), _p(out))
    return out


def vertical_forward(w: Workload, nj: int, first, count, forward: np.ndarray, physical: np.ndarray) -> np.ndarray:
    first = np.ascontiguousarray(first, np.uint64)
    count = np.ascontiguousarray(count, np.uint64)
    physical = np.ascontiguousarray(physical, np.complex128)
    forward = np.ascontiguousarray(forward, np.float64)
    nmodes = int(first[-1] + count[-1])
    out = np.zeros(nmodes * w.fields * nj, np.complex128)
    lib().or_vertical_forward(w.nz, nj, w.fields, len(first), _p(first), _p(count), _p(forward), _p(physical), _p(out))
    return out


def vertical_inverse(w: Workload, nj: int, first, count, inverse: np.ndarray, modal: np.ndarray) -> np.ndarray:
    first = np.ascontiguousarray(first, np.uint64)
    count = np.ascontiguousarray(count, np.uint64)
    modal = np.ascontiguousarray(modal, np.complex128)
    inverse = np.ascontiguousarray(inverse, np.float64)
    nmodes = int(first[-1] + count[-1])
    out = np.zeros(nmodes * w.planes, np.complex128)
    lib().or_vertical_inverse(w.nz, nj, w.fields, len(first), _p(first), _p(count), _p(inverse), _p(modal), _p(out))
    return out


def pointwise_advection(volume: int, scale: float, shared: np.ndarray, derivative: np.ndarray) -> np.ndarray:
    shared = np.ascontiguousarray(shared, np.float64)
    derivative = np.ascontiguousarray(derivative, np.float64)
    out = np.zeros(volume)
    lib().or_pointwise_advection(volume, scale, _p(shared), _p(derivative), _p(out))
    return out


def max_relative_error(actual, expected, actual_scale: float = 1.0) -> float:
    """maximumRelativeError (src/model.cpp:718-736)."""
    a = np.ascontiguousarray(actual)
    b = np.ascontiguousarray(expected)
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = np.ascontiguousarray(a, np.complex128) * actual_scale
        b = np.ascontiguousarray(b, np.complex128)
        return float(lib().or_max_relative_error_complex(_p(a), _p(b), a.size))
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return float(lib().or_max_relative_error_real(_p(a), _p(b), a.size, actual_scale))


def relative_l2_error(actual, expected) -> float:
    """relativeL2Error (src/model.cpp:738-751)."""
    a = np.ascontiguousarray(actual, np.complex128)
    b = np.ascontiguousarray(expected, np.complex128)
    return float(lib().or_relative_l2_error_complex(_p(a), _p(b), a.size))


# ---- FP64 FFT paths (pocketfft) ---------------------------------------------
def _workers(workers):
    """Threads for the pocketfft calls: the caller's count, else the CPUs this
    process may run on, at most 16 (one GPU's share of a shared host)."""
    return workers if workers else max(1, min(16, len(os.sched_getaffinity(0))))


def fft_r2c_wvm(w: Workload, real: np.ndarray, workers: int | None = None) -> np.ndarray:
    """Full batched 2-D r2c written in WVM order z + Nz*field + P*(kx + Nxh*ky)
    (geometry of src/fftw_provider.cpp:537-572)."""
    import scipy.fft

    a = np.asarray(real, np.float64).reshape(w.planes, w.ny, w.nx)
    s = scipy.fft.rfft2(a, axes=(1, 2), workers=_workers(workers))
    return np.ascontiguousarray(s.transpose(1, 2, 0)).reshape(-1)


def fft_c2r_wvm(w: Workload, spec: np.ndarray, workers: int | None = None) -> np.ndarray:
    """Unnormalised batched 2-D c2r from WVM order (src/fftw_provider.cpp:574-610):
    columns c2c along ky, then rows c2r along kx ignoring Im of the kx=0 and
    kx=Nx/2 entries, as FFTW's multi-dimensional c2r does."""
    import scipy.fft

    s = np.asarray(spec, np.complex128).reshape(w.ny, w.nx_half, w.planes).transpose(2, 0, 1)
    out = scipy.fft.irfft2(s, s=(w.ny, w.nx), axes=(1, 2), norm="forward", workers=_workers(workers))
    return np.ascontiguousarray(out).reshape(-1)


def retained_inverse(w: Workload, modes: np.ndarray, retained: np.ndarray, workers=None) -> np.ndarray:
    """T_h^- = F_xy^-1 E_h (src/pruned_fftw_provider.cpp:265-276)."""
    return fft_c2r_wvm(w, embed_retained(w, modes, retained), workers)


def retained_forward(w: Workload, modes: np.ndarray, real: np.ndarray, workers=None) -> np.ndarray:
    """T_h^+ = P_h F_xy (src/pruned_fftw_provider.cpp:230-241)."""
    return gather_retained(w, modes, fft_r2c_wvm(w, real, workers))


def retained_operator(w: Workload, modes, retained, nj, first, count, forward, workers=None) -> np.ndarray:
    """One retained-operator evaluation: pruned inverse c2r, forward r2c
    truncated to the retained modes, K^2-grouped vertical forward GEMM."""
    real = retained_inverse(w, modes, retained, workers)
    back = retained_forward(w, modes, real, workers)
    return vertical_forward(w, nj, first, count, forward, back)


def wvm_direct_vertical(w: Workload, modes: np.ndarray, nj: int, first, count, mats: np.ndarray, full: np.ndarray,
                        forward: bool = True) -> np.ndarray:
    """WvmDirectVerticalGemmProvider::executeForward / executeInverse
    (src/vertical_gemm.cpp:1133-1157, boundary copies 1107-1131) on FULL WVM
    arrays: for each retained mode at stored frequency q = kx + (Nx/2+1) ky the
    block full[q] (K x fields, K fastest) is multiplied by the mode's group
    matrix (M x K row-major: Nj x Nz forward, Nz x Nj inverse) -- on the STORED
    value, without the mode's conjugation flag -- and written at q; for
    kx = 0, ky != 0, 2 ky != Ny the conjugate is also written at (0, Ny - ky).
    Entries of other frequencies are zero (initializeModalOutput /
    initializeSpectrumOutput)."""
    nxh, F = w.nx // 2 + 1, w.fields
    K, M = (w.nz, nj) if forward else (nj, w.nz)
    first = np.asarray(first, np.int64)
    count = np.asarray(count, np.int64)
    src = np.asarray(full, np.complex128).reshape(nxh * w.ny, F, K)
    A = np.asarray(mats, np.float64).reshape(len(first), M, K)
    group_of = np.repeat(np.arange(len(first)), count)
    kx = modes["stored_kx"].astype(np.int64)
    ky = modes["stored_ky"].astype(np.int64)
    q = kx + nxh * ky
    out = np.zeros((nxh * w.ny, F, M), np.complex128)
    for m in range(len(modes)):
        out[q[m]] = src[q[m]] @ A[group_of[m]].T
    for m in np.nonzero((kx == 0) & (ky != 0) & (2 * ky != w.ny))[0]:
        out[nxh * ((w.ny - ky[m]) % w.ny)] = np.conj(out[q[m]])
    return out.reshape(-1)


def flux_modal_fixture(w: Workload, modes: np.ndarray, nj: int, fields: int, seed: int) -> np.ndarray:
    """Modal flux input: uniform[-0.125, 0.125) values, DC imaginary part zero
    (the distribution of src/benchmark.cpp:4043-4055, drawn from the portable
    generator of src/benchmark.cpp:187-196 instead of std::mt19937_64)."""
    n = len(modes) * fields * nj
    u = 0.125 * signed_units(seed, 2 * n)
    out = (u[0::2] + 1j * u[1::2]).reshape(len(modes), fields, nj)
    dc = (modes["k"] == 0) & (modes["l"] == 0)
    out[dc] = out[dc].real
    return out.reshape(-1)


def production_flux(w: Workload, modes: np.ndarray, nj: int, first, count, forward: np.ndarray,
                    inverse: np.ndarray, modal_in: np.ndarray, targets: int, use_direct: bool = False) -> np.ndarray:
    """The serial flux oracle of src/benchmark.cpp:4182-4234 restated: vertical
    inverse of the 3 + 3T input fields, pruned horizontal inverse of every
    field, per target -(u*qx + v*qy + w*qz)/(Nx*Ny)^2, raw forward with
    gather, vertical forward of the T outputs.  `use_direct` swaps pocketfft
    for the O(N^4) direct DFT (small grids)."""
    nin = 3 + 3 * targets
    win = Workload(w.nx, w.ny, w.nz, nin, w.lx, w.ly)
    wt = Workload(w.nx, w.ny, w.nz, targets, w.lx, w.ly)
    phys = vertical_inverse(win, nj, first, count, inverse, modal_in)
    if use_direct:
        real = direct_c2r(win, embed_retained(win, modes, phys))
    else:
        real = retained_inverse(win, modes, phys)
    vol = w.nx * w.ny * w.nz
    real = real.reshape(nin, vol)
    scale = 1.0 / float(w.nx * w.ny) ** 2
    out_real = np.empty((targets, vol))
    for t in range(targets):
        d = 3 + 3 * t
        out_real[t] = -scale * (real[0] * real[d] + real[1] * real[d + 1] + real[2] * real[d + 2])
    if use_direct:
        compact = gather_retained(wt, modes, direct_r2c(wt, out_real.reshape(-1)))
    else:
        compact = retained_forward(wt, modes, out_real.reshape(-1))
    return vertical_forward(wt, nj, first, count, forward, compact)


def vertical_derivative_matrix(nz: int, lz: float = 1.0) -> np.ndarray:
    """Second-order centred d/dz on nz uniform levels over [0, lz], one-sided
    second order at the ends (the matrix the provider builds; no reference
    counterpart)."""
    h = lz / (nz - 1)
    d = np.zeros((nz, nz))
    d[0, :3] = np.array([-1.5, 2.0, -0.5]) / h
    for k in range(1, nz - 1):
        d[k, k - 1], d[k, k + 1] = -0.5 / h, 0.5 / h
    d[nz - 1, nz - 3:] = np.array([0.5, -2.0, 1.5]) / h
    return d.reshape(-1)


def advection_flux(w: Workload, modes: np.ndarray, dz: np.ndarray, spectral: np.ndarray,
                   use_direct: bool = False, z_chunk: int | None = None) -> np.ndarray:
    """Advection flux of one tracer from compact retained (u, v, w, q):
    q_x = i 2pi k/lx q, q_y = i 2pi l/ly q (logical (k, l) of each mode),
    q_z = D q along z; unnormalised inverses of the six fields; product
    -(u q_x + v q_y + w q_z)/(Nx Ny)^2; raw forward truncated to the modes.

    After q_z every step acts on one level at a time, so `z_chunk` levels at
    a time give the same values with the full-grid arrays of only that many
    levels in memory (the 1024^2 x Nz257 grids)."""
    n = len(modes)
    sp = np.asarray(spectral, np.complex128).reshape(n, 4, w.nz)
    u, v, ww, q = sp[:, 0], sp[:, 1], sp[:, 2], sp[:, 3]
    k = modes["k"].astype(np.float64)[:, None]
    l = modes["l"].astype(np.float64)[:, None]
    D = np.asarray(dz, np.float64).reshape(w.nz, w.nz)
    six = np.stack([u, v, ww, 1j * (2 * np.pi / w.lx) * k * q, 1j * (2 * np.pi / w.ly) * l * q, q @ D.T], axis=1)
    if z_chunk and z_chunk < w.nz:
        out = np.empty((n, w.nz), np.complex128)
        for z0 in range(0, w.nz, z_chunk):
            z1 = min(w.nz, z0 + z_chunk)
            out[:, z0:z1] = _advection_levels(Workload(w.nx, w.ny, z1 - z0, 4, w.lx, w.ly), modes,
                                              np.ascontiguousarray(six[:, :, z0:z1]), use_direct).reshape(n, z1 - z0)
        return out.reshape(-1)
    return _advection_levels(w, modes, six, use_direct)


def _advection_levels(w: Workload, modes: np.ndarray, six: np.ndarray, use_direct: bool) -> np.ndarray:
    """The per-level part of advection_flux on the six spectral fields
    (mode, field, z) of w.nz levels."""
    w6 = Workload(w.nx, w.ny, w.nz, 6, w.lx, w.ly)
    w1 = Workload(w.nx, w.ny, w.nz, 1, w.lx, w.ly)
    if use_direct:
        real = direct_c2r(w6, embed_retained(w6, modes, six.reshape(-1)))
    else:
        real = retained_inverse(w6, modes, six.reshape(-1))
    vol = w.nx * w.ny * w.nz
    r = real.reshape(6, vol)
    prod = -(r[0] * r[3] + r[1] * r[4] + r[2] * r[5]) / float(w.nx * w.ny) ** 2
    if use_direct:
        return gather_retained(w1, modes, direct_r2c(w1, prod))
    return retained_forward(w1, modes, prod)


def vertical_forward_blas(w: Workload, nj: int, first, count, forward: np.ndarray, physical: np.ndarray) -> np.ndarray:
    """Same product as vertical_forward, through numpy's BLAS (multi-threaded):
    used only by the CPU baseline timing.  Equality with the C loop is
    checked in tests/test_oracle.py."""
    nmodes = int(first[-1] + count[-1])
    P = np.asarray(physical, np.complex128).reshape(nmodes, w.fields, w.nz)
    F = np.asarray(forward, np.float64).reshape(len(first), nj, w.nz)
    out = np.empty((nmodes, w.fields, nj), np.complex128)
    for g in range(len(first)):
        a, b = int(first[g]), int(first[g] + count[g])
        out[a:b] = P[a:b] @ F[g].T
    return out.reshape(-1)


def retained_operator_fast(w: Workload, modes, retained, nj, first, count, forward, workers=None) -> np.ndarray:
    """retained_operator with the BLAS vertical product (CPU baseline leg)."""
    real = retained_inverse(w, modes, retained, workers)
    back = retained_forward(w, modes, real, workers)
    return vertical_forward_blas(w, nj, first, count, forward, back)


def vertical_inverse_blas(w: Workload, nj: int, first, count, inverse: np.ndarray, modal: np.ndarray) -> np.ndarray:
    """Same product as vertical_inverse, through numpy's BLAS; equality with the
    C loop is checked in tests/test_oracle.py."""
    nmodes = int(first[-1] + count[-1])
    M = np.asarray(modal, np.complex128).reshape(nmodes, w.fields, nj)
    inv = np.asarray(inverse, np.float64).reshape(len(first), w.nz, nj)
    out = np.empty((nmodes, w.fields, w.nz), np.complex128)
    for g in range(len(first)):
        a, b = int(first[g]), int(first[g] + count[g])
        out[a:b] = M[a:b] @ inv[g].T
    return out.reshape(-1)


# ---- synthetic modal work and the spectral pipeline ---------------------------
def synthetic_modal_weights(w: Workload, modes: np.ndarray) -> np.ndarray:
    """syntheticModalWorkWeights (src/model.cpp:246-267): weight of (mode, j,
    field) = 0.75 + 0.5 * key / 96 with key = (17|k| + 29|l| + 7j + 11 field)
    mod 97, stored at modalSpectrumIndex j + Nj*(field + F*mode)
    (src/model.cpp:233-238).  Integer arithmetic in uint64."""
    nj, F = w.nj, w.fields
    k = np.abs(modes["k"].astype(np.int64)).astype(np.uint64)
    l = np.abs(modes["l"].astype(np.int64)).astype(np.uint64)
    j = np.arange(nj, dtype=np.uint64)
    f = np.arange(F, dtype=np.uint64)
    key = (np.uint64(17) * k[:, None, None] + np.uint64(29) * l[:, None, None] + np.uint64(11) * f[None, :, None]
           + np.uint64(7) * j[None, None, :]) % np.uint64(97)
    return (0.75 + 0.5 * key.astype(np.float64) / 96.0).reshape(-1)


def _stored_frequencies(w: Workload, modes: np.ndarray):
    kx = modes["stored_kx"].astype(np.int64)
    ky = modes["stored_ky"].astype(np.int64)
    return kx, ky, kx + w.nx_half * ky


def apply_modal_work_wvm(w: Workload, modes: np.ndarray, weights: np.ndarray, full_in: np.ndarray,
                         full_out: np.ndarray) -> np.ndarray:
    """applySyntheticModalWorkWvm (src/model.cpp:286-308): a copy of full_out
    whose retained entries are weight * full_in at the STORED frequency, and
    whose k_x = 0 partner (0, Ny - ky) entries (ky != 0, 2ky != Ny) hold the
    conjugate of that product."""
    nj, F = w.nj, w.fields
    kx, ky, q = _stored_frequencies(w, modes)
    src = np.asarray(full_in, np.complex128).reshape(w.nx_half * w.ny, F * nj)
    out = np.array(full_out, np.complex128).reshape(w.nx_half * w.ny, F * nj)
    out[q] = src[q] * np.asarray(weights, np.float64).reshape(len(modes), F * nj)
    part = np.nonzero((kx == 0) & (ky != 0) & (2 * ky != w.ny))[0]
    out[w.nx_half * ((w.ny - ky[part]) % w.ny)] = np.conj(out[q[part]])
    return out.reshape(-1)


def gather_retained_modal(w: Workload, modes: np.ndarray, full_modal: np.ndarray) -> np.ndarray:
    """gatherRetainedModal (src/model.cpp:341-355)."""
    nj, F = w.nj, w.fields
    _, _, q = _stored_frequencies(w, modes)
    out = np.asarray(full_modal, np.complex128).reshape(w.nx_half * w.ny, F * nj)[q].copy()
    cj = modes["conj"] != 0
    out[cj] = np.conj(out[cj])
    return out.reshape(-1)


def embed_retained_modal(w: Workload, modes: np.ndarray, modal: np.ndarray) -> np.ndarray:
    """embedRetainedModal (src/model.cpp:357-379): zero fill, stored value
    (conjugated for conjugating modes) at the stored frequency, and its
    conjugate at the k_x = 0 partner."""
    nj, F = w.nj, w.fields
    kx, ky, q = _stored_frequencies(w, modes)
    stored = np.array(modal, np.complex128).reshape(len(modes), F * nj)
    cj = modes["conj"] != 0
    stored[cj] = np.conj(stored[cj])
    out = np.zeros((w.nx_half * w.ny, F * nj), np.complex128)
    out[q] = stored
    part = np.nonzero((kx == 0) & (ky != 0) & (2 * ky != w.ny))[0]
    out[w.nx_half * ((w.ny - ky[part]) % w.ny)] = np.conj(stored[part])
    return out.reshape(-1)


def spectral_pipeline(w: Workload, modes: np.ndarray, first, count, forward: np.ndarray, inverse: np.ndarray,
                      weights: np.ndarray, real: np.ndarray, workers=None) -> dict:
    """The reference's canonical spectral-pipeline oracle
    (src/benchmark.cpp:3085-3114): unnormalised r2c in WVM order, gather of
    the retained modes, vertical forward, modal work (weights in the modal
    layout, applySyntheticModalWorkInterleaved, src/model.cpp:269-274),
    vertical inverse, embed into the zero-padded half spectrum, unnormalised
    c2r.  Returns the five arrays the reference's correctness metrics read
    (src/benchmark.cpp:3175-3198)."""
    nj = w.nj
    retained = retained_forward(w, modes, real, workers)
    fwd = vertical_forward_blas(w, nj, first, count, forward, retained)
    post = np.asarray(weights, np.float64) * fwd
    inv = vertical_inverse_blas(w, nj, first, count, inverse, post)
    out = retained_inverse(w, modes, inv, workers)
    return {"retained": retained, "forward": fwd, "post": post, "inverse": inv, "output": out}


# ---- constant-stratification type-I vertical transforms ----------------------
def direct_forward_dct(x: np.ndarray) -> np.ndarray:
    """directForwardDct (src/constant_stratification.cpp:237-256): the WVM
    normalised DCT-I of one complex column, as an explicit sum."""
    x = np.asarray(x, np.complex128)
    nz = x.size
    out = np.zeros(nz, np.complex128)
    for j in range(nz):
        v = x[0] + (1.0 if j % 2 == 0 else -1.0) * x[-1]
        for z in range(1, nz - 1):
            v += 2.0 * np.cos(np.pi * (j * z) / (nz - 1)) * x[z]
        v /= nz - 1
        if j == nz - 1:
            v *= 0.5
        out[j] = v
    return out


def direct_forward_dst(x: np.ndarray) -> np.ndarray:
    """directForwardDst (src/constant_stratification.cpp:258-269): the WVM
    normalised DST-I of the interior, zero endpoints."""
    x = np.asarray(x, np.complex128)
    nz = x.size
    out = np.zeros(nz, np.complex128)
    for j in range(1, nz - 1):
        v = 0.0
        for z in range(1, nz - 1):
            v += 2.0 * np.sin(np.pi * (j * z) / (nz - 1)) * x[z]
        out[j] = v / (nz - 1)
    return out


TYPE1_DCT, TYPE1_DST, TYPE1_DCT_FORWARD, TYPE1_DST_FORWARD = 0, 1, 2, 3


def type1(kind: int, data: np.ndarray, first: int = 0, count: int | None = None) -> np.ndarray:
    """One FFTW R2RPlan execution (src/constant_stratification.cpp:95-147) on
    complex data (rows, channels, nz), channels [first, first + count):
    REDFT00 over all nz points or RODFT00 over the interior (scipy's
    unnormalised DCT/DST type 1 are those transforms), real and imaginary
    parts alike; the forward kinds add normalizeForwardDct / Dst
    (196-235).  Returns a new array."""
    import scipy.fft

    out = np.array(data, np.complex128)
    count = out.shape[1] - first if count is None else count
    sl = out[:, first:first + count, :]
    nz = out.shape[2]
    if kind in (TYPE1_DCT, TYPE1_DCT_FORWARD):
        t = scipy.fft.dct(sl.real, type=1, axis=-1) + 1j * scipy.fft.dct(sl.imag, type=1, axis=-1)
        if kind == TYPE1_DCT_FORWARD:
            t /= nz - 1
            t[..., -1] *= 0.5
        sl[...] = t
    else:
        inner = sl[..., 1:-1]
        t = scipy.fft.dst(inner.real, type=1, axis=-1) + 1j * scipy.fft.dst(inner.imag, type=1, axis=-1)
        if kind == TYPE1_DST_FORWARD:
            sl[..., 1:-1] = t / (nz - 1)
            sl[..., 0] = 0.0
            sl[..., -1] = 0.0
        else:
            sl[..., 1:-1] = t
    return out


def type1_schedule(arena3: np.ndarray, arena1: np.ndarray):
    """executeCompleteVerticalGraph (src/constant_stratification.cpp:271-292)."""
    a3, a1 = np.array(arena3, np.complex128), np.array(arena1, np.complex128)
    for _ in range(3):
        a3 = type1(TYPE1_DCT, a3, 0, 2)
        a3 = type1(TYPE1_DST, a3, 2, 1)
    for _ in range(2):
        a3 = type1(TYPE1_DST, a3, 0, 2)
        a3 = type1(TYPE1_DCT, a3, 2, 1)
    for _ in range(2):
        a1 = type1(TYPE1_DCT_FORWARD, a1)
    for _ in range(2):
        a1 = type1(TYPE1_DST_FORWARD, a1)
    return a3, a1


def type1_fill(rows: int, channels: int, nz: int, modes: np.ndarray | None = None, nx_half: int = 0) -> np.ndarray:
    """fillFull / fillCompact (src/constant_stratification.cpp:237-269):
    (sin a, cos 1.7a), a = 0.017(1 + row mod 97) + 0.113(1 + channel) +
    0.071(1 + z); compact rows use the mode's stored frequency and conjugate
    conjugating modes."""
    if modes is None:
        row = np.arange(rows, dtype=np.int64)
        cj = np.zeros(rows, bool)
    else:
        row = modes["stored_kx"].astype(np.int64) + nx_half * modes["stored_ky"].astype(np.int64)
        cj = modes["conj"] != 0
    c = np.arange(channels)
    z = np.arange(nz)
    a = (0.017 * (1 + row % 97))[:, None, None] + (0.113 * (1 + c))[None, :, None] + (0.071 * (1 + z))[None, None, :]
    out = np.sin(a) + 1j * np.cos(1.7 * a)
    out[cj] = np.conj(out[cj])
    return out
