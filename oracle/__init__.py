"""CPU oracle for dtfft_amd -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline).  The product path (``dtfft_amd``) never imports it.

Contents
--------
* ``host_kernel`` -- the reference's host kernel plugin (kernel_host) restated:
  kernel-type dispatch of ``_dtfft_kernel_host_routines.inc:16-96`` plus the 2D
  BACKWARD->FORWARD remap and the dummy rule of
  ``src/dtfft_abstract_kernel.F90:219-288``.  The loops themselves live in
  ``dtfft_oracle.c`` (one C function per reference routine, file:line cited).
* ``fft_c2c / fft_r2c / fft_c2r`` -- FP64 (or FP32) batched 1D FFTs with the FFTW
  conventions dtFFT uses (``dtfft_executor_fftw_m.F90:149-223``).
* ``dft_naive`` -- O(n^2) long-double DFT used to pin the FFT.

Parity pinning lives in ``tests/test_oracle_transpose.py`` and
``tests/test_oracle_fft.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# kernel_type_t values, src/dtfft_abstract_kernel.F90:59-98
KERNEL_DUMMY = -1
KERNEL_PACK = 1
KERNEL_COPY_PIPELINED = 2
KERNEL_UNPACK = 3
KERNEL_COPY = 4
KERNEL_UNPACK_PIPELINED = 5
KERNEL_PACK_PIPELINED = 6
KERNEL_PERMUTE_FORWARD = 7
KERNEL_PERMUTE_BACKWARD = 8
KERNEL_PERMUTE_BACKWARD_START = 9
KERNEL_PERMUTE_BACKWARD_END = 10
KERNEL_PERMUTE_BACKWARD_END_PIPELINED = 11
KERNEL_PACK_FORWARD = 12
KERNEL_PACK_BACKWARD = 13
KERNEL_UNPACK_FORWARD = 15
KERNEL_UNPACK_FORWARD_PIPELINED = 16
KERNEL_UNPACK_BACKWARD = 17
KERNEL_UNPACK_BACKWARD_PIPELINED = 18

FFT_FORWARD = -1   # src/include/_dtfft_private.h:27
FFT_BACKWARD = +1  # src/include/_dtfft_private.h:28


def build() -> str:
    """Compile dtfft_oracle.c into liboracle.so (gcc, OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH) or (
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "dtfft_oracle.c"))
        ):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i, ip, lng, dp, fp = (ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p)
    sig = {
        "orc_permute_forward_write": [vp, vp, i, i, ip],
        "orc_permute_forward_read": [vp, vp, i, i, ip],
        "orc_permute_backward_write": [vp, vp, i, ip],
        "orc_permute_backward_read": [vp, vp, i, ip],
        "orc_permute_forward_block": [vp, vp, i, i, ip, i, i],
        "orc_permute_backward_block": [vp, vp, i, ip, i, i],
        "orc_permute_backward_start_write": [vp, vp, i, ip],
        "orc_permute_backward_start_read": [vp, vp, i, ip],
        "orc_permute_backward_end_pipelined_write": [vp, vp, i, ip, ip],
        "orc_permute_backward_end_pipelined_read": [vp, vp, i, ip, ip],
        "orc_unpack_pipelined": [vp, vp, i, i, ip, ip],
        "orc_pack_pipelined": [vp, vp, i, i, ip, ip],
        "orc_copy": [vp, vp, i, lng, lng, lng],
        "orc_unpack_forward_pipelined_write": [vp, vp, i, i, ip, ip],
        "orc_unpack_forward_pipelined_read": [vp, vp, i, i, ip, ip],
        "orc_unpack_backward_pipelined_write": [vp, vp, i, ip, ip],
        "orc_unpack_backward_pipelined_read": [vp, vp, i, ip, ip],
        "orc_pack_forward_write": [vp, vp, i, i, ip, ip],
        "orc_pack_forward_read": [vp, vp, i, i, ip, ip],
        "orc_pack_backward_write": [vp, vp, i, ip, ip],
        "orc_pack_backward_read": [vp, vp, i, ip, ip],
        "orc_fft_c2c_f64": [dp, dp, i, lng, i],
        "orc_fft_c2c_f32": [fp, fp, i, lng, i],
        "orc_fft_r2c_f64": [dp, dp, i, lng],
        "orc_fft_r2c_f32": [fp, fp, i, lng],
        "orc_fft_c2r_f64": [dp, dp, i, lng],
        "orc_fft_c2r_f32": [fp, fp, i, lng],
        "orc_dft_naive": [dp, dp, i, i],
        "orc_max_threads": [],
        "orc_set_threads": [i],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int if name == "orc_max_threads" else None


def _iarr(v):
    a = (ctypes.c_int * len(v))(*[int(x) for x in v])
    return a


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Kernel plugin restatement
# --------------------------------------------------------------------------
_PIPELINED_ONLY = (KERNEL_UNPACK_PIPELINED, KERNEL_PERMUTE_BACKWARD_END_PIPELINED,
                   KERNEL_COPY_PIPELINED, KERNEL_PACK_BACKWARD, KERNEL_PACK_FORWARD,
                   KERNEL_PACK_PIPELINED, KERNEL_UNPACK_FORWARD_PIPELINED,
                   KERNEL_UNPACK_BACKWARD_PIPELINED)


def effective_kernel_type(kernel_type: int, ndims: int) -> int:
    """2D remap of BACKWARD-type kernels, src/dtfft_abstract_kernel.F90:272-283."""
    if ndims == 2:
        return {
            KERNEL_PACK_BACKWARD: KERNEL_PACK_FORWARD,
            KERNEL_PERMUTE_BACKWARD: KERNEL_PERMUTE_FORWARD,
            KERNEL_UNPACK_BACKWARD: KERNEL_UNPACK_FORWARD,
            KERNEL_UNPACK_BACKWARD_PIPELINED: KERNEL_UNPACK_FORWARD_PIPELINED,
        }.get(kernel_type, kernel_type)
    return kernel_type


def host_kernel(kernel_type: int, dims, inbuf: np.ndarray, outbuf: np.ndarray,
                neighbor_data=None, neighbor: int | None = None, access: str = "write") -> None:
    """Run the host kernel of ``kernel_type`` on raw element buffers.

    ``dims`` are the kernel's local dims (fastest first); ``neighbor_data`` is an
    (P, 5) int array whose rows are (nxx, nyy, nzz, din, dout);
    ``neighbor`` is 1-based as in the reference.  Element size is
    ``inbuf.itemsize`` (4, 8 or 16 bytes).  Follows
    ``_dtfft_kernel_host_routines.inc:16-96``.
    """
    L = lib()
    dims = [int(d) for d in dims]
    ndims = len(dims)
    if any(d == 0 for d in dims):
        return  # dummy kernel, src/dtfft_abstract_kernel.F90:233-237
    if kernel_type == KERNEL_DUMMY:
        return
    es = inbuf.itemsize
    assert outbuf.itemsize == es and es in (4, 8, 16)
    kt = effective_kernel_type(kernel_type, ndims)
    ip, op = _ptr(inbuf), _ptr(outbuf)
    d = _iarr(dims + [1] * (3 - ndims))
    nd = None if neighbor_data is None else np.asarray(neighbor_data, dtype=np.int64).reshape(-1, 5)

    def loc(n):
        return _iarr(nd[n - 1])

    if kt in _PIPELINED_ONLY and neighbor is None:
        raise ValueError("neighbor is required for pipelined kernel types")

    if kt == KERNEL_UNPACK:
        for n in range(1, nd.shape[0] + 1):
            L.orc_unpack_pipelined(ip, op, es, ndims, d, loc(n))
    elif kt == KERNEL_UNPACK_PIPELINED:
        L.orc_unpack_pipelined(ip, op, es, ndims, d, loc(neighbor))
    elif kt == KERNEL_PACK:
        for n in range(1, nd.shape[0] + 1):
            L.orc_pack_pipelined(ip, op, es, ndims, d, loc(n))
    elif kt == KERNEL_PACK_PIPELINED:
        L.orc_pack_pipelined(ip, op, es, ndims, d, loc(neighbor))
    elif kt == KERNEL_COPY:
        L.orc_copy(ip, op, es, int(np.prod(dims)), 0, 0)
    elif kt == KERNEL_COPY_PIPELINED:
        row = nd[neighbor - 1]
        L.orc_copy(ip, op, es, int(row[0] * row[1] * row[2]), int(row[3]), int(row[4]))
    else:
        w = access == "write"
        if kt == KERNEL_PERMUTE_FORWARD:
            (L.orc_permute_forward_write if w else L.orc_permute_forward_read)(ip, op, es, ndims, d)
        elif kt == KERNEL_PERMUTE_BACKWARD:
            (L.orc_permute_backward_write if w else L.orc_permute_backward_read)(ip, op, es, d)
        elif kt == KERNEL_PERMUTE_BACKWARD_START:
            (L.orc_permute_backward_start_write if w else L.orc_permute_backward_start_read)(ip, op, es, d)
        elif kt == KERNEL_PERMUTE_BACKWARD_END:
            f = L.orc_permute_backward_end_pipelined_write if w else L.orc_permute_backward_end_pipelined_read
            for n in range(1, nd.shape[0] + 1):
                f(ip, op, es, d, loc(n))
        elif kt == KERNEL_PERMUTE_BACKWARD_END_PIPELINED:
            f = L.orc_permute_backward_end_pipelined_write if w else L.orc_permute_backward_end_pipelined_read
            f(ip, op, es, d, loc(neighbor))
        elif kt == KERNEL_UNPACK_FORWARD:
            f = L.orc_unpack_forward_pipelined_write if w else L.orc_unpack_forward_pipelined_read
            for n in range(1, nd.shape[0] + 1):
                f(ip, op, es, ndims, d, loc(n))
        elif kt == KERNEL_UNPACK_FORWARD_PIPELINED:
            f = L.orc_unpack_forward_pipelined_write if w else L.orc_unpack_forward_pipelined_read
            f(ip, op, es, ndims, d, loc(neighbor))
        elif kt == KERNEL_UNPACK_BACKWARD:
            f = L.orc_unpack_backward_pipelined_write if w else L.orc_unpack_backward_pipelined_read
            for n in range(1, nd.shape[0] + 1):
                f(ip, op, es, d, loc(n))
        elif kt == KERNEL_UNPACK_BACKWARD_PIPELINED:
            f = L.orc_unpack_backward_pipelined_write if w else L.orc_unpack_backward_pipelined_read
            f(ip, op, es, d, loc(neighbor))
        elif kt == KERNEL_PACK_FORWARD:
            f = L.orc_pack_forward_write if w else L.orc_pack_forward_read
            f(ip, op, es, ndims, d, loc(neighbor))
        elif kt == KERNEL_PACK_BACKWARD:
            f = L.orc_pack_backward_write if w else L.orc_pack_backward_read
            f(ip, op, es, d, loc(neighbor))
        else:
            raise ValueError(f"unknown kernel type {kernel_type}")


# ---------------------------------------------------------------------------
# Host kernel variants and their timed selection (kernel_host%create_host)
# ---------------------------------------------------------------------------
BLOCK_SIZES = (4, 8, 16, 32, 64)   # HOST_KERNEL_BLOCK_4 .. HOST_KERNEL_BLOCK_64


def permute_variant(kernel_type: int, dims, inbuf: np.ndarray, outbuf: np.ndarray, block: int = 0,
                    access: str = "write") -> None:
    """PERMUTE_FORWARD / PERMUTE_BACKWARD with a host-kernel variant: block 0 is BASE
    (_dtfft_kernel_host_routines.inc), block 4..64 the cache-blocked BLOCK_<block> routines
    (_dtfft_kernel_host_block_routines.inc); access "write" or "read" loop order."""
    dims = [int(d) for d in dims]
    kt = effective_kernel_type(kernel_type, len(dims))
    if block == 0:
        host_kernel(kt, dims, inbuf, outbuf, access=access)
        return
    es = inbuf.itemsize
    d = _iarr(dims + [1] * (3 - len(dims)))
    read = 1 if access == "read" else 0
    if kt == KERNEL_PERMUTE_FORWARD:
        lib().orc_permute_forward_block(_ptr(inbuf), _ptr(outbuf), es, len(dims), d, int(block), read)
    elif kt == KERNEL_PERMUTE_BACKWARD:
        lib().orc_permute_backward_block(_ptr(inbuf), _ptr(outbuf), es, d, int(block), read)
    else:
        raise ValueError(f"no blocked variant restated for kernel type {kernel_type}")


def autotune_permute(kernel_type: int, dims, es: int, warmup: int = 2, iters: int = 5):
    """Timed choice of a permute's host kernel, as kernel_host%create_host does at
    DTFFT_EXHAUSTIVE (src/dtfft_kernel_host.F90:178-291): BASE is timed with WRITE and with
    READ access (select_access_mode, _dtfft_kernel_host_routines.inc:98-135) and keeps the
    faster mode; BLOCK_4 .. BLOCK_64 are then timed in that mode; the fastest wins.  Warm-up
    and timed iterations default to the reference's measure_warmup_iters / measure_iters.
    Returns (block, access) with block 0 for BASE."""
    import time
    n = int(np.prod([int(d) for d in dims]))
    dt = elem_dtype(es)
    a = np.zeros(n, dtype=dt)
    b = np.zeros(n, dtype=dt)

    def bench(block, access):
        for _ in range(warmup):
            permute_variant(kernel_type, dims, a, b, block, access)
        t0 = time.perf_counter()
        for _ in range(iters):
            permute_variant(kernel_type, dims, a, b, block, access)
        return (time.perf_counter() - t0) / iters

    tw, tr = bench(0, "write"), bench(0, "read")
    access = "write" if tw < tr else "read"
    best, best_t = 0, min(tw, tr)
    for blk in BLOCK_SIZES:
        t = bench(blk, access)
        if t < best_t:
            best, best_t = blk, t
    return best, access


def elem_dtype(es: int):
    """A numpy dtype whose items are `es` raw bytes (for bit-exact moves)."""
    return {4: np.uint32, 8: np.uint64, 16: np.dtype((np.void, 16))}[es]


# --------------------------------------------------------------------------
# FFT oracle
# --------------------------------------------------------------------------
def _as_interleaved(x: np.ndarray, real):
    x = np.ascontiguousarray(x)
    return x.view(real)


def fft_c2c(x: np.ndarray, sign: int = FFT_FORWARD) -> np.ndarray:
    """Batched 1D C2C along the last numpy axis (the fastest one)."""
    dt = np.complex128 if x.dtype == np.complex128 else np.complex64
    x = np.ascontiguousarray(x, dtype=dt)
    n = x.shape[-1]
    out = np.empty_like(x)
    if n == 0 or x.size == 0:
        return out
    hm = x.size // n
    if dt == np.complex128:
        lib().orc_fft_c2c_f64(_ptr(x), _ptr(out), n, hm, sign)
    else:
        lib().orc_fft_c2c_f32(_ptr(x), _ptr(out), n, hm, sign)
    return out


def fft_r2c(x: np.ndarray) -> np.ndarray:
    dt = np.float64 if x.dtype == np.float64 else np.float32
    x = np.ascontiguousarray(x, dtype=dt)
    n = x.shape[-1]
    out = np.empty(x.shape[:-1] + (n // 2 + 1,), dtype=np.complex128 if dt == np.float64 else np.complex64)
    hm = x.size // n
    (lib().orc_fft_r2c_f64 if dt == np.float64 else lib().orc_fft_r2c_f32)(_ptr(x), _ptr(out), n, hm)
    return out


def fft_c2r(x: np.ndarray, n: int) -> np.ndarray:
    dt = np.complex128 if x.dtype == np.complex128 else np.complex64
    x = np.ascontiguousarray(x, dtype=dt)
    assert x.shape[-1] == n // 2 + 1
    out = np.empty(x.shape[:-1] + (n,), dtype=np.float64 if dt == np.complex128 else np.float32)
    hm = x.size // x.shape[-1]
    (lib().orc_fft_c2r_f64 if dt == np.complex128 else lib().orc_fft_c2r_f32)(_ptr(x), _ptr(out), n, hm)
    return out


def dft_naive(x: np.ndarray, sign: int = FFT_FORWARD) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128).ravel()
    out = np.empty_like(x)
    lib().orc_dft_naive(_ptr(x), _ptr(out), x.size, sign)
    return out


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())
