"""R2R (DCT / DST types 1-4) oracle -- TEST INFRASTRUCTURE ONLY.

dtFFT hands R2R transforms to FFTW (``fftw_plan_many_r2r``,
src/interfaces/fft/fftw/dtfft_executor_fftw_m.F90:149-223); its kind ids are FFTW's
REDFT/RODFT codes (include/dtfft_config.h.in:66-73).  FFTW is a third-party library that
is not under the reference tree and not installed here, so this module restates FFTW's
published definitions of the eight unnormalised transforms directly, as dense cosine /
sine matrices evaluated in extended precision (O(n^2) per line, for the small and
moderate sizes the tests use).  Pinned by tests/test_oracle_r2r.py against scipy's
golden vectors (tests/golden/fft_golden.npz) and the inverse-pair identities.

Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

# dtFFT kind ids (include/dtfft_config.h.in:66-73) = FFTW kinds
DCT_1, DCT_3, DCT_2, DCT_4 = 3, 4, 5, 6
DST_1, DST_3, DST_2, DST_4 = 7, 8, 9, 10
KINDS = (DCT_1, DCT_2, DCT_3, DCT_4, DST_1, DST_2, DST_3, DST_4)

# inverse kinds, src/dtfft_utils.F90:447-471
INVERSE = {DCT_1: DCT_1, DCT_2: DCT_3, DCT_3: DCT_2, DCT_4: DCT_4,
           DST_1: DST_1, DST_2: DST_3, DST_3: DST_2, DST_4: DST_4}


def logical_size(kind: int, n: int) -> int:
    """FFTW's logical size N: kind(inverse(kind)(x)) = N x."""
    if kind == DCT_1:
        return 2 * (n - 1)
    if kind == DST_1:
        return 2 * (n + 1)
    return 2 * n


def matrix(kind: int, n: int) -> np.ndarray:
    """M with y = M @ x for FFTW's definition of `kind` on n points (float64 result of a
    long-double evaluation)."""
    j = np.arange(n, dtype=np.longdouble)[None, :]
    k = np.arange(n, dtype=np.longdouble)[:, None]
    pi = np.longdouble(np.pi)
    if kind == DCT_1:  # REDFT00: x0 + (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi j k / (n-1))
        if n < 2:
            raise ValueError("DCT-I needs n >= 2")
        m = 2 * np.cos(pi * j * k / (n - 1))
        m[:, 0] = 1
        m[:, n - 1] = np.where(np.arange(n) % 2 == 0, 1, -1)
    elif kind == DCT_2:  # REDFT10: 2 sum x_j cos(pi (j + 1/2) k / n)
        m = 2 * np.cos(pi * (j + 0.5) * k / n)
    elif kind == DCT_3:  # REDFT01: x0 + 2 sum_{j>=1} x_j cos(pi j (k + 1/2) / n)
        m = 2 * np.cos(pi * j * (k + 0.5) / n)
        m[:, 0] = 1
    elif kind == DCT_4:  # REDFT11: 2 sum x_j cos(pi (j + 1/2)(k + 1/2) / n)
        m = 2 * np.cos(pi * (j + 0.5) * (k + 0.5) / n)
    elif kind == DST_1:  # RODFT00: 2 sum x_j sin(pi (j + 1)(k + 1) / (n + 1))
        m = 2 * np.sin(pi * (j + 1) * (k + 1) / (n + 1))
    elif kind == DST_2:  # RODFT10: 2 sum x_j sin(pi (j + 1/2)(k + 1) / n)
        m = 2 * np.sin(pi * (j + 0.5) * (k + 1) / n)
    elif kind == DST_3:  # RODFT01: (-1)^k x_{n-1} + 2 sum_{j<n-1} x_j sin(pi (j + 1)(k + 1/2) / n)
        m = 2 * np.sin(pi * (j + 1) * (k + 0.5) / n)
        m[:, n - 1] = np.where(np.arange(n) % 2 == 0, 1, -1)
    elif kind == DST_4:  # RODFT11: 2 sum x_j sin(pi (j + 1/2)(k + 1/2) / n)
        m = 2 * np.sin(pi * (j + 0.5) * (k + 0.5) / n)
    else:
        raise ValueError(f"unknown R2R kind {kind}")
    return m.astype(np.float64)


def r2r(x: np.ndarray, kind: int) -> np.ndarray:
    """Unnormalised R2R of `kind` along the last axis (FP64)."""
    x = np.asarray(x, dtype=np.float64)
    return x @ matrix(kind, x.shape[-1]).T


def execute_generic_r2r(x: np.ndarray, kinds, forward: bool = True) -> np.ndarray:
    """3D R2R cycle on one rank (src/dtfft_plan.F90:1057-1101 with R2R executors):
    forward R2R_X, X->Y, R2R_Y, Y->Z, R2R_Z; kinds = (kind_x, kind_y, kind_z).  The
    backward runs the inverse kinds (src/dtfft_utils.F90:447-471) in reverse order.
    Forward input (nz, ny, nx) -> Z-pencil (ny, nx, nz); backward the reverse."""
    kx, ky, kz = kinds
    if forward:
        a = r2r(x, kx)                                   # X-pencil (nz, ny, nx)
        a = np.ascontiguousarray(np.transpose(a, (2, 0, 1)))   # Y-pencil (nx, nz, ny)
        a = r2r(a, ky)
        a = np.ascontiguousarray(np.transpose(a, (2, 0, 1)))   # Z-pencil (ny, nx, nz)
        return r2r(a, kz)
    a = r2r(x, INVERSE[kz])                              # Z-pencil (ny, nx, nz)
    a = np.ascontiguousarray(np.transpose(a, (1, 2, 0)))   # Y-pencil (nx, nz, ny)
    a = r2r(a, INVERSE[ky])
    a = np.ascontiguousarray(np.transpose(a, (1, 2, 0)))   # X-pencil (nz, ny, nx)
    return r2r(a, INVERSE[kx])


def execute_2d_r2r(x: np.ndarray, kinds, forward: bool = True) -> np.ndarray:
    """2D R2R (execute_2d, src/dtfft_plan.F90:877-913): R2R_X, X->Y, R2R_Y.  Forward input
    (ny, nx) -> (nx, ny); backward the reverse with inverse kinds."""
    kx, ky = kinds
    if forward:
        a = r2r(x, kx)
        return r2r(np.ascontiguousarray(a.T), ky)
    a = r2r(x, INVERSE[ky])
    return r2r(np.ascontiguousarray(a.T), INVERSE[kx])
