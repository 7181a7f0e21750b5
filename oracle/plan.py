"""Plan-level composition of the oracle -- TEST INFRASTRUCTURE ONLY.

Restates the single-rank execute sequences of ``src/dtfft_plan.F90`` with the
host kernels and the FFT oracle, on numpy arrays in C order (so a Fortran
``[X, Y, Z]`` pencil is a numpy array of shape ``(nz, ny, nx)``).

* ``execute_2d``      -- ``src/dtfft_plan.F90:877-913``
* ``execute_z_slab``  -- ``src/dtfft_plan.F90:974-1016``
* ``execute_generic`` -- ``src/dtfft_plan.F90:1057-1101``

On one rank every transpose is a single permute kernel
(``src/dtfft_reshape_handle_generic.F90:232-253``): X->Y, Y->Z, Z->X use
PERMUTE_FORWARD and the reverse directions PERMUTE_BACKWARD.
"""
from __future__ import annotations

import numpy as np

from . import (FFT_BACKWARD, FFT_FORWARD, KERNEL_PERMUTE_BACKWARD, KERNEL_PERMUTE_FORWARD,
               autotune_permute, fft_c2c, fft_c2r, fft_r2c, permute_variant)


def _kernel_key(kernel_type, fdims) -> str:
    name = "PERMUTE_FORWARD" if kernel_type == KERNEL_PERMUTE_FORWARD else "PERMUTE_BACKWARD"
    return name + " " + "x".join(str(int(d)) for d in fdims)


def _permute(kernel_type, a: np.ndarray, fdims, kernels=None) -> np.ndarray:
    """Apply a permute kernel to a pencil whose Fortran dims are ``fdims``; ``kernels`` maps
    "<kind> <dims>" to the host-kernel variant (block, access) to run (default BASE/WRITE)."""
    fdims = [int(d) for d in fdims]
    flat = np.ascontiguousarray(a).ravel()
    out = np.empty_like(flat)
    view_in = flat.view(_raw(flat.itemsize))
    view_out = out.view(_raw(flat.itemsize))
    block, access = (kernels or {}).get(_kernel_key(kernel_type, fdims), (0, "write"))
    permute_variant(kernel_type, fdims, view_in, view_out, block, access)
    if len(fdims) == 2:
        shape = (fdims[0], fdims[1])  # out Fortran [b, a] -> numpy (a, b)
    elif kernel_type == KERNEL_PERMUTE_FORWARD:
        shape = (fdims[0], fdims[2], fdims[1])  # out [b, c, a]
    else:
        shape = (fdims[1], fdims[0], fdims[2])  # out [c, a, b]
    return out.reshape(shape)


def _raw(es):
    return {4: np.uint32, 8: np.uint64, 16: np.dtype((np.void, 16))}[es]


def _fft_last(a, sign):
    return fft_c2c(a, sign)


def _fft_2d(a, sign):
    """2D FFT over the two fastest axes (FFT_2D executor, dtfft_abstract_executor.F90:146-150)."""
    t = _fft_last(a, sign)
    t = np.swapaxes(t, -1, -2).copy()
    t = _fft_last(t, sign)
    return np.swapaxes(t, -1, -2).copy()


def execute_generic(x: np.ndarray, forward: bool = True, kernels=None) -> np.ndarray:
    """3D C2C generic cycle on one rank. Forward input is the X-pencil (nz, ny, nx);
    output the Z-pencil (ny, nx, nz).  Backward is the reverse.  ``kernels``: host-kernel
    variants of the permutes (from autotune_generic); default BASE/WRITE."""
    k = kernels
    if forward:
        nz, ny, nx = x.shape
        a = _fft_last(x, FFT_FORWARD)
        a = _permute(KERNEL_PERMUTE_FORWARD, a, (nx, ny, nz), k)    # X->Y: (nx, nz, ny)
        a = _fft_last(a, FFT_FORWARD)
        a = _permute(KERNEL_PERMUTE_FORWARD, a, (ny, nz, nx), k)    # Y->Z: (ny, nx, nz)
        return _fft_last(a, FFT_FORWARD)
    ny, nx, nz = x.shape
    a = _fft_last(x, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_BACKWARD, a, (nz, nx, ny), k)       # Z->Y: (nx, nz, ny)
    a = _fft_last(a, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_BACKWARD, a, (ny, nz, nx), k)       # Y->X: (nz, ny, nx)
    return _fft_last(a, FFT_BACKWARD)


def autotune_generic(dims, es: int, warmup: int = 2, iters: int = 5) -> dict:
    """Host-kernel variant of each permute of the 3D generic cycle of Fortran dims
    (nx, ny, nz), picked by timing as kernel_host%create_host does at DTFFT_EXHAUSTIVE
    (src/dtfft_kernel_host.F90:178-291).  Returns {"<kind> <dims>": (block, access)}."""
    nx, ny, nz = [int(d) for d in dims]
    out = {}
    for kt, fd in ((KERNEL_PERMUTE_FORWARD, (nx, ny, nz)), (KERNEL_PERMUTE_FORWARD, (ny, nz, nx)),
                   (KERNEL_PERMUTE_BACKWARD, (nz, nx, ny)), (KERNEL_PERMUTE_BACKWARD, (ny, nz, nx))):
        key = _kernel_key(kt, fd)
        if key not in out:
            out[key] = autotune_permute(kt, fd, es, warmup, iters)
    return out


def execute_z_slab(x: np.ndarray, forward: bool = True) -> np.ndarray:
    """3D C2C Z-slab cycle on one rank: 2D FFT(X,Y), X->Z, FFT(Z)."""
    if forward:
        nz, ny, nx = x.shape
        a = _fft_2d(x, FFT_FORWARD)
        a = _permute(KERNEL_PERMUTE_BACKWARD, a, (nx, ny, nz))      # X->Z: (ny, nx, nz)
        return _fft_last(a, FFT_FORWARD)
    ny, nx, nz = x.shape
    a = _fft_last(x, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_FORWARD, a, (nz, nx, ny))           # Z->X: (nz, ny, nx)
    return _fft_2d(a, FFT_BACKWARD)


def execute_2d(x: np.ndarray, forward: bool = True) -> np.ndarray:
    """2D C2C cycle: FFT(X), X->Y, FFT(Y).  Forward input (ny, nx) -> output (nx, ny)."""
    if forward:
        ny, nx = x.shape
        a = _fft_last(x, FFT_FORWARD)
        a = _permute(KERNEL_PERMUTE_FORWARD, a, (nx, ny))
        return _fft_last(a, FFT_FORWARD)
    nx, ny = x.shape
    a = _fft_last(x, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_FORWARD, a, (ny, nx))  # 2D backward remaps to forward
    return _fft_last(a, FFT_BACKWARD)


def execute_generic_r2c(x: np.ndarray) -> np.ndarray:
    """3D R2C generic forward: R2C along X (n -> n/2+1), then the C2C cycle on the
    complex pencil (src/dtfft_plan.F90:1057-1101 with fft(1) an R2C executor)."""
    nz, ny, nx = x.shape
    a = fft_r2c(x)
    ncx = nx // 2 + 1
    a = _permute(KERNEL_PERMUTE_FORWARD, a, (ncx, ny, nz))
    a = _fft_last(a, FFT_FORWARD)
    a = _permute(KERNEL_PERMUTE_FORWARD, a, (ny, nz, ncx))
    return _fft_last(a, FFT_FORWARD)


def execute_generic_c2r(z: np.ndarray, nx: int) -> np.ndarray:
    ny, ncx, nz = z.shape
    a = _fft_last(z, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_BACKWARD, a, (nz, ncx, ny))
    a = _fft_last(a, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_BACKWARD, a, (ny, nz, ncx))
    return fft_c2r(a, nx)


def execute_2d_r2c(x: np.ndarray) -> np.ndarray:
    """2D R2C forward (execute_2d, src/dtfft_plan.F90:877-913, fft(1) an R2C executor):
    R2C along X (n -> n/2+1), X->Y permute, C2C along Y.  Input (ny, nx) real; output the
    Y-pencil (ncx, ny)."""
    ny, nx = x.shape
    a = fft_r2c(x)
    a = _permute(KERNEL_PERMUTE_FORWARD, a, (nx // 2 + 1, ny))
    return _fft_last(a, FFT_FORWARD)


def execute_2d_c2r(z: np.ndarray, nx: int) -> np.ndarray:
    """Inverse of execute_2d_r2c (unnormalised): C2C along Y, Y->X permute, C2R along X."""
    ncx, ny = z.shape
    a = _fft_last(z, FFT_BACKWARD)
    a = _permute(KERNEL_PERMUTE_FORWARD, a, (ny, ncx))  # 2D backward remaps to forward
    return fft_c2r(a, nx)

