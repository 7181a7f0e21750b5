"""GPU parity of the fused Z->X (3D) / Y->X (2D) transpose + C2R kernels that load their input
rows straight into LDS (each input element read from HBM once, the untangling in place in
LDS), against the FP64 oracle's C2R sequence (src/dtfft_plan.F90:1057-1101 backward, FFTW c2r
semantics) applied to a random half spectrum -- Im X_0 and Im X_{n/2} must not contribute:
the persistent double buffer (fft_c2r_dma2_kernel, lds_dma = 2, the default) and one tile
per block (fft_c2r_dma_kernel, lds_dma = 3) with the FFT twiddles read from global memory
(twl = 1) or kept in LDS (twl = 2).  Bars: FP32 1e-6, FP64 1e-12 relative L2.
Shapes are those whose C2R stage takes this kernel (lines a multiple of 16 FP32 / 8 FP64,
64 threads or more per block); the compile-only CPU tests check that they do."""
import numpy as np
import pytest

import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision
TOL = {P.DOUBLE: 1e-12, P.SINGLE: 1e-6}

SHAPES_3D = {P.SINGLE: [(256, 12, 32), (1024, 6, 16), (256, 3, 48)],
             P.DOUBLE: [(512, 6, 16), (1024, 4, 8), (512, 5, 24)]}
SHAPES_2D = {P.SINGLE: [(256, 64), (1024, 32)], P.DOUBLE: [(512, 16), (1024, 24)]}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdt(prec):
    return np.complex128 if prec == P.DOUBLE else np.complex64


def rdt(prec):
    return np.float64 if prec == P.DOUBLE else np.float32


def half_spectrum(shape, prec, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(cdt(prec))


VARIANTS = [(2, 1), (3, 1), (3, 2)]  # (lds_dma, twl)


@pytest.mark.parametrize("lds_dma,twl", VARIANTS)
@pytest.mark.parametrize("prec,dims", [(p, s) for p in (P.SINGLE, P.DOUBLE) for s in SHAPES_3D[p]])
def test_fused_c2r_3d(prec, dims, lds_dma, twl):
    nx, ny, nz = dims
    nc = nx // 2 + 1
    print(f"case: fused C2R 3D {dims} {prec.name} lds_dma={lds_dma} twl={twl}", flush=True)
    D.set_config(enable_fused=True, lds_dma=lds_dma, twl=twl)
    try:
        plan = D.PlanR2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    z = half_spectrum((ny, nc, nz), prec, nx + ny + nz)  # Z-pencil [Z, X, Y]
    a = D.DeviceArray.from_numpy(z, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.BACKWARD)
    got = b.download((nz, ny, nx), rdt(prec)).astype(np.float64)
    want = OP.execute_generic_c2r(z.astype(np.complex128), nx)
    assert rel(got, want) < TOL[prec]
    # and the round trip through the forward
    x = np.random.default_rng(7).standard_normal((nz, ny, nx)).astype(rdt(prec))
    a.upload(x)
    plan.execute(a, b, D.Execute.FORWARD)
    plan.execute(b, a, D.Execute.BACKWARD)
    back = a.download((nz, ny, nx), rdt(prec)).astype(np.float64) / x.size
    assert rel(back, x.astype(np.float64)) < 2 * TOL[prec]


@pytest.mark.parametrize("lds_dma,twl", VARIANTS)
@pytest.mark.parametrize("prec,dims", [(p, s) for p in (P.SINGLE, P.DOUBLE) for s in SHAPES_2D[p]])
def test_fused_c2r_2d(prec, dims, lds_dma, twl):
    nx, ny = dims
    nc = nx // 2 + 1
    print(f"case: fused C2R 2D {dims} {prec.name} lds_dma={lds_dma} twl={twl}", flush=True)
    D.set_config(enable_fused=True, lds_dma=lds_dma, twl=twl)
    try:
        plan = D.PlanR2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    z = half_spectrum((nc, ny), prec, nx * ny)  # Y-pencil [Y, X]
    a = D.DeviceArray.from_numpy(z, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.BACKWARD)
    got = b.download((ny, nx), rdt(prec)).astype(np.float64)
    assert rel(got, OP.execute_2d_c2r(z.astype(np.complex128), nx)) < TOL[prec]
