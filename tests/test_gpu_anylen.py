"""GPU parity of transforms of any length against the FP64 oracle: lengths with a prime
factor above 13 (Bluestein's chirp z-transform, device/anylen.hip) and real transforms of
odd length (the direct odd-length kernel).  cuFFT and FFTW take every length
(src/interfaces/fft/cufft/dtfft_executor_cufft_m.F90:77-93), and the reference's own tests
run such shapes: 2D C2C 333 x 444 and 11 x 39 (tests/c/test_c2c_2d_float_c.c:33-38), 2D R2C
35 x 44 (tests/c/test_r2c_2d_float_c.c:35).

Bars: FP64 within 1e-12 relative L2 of the oracle, FP32 within 1e-6 (Bluestein runs two
transforms of length >= 2n - 1, so the FP32 bar is the plain kernels' bar)."""
import numpy as np
import pytest

import oracle as O
import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision
TOL = {P.DOUBLE: 1e-12, P.SINGLE: 1e-6}
RT = {P.DOUBLE: 1e-12, P.SINGLE: 2e-6}  # round trips: a forward and a backward transform


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdt(prec):
    return np.complex128 if prec == P.DOUBLE else np.complex64


def rdt(prec):
    return np.float64 if prec == P.DOUBLE else np.float32


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [17, 37, 101, 1022, 4099, 4444, 6561])
def test_c2c_1d_any_length(n, prec):
    """6561 = 3^8 runs eight radix-3 passes directly; the others run Bluestein."""
    print(f"case: C2C n={n} {prec.name}", flush=True)
    hm = max(3, 2048 // n)
    rng = np.random.default_rng(n + 11)
    x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(cdt(prec))
    ex = D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, hm, [n])
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    x64 = x.astype(np.complex128)
    for sign in (D.FFT_FORWARD, D.FFT_BACKWARD):
        ex.execute(a, b, sign)
        got = b.download(x.shape, x.dtype).astype(np.complex128)
        err = rel(got, O.fft_c2c(x64, sign))
        assert err < TOL[prec], (n, sign, err)
    ex.execute(a, a, D.FFT_FORWARD)  # in place
    assert rel(a.download(x.shape, x.dtype).astype(np.complex128), O.fft_c2c(x64)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [1, 3, 7, 35, 74, 146, 1023, 2002])
def test_r2c_c2r_1d_any_length(n, prec):
    print(f"case: R2C/C2R n={n} {prec.name}", flush=True)
    hm = max(3, 1024 // n)
    nc = n // 2 + 1
    rng = np.random.default_rng(n + 17)
    x = rng.standard_normal((hm, n)).astype(rdt(prec))
    ex = D.FFTExecutor(1, D.FFTType.R2C, prec, n, nc, hm, [n], [n], [nc])
    a = D.DeviceArray.from_numpy(x)
    c = D.DeviceArray(hm * nc * np.dtype(cdt(prec)).itemsize)
    ex.execute(a, c, D.FFT_FORWARD)
    got = c.download((hm, nc), cdt(prec)).astype(np.complex128)
    assert rel(got, O.fft_r2c(x.astype(np.float64))) < TOL[prec]
    # C2R of a random half spectrum: Im X_0 (and Im X_{n/2} for even n) must not contribute
    h = (rng.standard_normal((hm, nc)) + 1j * rng.standard_normal((hm, nc))).astype(cdt(prec))
    c.upload(h)
    r = D.DeviceArray(x.nbytes)
    ex.execute(c, r, D.FFT_BACKWARD)
    got = r.download((hm, n), rdt(prec)).astype(np.float64)
    assert rel(got, O.fft_c2r(h.astype(np.complex128), n)) < TOL[prec]


def make_plan(cls, dims, prec, fused):
    D.set_config(enable_fused=fused)
    try:
        return cls(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dims", [(333, 444), (11, 39)])
def test_c2c_2d_reference_shapes(dims, fused, prec):
    nx, ny = dims
    print(f"case: 2D C2C {nx}x{ny} {prec.name} fused={fused}", flush=True)
    rng = np.random.default_rng(nx * ny)
    x = (rng.uniform(-1, 1, (ny, nx)) + 1j * rng.uniform(-1, 1, (ny, nx))).astype(cdt(prec))
    plan = make_plan(D.PlanC2C, dims, prec, fused)
    a = D.DeviceArray.from_numpy(x)
    b, c = D.DeviceArray(x.nbytes), D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((nx, ny), x.dtype).astype(np.complex128)
    assert rel(z, OP.execute_2d(x.astype(np.complex128), True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((ny, nx), x.dtype).astype(np.complex128) / x.size
    assert rel(back, x.astype(np.complex128)) < RT[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dims", [(35, 44), (74, 20)])
def test_r2c_2d_any_length(dims, fused, prec):
    nx, ny = dims
    print(f"case: 2D R2C {nx}x{ny} {prec.name} fused={fused}", flush=True)
    rng = np.random.default_rng(nx + ny)
    x = rng.standard_normal((ny, nx)).astype(rdt(prec))
    plan = make_plan(D.PlanR2C, dims, prec, fused)
    nc = nx // 2 + 1
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b, c = D.DeviceArray(plan.alloc_bytes), D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((nc, ny), cdt(prec)).astype(np.complex128)
    assert rel(z, OP.execute_2d_r2c(x.astype(np.float64))) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((ny, nx), rdt(prec)).astype(np.float64) / x.size
    assert rel(back, x.astype(np.float64)) < RT[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dims", [(37, 20, 17), (12, 101, 6)])
def test_c2c_3d_any_length(dims, fused, prec):
    """Bluestein axes in the strided (fused transpose) layouts as well as in rows."""
    nx, ny, nz = dims
    print(f"case: 3D C2C {dims} {prec.name} fused={fused}", flush=True)
    rng = np.random.default_rng(nx + ny + nz)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(cdt(prec))
    plan = make_plan(D.PlanC2C, dims, prec, fused)
    a = D.DeviceArray.from_numpy(x)
    b, c = D.DeviceArray(x.nbytes), D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), x.dtype).astype(np.complex128)
    assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), x.dtype).astype(np.complex128) / x.size
    assert rel(back, x.astype(np.complex128)) < RT[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dims", [(74, 6, 8), (35, 10, 6)])
def test_r2c_3d_any_length(dims, fused, prec):
    """Even nx with n/2 = 37 runs the fused Z->X C2R through the Bluestein kernel; odd nx
    runs the plain sequence with the odd-length kernel."""
    nx, ny, nz = dims
    print(f"case: 3D R2C {dims} {prec.name} fused={fused}", flush=True)
    rng = np.random.default_rng(nx * ny + nz)
    x = rng.standard_normal((nz, ny, nx)).astype(rdt(prec))
    plan = make_plan(D.PlanR2C, dims, prec, fused)
    nc = nx // 2 + 1
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b, c = D.DeviceArray(plan.alloc_bytes), D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nc, nz), cdt(prec)).astype(np.complex128)
    assert rel(z, OP.execute_generic_r2c(x.astype(np.float64))) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), rdt(prec)).astype(np.float64) / x.size
    assert rel(back, x.astype(np.float64)) < RT[prec]
