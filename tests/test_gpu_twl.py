"""GPU parity of the FFT kernels that keep their twiddle tables in LDS (persistent blocks),
forced on at small shapes (config twl = 2), against the FP64 oracle.

By default these kernels run only on stages of >= 2048 tiles (FP32 rows of n >= 2048, the
fused unpack+FFT, column-in passes without LDS-DMA, the fused Z->X C2R); forcing them covers
every variant -- rows, segmented rows, column-in / row-out, column-in / column-out in place,
C2R rows and the fused column-in C2R -- in both precisions and both directions.  The same
default kernels at their production shapes are checked in tests/test_gpu_configs.py.

Bars: FP64 within 1e-12 relative L2 of the oracle, FP32 within 1e-6."""
import numpy as np
import pytest

import oracle as O
import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision
TOL = {P.DOUBLE: 1e-12, P.SINGLE: 1e-6}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdt(prec):
    return np.complex128 if prec == P.DOUBLE else np.complex64


def rdt(prec):
    return np.float64 if prec == P.DOUBLE else np.float32


@pytest.fixture(autouse=True)
def forced_twl():
    D.set_config(twl=2)
    yield
    D.set_config()


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [2048, 512, 60, 7])
def test_c2c_rows(n, prec):
    hm = 37
    rng = np.random.default_rng(n + 3)
    x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(cdt(prec))
    ex = D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, hm, [n])
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    for sign in (D.FFT_FORWARD, D.FFT_BACKWARD):
        ex.execute(a, b, sign)
        got = b.download(x.shape, x.dtype).astype(np.complex128)
        assert rel(got, O.fft_c2c(x.astype(np.complex128), sign)) < TOL[prec], (n, sign)


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [2048, 256, 6])
def test_c2r_rows(n, prec):
    hm = 19
    nc = n // 2 + 1
    rng = np.random.default_rng(n + 5)
    ex = D.FFTExecutor(1, D.FFTType.R2C, prec, n, nc, hm, [n], [n], [nc])
    h = (rng.standard_normal((hm, nc)) + 1j * rng.standard_normal((hm, nc))).astype(cdt(prec))
    c = D.DeviceArray.from_numpy(h)
    r = D.DeviceArray(hm * n * np.dtype(rdt(prec)).itemsize)
    ex.execute(c, r, D.FFT_BACKWARD)
    got = r.download((hm, n), rdt(prec)).astype(np.float64)
    assert rel(got, O.fft_c2r(h.astype(np.complex128), n)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims", [(60, 84, 35), (48, 40, 24), (512, 16, 8)])
def test_c2c_3d_fused_columns(dims, prec):
    """Fused 3D cycle with LDS-DMA off, so every column-in pass (X->Y+FFT_Y, Y->Z+FFT_Z,
    Z->X+FFT_X) and the in-place column pass (FFT_Y) run the LDS-twiddle column kernel."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + 2 * ny + 3 * nz)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(cdt(prec))
    D.set_config(enable_fused=True, lds_dma=0, twl=2)
    plan = D.PlanC2C(dims, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), x.dtype).astype(np.complex128)
    assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), x.dtype).astype(np.complex128)
    assert rel(back, OP.execute_generic(z, False)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims", [(64, 32, 16), (60, 21, 10), (1024, 8, 4)])
def test_r2c_3d_fused_c2r(dims, prec):
    """Fused R2C: the backward's Z->X transpose fused with the C2R along X runs the
    LDS-twiddle C2R column kernel (and the complex passes the LDS-twiddle C2C kernels)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 3 + ny)
    x = rng.standard_normal((nz, ny, nx)).astype(rdt(prec))
    D.set_config(enable_fused=True, lds_dma=0, twl=2)
    plan = D.PlanR2C(dims, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    nc = nx // 2 + 1
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nc, nz), cdt(prec)).astype(np.complex128)
    assert rel(z, OP.execute_generic_r2c(x.astype(np.float64))) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), rdt(prec)).astype(np.float64)
    assert rel(back, OP.execute_generic_c2r(z, nx)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims", [(16, 10), (60, 21), (2048, 12)])
def test_r2c_2d_fused_c2r(dims, prec):
    nx, ny = dims
    rng = np.random.default_rng(nx + ny)
    x = rng.standard_normal((ny, nx)).astype(rdt(prec))
    D.set_config(enable_fused=True, lds_dma=0, twl=2)
    plan = D.PlanR2C(dims, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    nc = nx // 2 + 1
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((nc, ny), cdt(prec)).astype(np.complex128)
    assert rel(z, OP.execute_2d_r2c(x.astype(np.float64))) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((ny, nx), rdt(prec)).astype(np.float64)
    assert rel(back, OP.execute_2d_c2r(z, nx)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("glob", [64, 2048])
def test_unpack_fft_segmented_rows(glob, prec):
    """Fused unpack+FFT (segmented rows read from the per-peer chunks of a receive buffer)
    with the LDS twiddle table; the FFT axis is the full global extent."""
    Pn, me = 4, 2
    a, b, c = 64, glob, 4
    sh = D.ExchangeShare(a, b, c, Pn, me, prec, fused=True)
    dt = cdt(prec)
    rng = np.random.default_rng(glob)
    n_a = a * (b // Pn) * c
    A = (rng.uniform(-1, 1, n_a) + 1j * rng.uniform(-1, 1, n_a)).astype(dt)
    raw = A.view(O.elem_dtype(sh.es))
    send = np.zeros(sh.bytes // sh.es, dtype=raw.dtype)
    for i in range(1, Pn + 1):
        O.host_kernel(O.KERNEL_PACK_FORWARD, sh.a_dims, raw, send, sh.send_rows, i)
    Bw = np.zeros(int(np.prod(sh.b_dims)), dtype=raw.dtype)
    O.host_kernel(O.KERNEL_UNPACK, sh.b_dims, send, Bw, sh.recv_rows)
    dA = D.DeviceArray.from_numpy(A)
    ds, dr = D.DeviceArray(sh.bytes), D.DeviceArray(sh.bytes)
    dB = D.DeviceArray(Bw.size * sh.es)
    for sign in (D.FFT_FORWARD, D.FFT_BACKWARD):
        sh.run(dA, ds, dr, dB, sign)
        got = dB.download(Bw.size, dt).astype(np.complex128).reshape(-1, b)
        want = O.fft_c2c(Bw.view(dt).astype(np.complex128).reshape(-1, b), sign)
        assert rel(got, want) < TOL[prec], sign
