"""GPU parity of single-rank plans (plain and fused paths, generic and Z-slab, 2D and
3D, C2C and R2C) against the oracle's restatement of src/dtfft_plan.F90's execute
sequences, plus full-size size-independent checks of the transpose-only cycle.

Bars: transposes bit-exact; FP64 results within 1e-12 relative L2 of the oracle,
FP32 within 1e-6; round trips bwd(fwd(x))/N == x within 1e-13 (FP64) / 1e-6 (FP32)."""
import numpy as np
import pytest

import oracle as O
import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision
TOL = {P.DOUBLE: 1e-12, P.SINGLE: 1e-6}
RT = {P.DOUBLE: 1e-13, P.SINGLE: 1e-6}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdt(prec):
    return np.complex128 if prec == P.DOUBLE else np.complex64


def rdt(prec):
    return np.float64 if prec == P.DOUBLE else np.float32


def make_plan(cls, dims, prec, fused=True, z_slab=False, executor=D.Executor.HIP):
    D.set_config(enable_z_slab=z_slab, enable_fused=fused)
    try:
        return cls(dims, prec, D.Effort.ESTIMATE, executor)
    finally:
        D.set_config()


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("z_slab", [False, True])
@pytest.mark.parametrize("dims", [(12, 10, 6), (48, 40, 24), (60, 84, 35)])
def test_c2c_3d(dims, z_slab, fused, prec):
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * ny * nz)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(cdt(prec))
    plan = make_plan(D.PlanC2C, dims, prec, fused, z_slab)
    assert plan.z_slab_enabled == z_slab
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), x.dtype)
    x64 = x.astype(np.complex128)
    want = OP.execute_z_slab(x64, True) if z_slab else OP.execute_generic(x64, True)
    assert rel(z.astype(np.complex128), want) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), x.dtype).astype(np.complex128) / x.size
    assert rel(back, x64) < RT[prec] * 10


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
def test_c2c_2d_256(fused, prec):
    """Config: 2D C2C 256x256, one rank; backward(forward(x))/N == x to 1e-13 (FP64)."""
    nx = ny = 256
    rng = np.random.default_rng(256)
    x = (rng.uniform(-1, 1, (ny, nx)) + 1j * rng.uniform(-1, 1, (ny, nx))).astype(cdt(prec))
    plan = make_plan(D.PlanC2C, (nx, ny), prec, fused)
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    y = b.download((nx, ny), x.dtype).astype(np.complex128)
    assert rel(y, OP.execute_2d(x.astype(np.complex128), True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((ny, nx), x.dtype).astype(np.complex128) / x.size
    assert np.abs(back - x).max() < (1e-13 if prec == P.DOUBLE else 1e-5)


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("dims", [(32, 64, 128), (64, 128, 256), (64, 32, 48)])
def test_c2c_3d_lds_dma(dims, mode, prec):
    """Fused cycle with every eligible column-in pass forced onto the LDS-DMA kernels
    (mode 2: double buffer, mode 3: single buffer, two blocks per CU); (64, 128, 256) gives
    several tiles per persistent block, so the next-tile loads overlap the current tile."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 7 + ny)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(cdt(prec))
    D.set_config(enable_fused=True, lds_dma=mode)
    try:
        plan = D.PlanC2C(dims, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), x.dtype).astype(np.complex128)
    assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), x.dtype).astype(np.complex128) / x.size
    assert rel(back, x.astype(np.complex128)) < RT[prec] * 10


def test_transpose_only_2d_bit_exact():
    nx, ny = 256, 256
    rng = np.random.default_rng(5)
    x = (rng.uniform(-1, 1, (ny, nx)) + 1j * rng.uniform(-1, 1, (ny, nx)))
    plan = make_plan(D.PlanC2C, (nx, ny), P.DOUBLE, executor=D.Executor.NONE)
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    assert b.download((nx, ny), x.dtype).tobytes() == np.ascontiguousarray(x.T).tobytes()
    want = np.empty(x.size, dtype=O.elem_dtype(16))
    O.host_kernel(O.KERNEL_PERMUTE_FORWARD, (nx, ny), x.ravel().view(O.elem_dtype(16)), want)
    assert b.download(x.size, O.elem_dtype(16)).tobytes() == want.tobytes()


@pytest.mark.parametrize("z_slab", [False, True])
def test_transpose_api_bit_exact(z_slab):
    dims = (33, 77, 21)
    nx, ny, nz = dims
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))
    plan = make_plan(D.PlanC2C, dims, P.DOUBLE, z_slab=z_slab, executor=D.Executor.NONE)
    raw = x.ravel().view(O.elem_dtype(16))
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    T = D.Transpose
    steps = ([(T.X_TO_Z, O.KERNEL_PERMUTE_BACKWARD, dims), (T.Z_TO_X, O.KERNEL_PERMUTE_FORWARD, (nz, nx, ny))]
             if z_slab else
             [(T.X_TO_Y, O.KERNEL_PERMUTE_FORWARD, dims), (T.Y_TO_Z, O.KERNEL_PERMUTE_FORWARD, (ny, nz, nx)),
              (T.Z_TO_Y, O.KERNEL_PERMUTE_BACKWARD, (nz, nx, ny)),
              (T.Y_TO_X, O.KERNEL_PERMUTE_BACKWARD, (ny, nz, nx))])
    cur = raw
    for tt, kt, kd in steps:
        want = np.empty_like(cur)
        O.host_kernel(kt, kd, cur, want)
        plan.transpose(a, b, tt)
        got = b.download(cur.size, cur.dtype)
        assert got.tobytes() == want.tobytes(), tt
        a, b = b, a
        cur = want
    assert cur.tobytes() == raw.tobytes()  # the cycle returns to the input
    with pytest.raises(D.DtfftError) as e:
        plan.transpose(a, a, steps[0][0])
    assert e.value.code == 14  # DTFFT_ERROR_INPLACE_TRANSPOSE


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("z_slab", [False, True])
@pytest.mark.parametrize("dims", [(16, 10, 6), (64, 32, 16), (60, 21, 10)])
def test_r2c_3d(dims, z_slab, fused, prec):
    """3D R2C on one rank against the oracle's generic R2C sequence; fused: R2C along X,
    X->Y+FFT_Y, Y->Z+FFT_Z forward, and Z, Y in place, Z->X fused with the C2R backward."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + ny + nz)
    x = rng.standard_normal((nz, ny, nx)).astype(rdt(prec))
    plan = make_plan(D.PlanR2C, dims, prec, fused, z_slab=z_slab)
    nc = nx // 2 + 1
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nc, nz), cdt(prec)).astype(np.complex128)
    # both paths end in the Z-pencil [Z, X, Y] holding the full 3D R2C transform
    want = OP.execute_generic_r2c(x.astype(np.float64))
    assert rel(z, want) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), rdt(prec)).astype(np.float64) / x.size
    assert rel(back, x.astype(np.float64)) < RT[prec] * 10


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dims", [(16, 10), (256, 256), (60, 21)])
def test_r2c_2d(dims, fused, prec):
    """2D R2C on one rank (execute_2d order); fused: the Y->X transpose is fused into the
    C2R along X."""
    nx, ny = dims
    rng = np.random.default_rng(nx * ny)
    x = rng.standard_normal((ny, nx)).astype(rdt(prec))
    plan = make_plan(D.PlanR2C, dims, prec, fused)
    nc = nx // 2 + 1
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((nc, ny), cdt(prec)).astype(np.complex128)
    assert rel(z, OP.execute_2d_r2c(x.astype(np.float64))) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((ny, nx), rdt(prec)).astype(np.float64) / x.size
    assert rel(back, x.astype(np.float64)) < RT[prec] * 10


def test_transpose_cycle_512_c128_full_size():
    """Config: 512^3 complex128 transpose-only XYZ->YZX->ZXY->XYZ, bit-exact.  The first
    permute is compared with the oracle in full; the cycle must return the input."""
    n = 512
    rng = np.random.default_rng(2)
    x = np.empty(n ** 3, dtype=np.complex128)
    x.real = rng.uniform(-1, 1, n ** 3)
    x.imag = rng.uniform(-1, 1, n ** 3)
    raw = x.view(O.elem_dtype(16))
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    k = D.Kernel(D.KernelType.PERMUTE_FORWARD, (n, n, n), 16)
    k.execute(a, b, sync=True)
    want = np.empty_like(raw)
    O.host_kernel(O.KERNEL_PERMUTE_FORWARD, (n, n, n), raw, want)
    got = b.download(raw.size, raw.dtype)
    assert got.tobytes() == want.tobytes()
    del got, want
    k.execute(b, a)
    k.execute(a, b, sync=True)
    assert b.download(raw.size, raw.dtype).tobytes() == raw.tobytes()


def test_c2c_512_fp32_lds_dma_matches_default():
    """512^3 FP32 fused cycle with the LDS double-buffer column kernel forced (64 tiles per
    persistent block) against the default kernels: forward outputs within 1e-6 of each
    other, round trip within 1e-6."""
    n = 512
    rng = np.random.default_rng(9)
    x = np.empty(n ** 3, dtype=np.complex64)
    x.real = rng.uniform(-1, 1, n ** 3)
    x.imag = rng.uniform(-1, 1, n ** 3)
    outs = []
    for mode in (0, 2, 3):
        D.set_config(enable_fused=True, lds_dma=mode)
        try:
            plan = D.PlanC2C((n, n, n), P.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)
        finally:
            D.set_config()
        a = D.DeviceArray.from_numpy(x)
        b = D.DeviceArray(x.nbytes)
        plan.execute(a, b, D.Execute.FORWARD)
        outs.append(b.download(n ** 3, np.complex64))
        if mode >= 2:
            c = D.DeviceArray(x.nbytes)
            plan.execute(b, c, D.Execute.BACKWARD)
            back = c.download(n ** 3, np.complex64).astype(np.complex128) / n ** 3
            assert rel(back, x.astype(np.complex128)) < 1e-6
            del c
        del a, b, plan
    for o in outs[1:]:
        assert rel(o.astype(np.complex128), outs[0].astype(np.complex128)) < 1e-6


def test_c2c_512_fp32_round_trip_full_size():
    """Metric workload at full size: 512^3 C2C FP32 fused, bwd(fwd(x))/N == x within 1e-6 rel L2."""
    n = 512
    rng = np.random.default_rng(3)
    x = np.empty(n ** 3, dtype=np.complex64)
    x.real = rng.uniform(-1, 1, n ** 3).astype(np.float32)
    x.imag = rng.uniform(-1, 1, n ** 3).astype(np.float32)
    plan = make_plan(D.PlanC2C, (n, n, n), P.SINGLE, fused=True)
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    c = D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download(x.shape, x.dtype)
    err = np.linalg.norm((back / np.float32(n ** 3)).astype(np.complex128) - x) / np.linalg.norm(x)
    assert err < 1e-6, err
