"""GPU parity of the chunk-blocked row -> column-in pairs of the fused C2C sequences.

A row FFT pass whose output the next pass reads as column tiles writes it chunk-blocked
(each tile's lines as one contiguous block); the intermediate layout is private to the plan.
The blocked sequence runs the same butterflies as the natural one on other addresses; both
must match the FP64 oracle (src/dtfft_plan.F90's execute sequences, restated in
oracle/plan.py) within 1e-6 (FP32) / 1e-12 (FP64) relative L2, forward and backward, and
each other within a few units in the last place.  They are not always bit-identical: the row
kernel's two instantiations (natural and chunked output) may contract a different product
of a butterfly into an FMA (measured on MI355X: 256-point FP64 rows, outputs past the first
radix-16 group, 1-ulp differences at 3e-16 of the oracle either way).  Shapes cover the LDS-DMA column kernel
(FP32, FFT lengths divisible by 8), the plain column kernel (FP64, and lengths such as 20),
pairs built in one direction only, partial row tiles, 2D plans, and the padded pair of the FP32
R2C forward (R2C rows -> X->Y)."""
import numpy as np
import pytest

import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision
TOL = {P.DOUBLE: 1e-12, P.SINGLE: 1e-6}
ULPS = {P.DOUBLE: 1e-15, P.SINGLE: 5e-7}   # blocked vs natural, relative L2


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdt(prec):
    return np.complex128 if prec == P.DOUBLE else np.complex64


def plan_and_stages(dims, prec, blocked, capfd):
    D.set_config(enable_fused=True, blocked=blocked)
    try:
        p = D.PlanC2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    capfd.readouterr()
    p.report()
    names = [ln.split()[0] for ln in capfd.readouterr().out.splitlines()[1:] if ln.strip()]
    return p, names


def run(p, x, out_shape, fwd):
    a = D.DeviceArray.from_numpy(x, p.alloc_bytes)
    b = D.DeviceArray(p.alloc_bytes)
    p.execute(a, b, D.Execute.FORWARD if fwd else D.Execute.BACKWARD)
    return b.download(out_shape, x.dtype)


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims", [
    (48, 40, 32),     # plain column kernels (lengths 40 / 48 with radix 5 / 3)
    (64, 20, 16),     # LDS-DMA column kernel backward (FP32), plain forward
    (32, 24, 40),     # FP32: Z = 40 is not a multiple of the 16-line chunk, backward stays natural
    (16, 512, 48),    # one chunk along X, LDS-DMA 512-point columns (FP32)
    (96, 13, 16),     # 13 rows per plane: partial row tiles in the blocked row pass
])
def test_blocked_3d_matches_natural_and_oracle(dims, prec, capfd):
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + 7 * ny + 31 * nz)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(cdt(prec))
    pb, names_b = plan_and_stages(dims, prec, True, capfd)
    pn, names_n = plan_and_stages(dims, prec, False, capfd)
    # chunks are 128 bytes of lines (16 FP32 / 8 FP64) or a plain column kernel's 8 / 16 lines
    assert "fwd_rows" in names_b, names_b
    if nz % 16 == 0:
        assert "bwd_rows" in names_b, names_b
    if prec == P.SINGLE and nz % 16 != 0:
        assert "bwd_rows" not in names_b, names_b
    assert "fwd_rows" not in names_n and "bwd_rows" not in names_n, names_n
    x64 = x.astype(np.complex128)
    zb = run(pb, x, (ny, nx, nz), True)
    zn = run(pn, x, (ny, nx, nz), True)
    assert rel(zb.astype(np.complex128), zn.astype(np.complex128)) < ULPS[prec]
    want = OP.execute_generic(x64, True)
    assert rel(zb.astype(np.complex128), want) < TOL[prec]
    # backward from a Z-pencil of its own
    y = (rng.uniform(-1, 1, (ny, nx, nz)) + 1j * rng.uniform(-1, 1, (ny, nx, nz))).astype(cdt(prec))
    xb = run(pb, y, (nz, ny, nx), False)
    xn = run(pn, y, (nz, ny, nx), False)
    assert rel(xb.astype(np.complex128), xn.astype(np.complex128)) < ULPS[prec]
    assert rel(xb.astype(np.complex128), OP.execute_generic(y.astype(np.complex128), False)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
def test_blocked_forward_in_place_takes_natural_layout(prec, capfd):
    """in == out: the blocked row pass would overwrite rows it has not read yet, so the plan
    runs the natural pair; the result still matches the oracle."""
    dims = (48, 40, 32)
    nx, ny, nz = dims
    rng = np.random.default_rng(3)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(cdt(prec))
    p, names = plan_and_stages(dims, prec, True, capfd)
    assert "fwd_rows" in names
    a = D.DeviceArray.from_numpy(x, p.alloc_bytes)
    p.execute(a, a, D.Execute.FORWARD)
    z = a.download((ny, nx, nz), x.dtype).astype(np.complex128)
    assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < TOL[prec]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims", [(64, 48), (256, 256), (32, 40)])
def test_blocked_2d_matches_natural_and_oracle(dims, prec, capfd):
    nx, ny = dims
    rng = np.random.default_rng(nx * ny)
    x = (rng.uniform(-1, 1, (ny, nx)) + 1j * rng.uniform(-1, 1, (ny, nx))).astype(cdt(prec))
    pb, names_b = plan_and_stages(dims, prec, True, capfd)
    pn, _ = plan_and_stages(dims, prec, False, capfd)
    assert "fwd_rows" in names_b, names_b
    if ny % 16 == 0:
        assert "bwd_rows" in names_b, names_b
    yb = run(pb, x, (nx, ny), True)
    yn = run(pn, x, (nx, ny), True)
    assert rel(yb.astype(np.complex128), yn.astype(np.complex128)) < ULPS[prec]
    assert rel(yb.astype(np.complex128), OP.execute_2d(x.astype(np.complex128), True)) < TOL[prec]
    y = (rng.uniform(-1, 1, (nx, ny)) + 1j * rng.uniform(-1, 1, (nx, ny))).astype(cdt(prec))
    xb = run(pb, y, (ny, nx), False)
    xn = run(pn, y, (ny, nx), False)
    assert rel(xb.astype(np.complex128), xn.astype(np.complex128)) < ULPS[prec]
    assert rel(xb.astype(np.complex128), OP.execute_2d(y.astype(np.complex128), False)) < TOL[prec]


@pytest.mark.parametrize("dims,prec", [
    ((512, 32, 16), P.SINGLE),    # nc = 257 complex lines per row, padded to 272 (17 chunks)
    ((480, 64, 8), P.SINGLE),     # nc = 241, padded to 256
    ((960, 32, 8), P.SINGLE),     # nc = 481, padded to 496; radix-3 and -5 rows
    ((480, 512, 4), P.DOUBLE),    # FP64: 16-line tiles of the plain column kernel (512 points)
])
def test_blocked_r2c_matches_natural_and_oracle(dims, prec, capfd):
    """3D R2C, fused: the R2C rows write the complex X-pencil chunk-blocked and padded to whole
    16-line chunks (nc = nx/2 + 1 lines -> ncp), so that the X->Y pass reads contiguous tiles
    (FP32: the LDS-DMA column kernel); the padding lines are never written and their transforms
    land in Y-pencil planes that the Y->Z pass does not read.  The plan's buffers grow to
    ncp * ny * nz complex values.  Forward against the natural layout and the FP64 oracle
    (src/dtfft_plan.F90's R2C execute sequence, oracle/plan.py), and the round trip."""
    nx, ny, nz = dims
    nc = nx // 2 + 1
    ncp = (nc + 15) // 16 * 16
    rdt = np.float64 if prec == P.DOUBLE else np.float32
    rng = np.random.default_rng(nx + ny + nz)
    x = rng.standard_normal((nz, ny, nx)).astype(rdt)
    plans = {}
    for blocked in (True, False):
        D.set_config(enable_fused=True, blocked=blocked)
        try:
            plans[blocked] = D.PlanR2C([nx, ny, nz], prec, D.Effort.ESTIMATE, D.Executor.HIP)
        finally:
            D.set_config()
    capfd.readouterr()
    plans[True].report()
    names = [ln.split()[0] for ln in capfd.readouterr().out.splitlines()[1:] if ln.strip()]
    assert "fwd_rows" in names, names
    assert plans[True].alloc_size == max(nx * ny * nz, 2 * ncp * ny * nz)
    assert plans[False].alloc_size == max(nx * ny * nz, 2 * nc * ny * nz)
    outs = {}
    for blocked, p in plans.items():
        a = D.DeviceArray.from_numpy(x, p.alloc_bytes)
        b = D.DeviceArray(p.alloc_bytes)
        c = D.DeviceArray(p.alloc_bytes)
        p.execute(a, b, D.Execute.FORWARD)
        z = b.download((ny, nc, nz), cdt(prec)).astype(np.complex128)
        p.execute(b, c, D.Execute.BACKWARD)
        back = c.download((nz, ny, nx), rdt).astype(np.float64) / x.size
        outs[blocked] = (z, back)
    want = OP.execute_generic_r2c(x.astype(np.float64))
    assert rel(outs[True][0], want) < TOL[prec]
    assert rel(outs[True][0], outs[False][0]) < ULPS[prec]
    assert rel(outs[True][1], x.astype(np.float64)) < TOL[prec] * 10
