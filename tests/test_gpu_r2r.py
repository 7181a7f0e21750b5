"""GPU parity of the R2R (DCT / DST 1-4) stages against the R2R oracle (FFTW's
REDFT/RODFT definitions, oracle/r2r.py): the executor plugin at rank 1 and 2, and the
single-rank R2R plans (execute_generic / execute_2d order with R2R stages)."""
import numpy as np
import pytest

import oracle.r2r as R

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision
TOL = {P.DOUBLE: 1e-12, P.SINGLE: 2e-6}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rdt(prec):
    return np.float64 if prec == P.DOUBLE else np.float32


def fused_form(kind, n):
    """The fused R2R kernel needs a complex work length that factors into primes <= 13;
    other work lengths run the three launches with a Bluestein C2C stage."""
    m = R.logical_size(kind, n) if kind in (R.DCT_1, R.DST_1) else (2 * n if kind in (R.DCT_4, R.DST_4) else n)
    for p in (2, 3, 5, 7, 11, 13):
        while m % p == 0:
            m //= p
    return m == 1


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [2, 8, 17, 23, 64])
@pytest.mark.parametrize("kind", R.KINDS)
def test_executor_rank1(kind, n, prec):
    hm = 37
    rng = np.random.default_rng(kind * 100 + n)
    x = rng.standard_normal((hm, n)).astype(rdt(prec))
    ex = D.FFTExecutor(1, D.FFTType.R2R, prec, n, n, hm, [n], r2r_kinds=[kind])
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    ex.execute(a, b, D.FFT_FORWARD)
    y = b.download((hm, n), x.dtype).astype(np.float64)
    assert rel(y, R.r2r(x.astype(np.float64), kind)) < TOL[prec]
    ex.execute(b, a, D.FFT_BACKWARD)  # inverse kind
    back = a.download((hm, n), x.dtype).astype(np.float64) / R.logical_size(kind, n)
    assert rel(back, x.astype(np.float64)) < TOL[prec] * 4


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("kinds", [(5, 6), (3, 9), (8, 4), (10, 7)])
def test_executor_rank2(kinds, prec):
    n0, n1, hm = 12, 20, 3  # FFTW order: n0 slow, n1 fast; kinds[i] along fft_sizes[i]
    rng = np.random.default_rng(sum(kinds))
    x = rng.standard_normal((hm, n0, n1)).astype(rdt(prec))
    ex = D.FFTExecutor(2, D.FFTType.R2R, prec, n0 * n1, n0 * n1, hm, [n0, n1], r2r_kinds=list(kinds))
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    ex.execute(a, a, D.FFT_FORWARD)  # in place
    y = a.download((hm, n0, n1), x.dtype).astype(np.float64)
    want = R.r2r(x.astype(np.float64), kinds[1])
    want = np.swapaxes(R.r2r(np.swapaxes(want, 1, 2), kinds[0]), 1, 2)
    assert rel(y, want) < TOL[prec]
    ex.execute(a, b, D.FFT_BACKWARD)
    scale = R.logical_size(kinds[0], n0) * R.logical_size(kinds[1], n1)
    back = b.download((hm, n0, n1), x.dtype).astype(np.float64) / scale
    assert rel(back, x.astype(np.float64)) < TOL[prec] * 4


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims,kinds", [((16, 10, 6), (5, 6, 9)), ((17, 12, 9), (3, 8, 10)),
                                        ((32, 24, 20), (4, 7, 5))])
def test_plan_r2r_3d(dims, kinds, prec):
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + ny + nz)
    x = rng.standard_normal((nz, ny, nx)).astype(rdt(prec))
    plan = D.PlanR2R(dims, kinds, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    assert plan.element_size == x.itemsize and plan.alloc_size == x.size
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    c = D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), x.dtype).astype(np.float64)
    assert rel(z, R.execute_generic_r2r(x.astype(np.float64), kinds, True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    scale = np.prod([R.logical_size(k, n) for k, n in zip(kinds, dims)])
    back = c.download((nz, ny, nx), x.dtype).astype(np.float64) / scale
    assert rel(back, x.astype(np.float64)) < TOL[prec] * 4


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
def test_plan_r2r_2d(prec):
    nx, ny = 24, 18
    kinds = (R.DCT_2, R.DST_4)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((ny, nx)).astype(rdt(prec))
    plan = D.PlanR2R((nx, ny), kinds, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    c = D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    y = b.download((nx, ny), x.dtype).astype(np.float64)
    assert rel(y, R.execute_2d_r2r(x.astype(np.float64), kinds, True)) < TOL[prec]
    plan.execute(b, c, D.Execute.BACKWARD)
    scale = R.logical_size(kinds[0], nx) * R.logical_size(kinds[1], ny)
    back = c.download((ny, nx), x.dtype).astype(np.float64) / scale
    assert rel(back, x.astype(np.float64)) < TOL[prec] * 4


def test_transpose_only_r2r_plan_bit_exact():
    """Transpose-only R2R plan (no kinds): real elements moved exactly, X -> Z pencil."""
    dims = (20, 12, 7)
    nx, ny, nz = dims
    x = np.random.default_rng(4).standard_normal((nz, ny, nx)).astype(np.float32)
    plan = D.PlanR2R(dims, None, P.SINGLE, D.Effort.ESTIMATE, D.Executor.NONE)
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), np.float32)
    assert z.tobytes() == np.ascontiguousarray(np.transpose(x, (1, 2, 0))).tobytes()


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [64, 512])
@pytest.mark.parametrize("kind", R.KINDS)
def test_fused_stage_matches_three_launch_stage(kind, n, prec):
    """Contiguous lines with one layout in and out run the fused kernel (real lines through
    LDS); an output line stride that differs from the input's runs the three launches through
    the complex work buffer.  Both must give the oracle's result on the same lines."""
    print(f"case: R2R kind {kind} n={n} {prec.name} (fused form: {fused_form(kind, n)})", flush=True)
    hm = 129  # not a multiple of the lines per block: a partial last tile
    rng = np.random.default_rng(kind * 1000 + n)
    x = rng.standard_normal((hm, n)).astype(rdt(prec))
    want = R.r2r(x.astype(np.float64), kind)
    fused = D.FFTExecutor(1, D.FFTType.R2R, prec, n, n, hm, [n], r2r_kinds=[kind])
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    fused.execute(a, b, D.FFT_FORWARD)
    y_fused = b.download((hm, n), x.dtype).astype(np.float64)
    od = n + 3
    three = D.FFTExecutor(1, D.FFTType.R2R, prec, n, od, hm, [n], r2r_kinds=[kind])
    c = D.DeviceArray(hm * od * x.itemsize)
    three.execute(a, c, D.FFT_FORWARD)
    y_three = c.download((hm, od), x.dtype)[:, :n].astype(np.float64)
    assert rel(y_fused, want) < TOL[prec]
    assert rel(y_three, want) < TOL[prec]
    # in place through the fused kernel
    fused.execute(a, a, D.FFT_FORWARD)
    assert a.download((hm, n), x.dtype).tobytes() == b.download((hm, n), x.dtype).tobytes()


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
def test_executor_rank2_padded_batches(prec):
    """Batches farther apart than one grid (idist = odist > n0 * n1): the fast-axis stage cannot
    fold its batches into one evenly spaced range of lines, so it runs the three launches; the
    slow-axis stage runs them anyway (strided lines).  The result is still the oracle's."""
    n0, n1, hm = 10, 16, 3
    dist = n0 * n1 + 24
    kinds = (R.DCT_2, R.DST_4)
    rng = np.random.default_rng(11)
    x = np.zeros((hm, dist), dtype=rdt(prec))
    x[:, :n0 * n1] = rng.standard_normal((hm, n0 * n1))
    ex = D.FFTExecutor(2, D.FFTType.R2R, prec, dist, dist, hm, [n0, n1], r2r_kinds=list(kinds))
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    ex.execute(a, b, D.FFT_FORWARD)
    y = b.download((hm, dist), x.dtype)[:, :n0 * n1].reshape(hm, n0, n1).astype(np.float64)
    g = x[:, :n0 * n1].reshape(hm, n0, n1).astype(np.float64)
    want = R.r2r(g, kinds[1])
    want = np.swapaxes(R.r2r(np.swapaxes(want, 1, 2), kinds[0]), 1, 2)
    assert rel(y, want) < TOL[prec]


def _plan_r2r(dims, kinds, prec, fused, capfd):
    D.set_config(enable_fused=fused)
    try:
        p = D.PlanR2R(dims, kinds, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    capfd.readouterr()
    p.report()
    return p, "fused transposes" in capfd.readouterr().out


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims,kinds", [((64, 32, 96), (5, 5, 5)), ((32, 64, 32), (4, 9, 8)),
                                        ((96, 32, 64), (8, 4, 5))])
def test_plan_r2r_3d_fused_transposes(dims, kinds, prec, capfd):
    """The fused R2R sequence (transposes folded into the stages that read across lines: X->Y
    + R2R_Y, Y->Z + R2R_Z forward; Z->X + R2R_X and R2R_Y in place backward) against the
    oracle's reference sequence, forward and backward, and against the plan's own sequence with
    permutes (the same butterflies; a few units in the last place apart)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    x = rng.standard_normal((nz, ny, nx)).astype(rdt(prec))
    pf, is_fused = _plan_r2r(dims, kinds, prec, True, capfd)
    pp, plain_fused = _plan_r2r(dims, kinds, prec, False, capfd)
    assert is_fused and not plain_fused
    outs = {}
    for name, p in (("fused", pf), ("plain", pp)):
        a = D.DeviceArray.from_numpy(x)
        b = D.DeviceArray(x.nbytes)
        p.execute(a, b, D.Execute.FORWARD)
        outs[name] = b.download((ny, nx, nz), x.dtype).astype(np.float64)
    assert rel(outs["fused"], R.execute_generic_r2r(x.astype(np.float64), kinds, True)) < TOL[prec]
    assert rel(outs["fused"], outs["plain"]) < TOL[prec] / 4
    y = rng.standard_normal((ny, nx, nz)).astype(rdt(prec))
    for name, p in (("fused", pf), ("plain", pp)):
        a = D.DeviceArray.from_numpy(y)
        b = D.DeviceArray(y.nbytes)
        p.execute(a, b, D.Execute.BACKWARD)
        outs[name] = b.download((nz, ny, nx), y.dtype).astype(np.float64)
    assert rel(outs["fused"], R.execute_generic_r2r(y.astype(np.float64), kinds, False)) < TOL[prec]
    assert rel(outs["fused"], outs["plain"]) < TOL[prec] / 4


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims,kinds", [((64, 96), (5, 4)), ((32, 32), (9, 8))])
def test_plan_r2r_2d_fused_transposes(dims, kinds, prec, capfd):
    nx, ny = dims
    rng = np.random.default_rng(nx + ny)
    x = rng.standard_normal((ny, nx)).astype(rdt(prec))
    pf, is_fused = _plan_r2r(dims, kinds, prec, True, capfd)
    assert is_fused
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    pf.execute(a, b, D.Execute.FORWARD)
    assert rel(b.download((nx, ny), x.dtype).astype(np.float64), R.execute_2d_r2r(x.astype(np.float64), kinds, True)) < TOL[prec]
    y = rng.standard_normal((nx, ny)).astype(rdt(prec))
    a = D.DeviceArray.from_numpy(y)
    pf.execute(a, b, D.Execute.BACKWARD)
    assert rel(b.download((ny, nx), y.dtype).astype(np.float64), R.execute_2d_r2r(y.astype(np.float64), kinds, False)) < TOL[prec]


@pytest.mark.parametrize("dims,kinds", [((48, 32, 64), (5, 5, 5)), ((64, 32, 64), (5, 6, 5))])
def test_plan_r2r_fused_falls_back_to_permutes(dims, kinds, capfd):
    """A stage without the fused column form (a line count that is not a multiple of the
    32-line tile row, a DCT-4) keeps the sequence with permutes; it still matches the oracle."""
    nx, ny, nz = dims
    x = np.random.default_rng(9).standard_normal((nz, ny, nx)).astype(np.float32)
    p, is_fused = _plan_r2r(dims, kinds, P.SINGLE, True, capfd)
    assert not is_fused
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    p.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), x.dtype).astype(np.float64)
    assert rel(z, R.execute_generic_r2r(x.astype(np.float64), kinds, True)) < TOL[P.SINGLE]


def _plan_r2r_blocked(dims, kinds, prec, blocked, capfd):
    D.set_config(enable_fused=True, blocked=blocked)
    try:
        p = D.PlanR2R(dims, kinds, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    capfd.readouterr()
    p.report()
    return p, [ln.split()[0] for ln in capfd.readouterr().out.splitlines()[1:] if ln.strip()]


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("dims,kinds", [((64, 64, 64), (5, 5, 5)), ((128, 64, 64), (4, 9, 8)),
                                        ((256, 256), (5, 4)), ((64, 128), (9, 8))])
def test_plan_r2r_blocked_pairs(dims, kinds, prec, capfd):
    """The fused R2R sequence with chunk-blocked pairs (the row stage writes its lines in chunks
    of the column stage's tile width, so that the column stage reads each tile as one contiguous
    block: forward R2R_X rows -> X->Y + R2R_Y, backward R2R_Z (3D) / R2R_Y (2D) rows -> the
    column stage into X) against the same sequence in the natural layout and the oracle's
    reference sequence, both directions; forward in place falls back to the natural pair.
    Shapes where the row stage's tiles stay within one plane (its lines per block divide the
    plane's lines), as the pair requires."""
    p3 = len(dims) == 3
    shape_in = tuple(reversed(dims))
    shape_out = (dims[1], dims[0], dims[2]) if p3 else (dims[0], dims[1])
    rng = np.random.default_rng(sum(dims) + 11)
    pb, names = _plan_r2r_blocked(dims, kinds, prec, True, capfd)
    pn, names_n = _plan_r2r_blocked(dims, kinds, prec, False, capfd)
    assert "r2r_frow" in names and "r2r_bcol" in names, names
    assert "r2r_frow" not in names_n and "r2r_brow" not in names_n
    x = rng.standard_normal(shape_in).astype(rdt(prec))
    outs = {}
    for name, p in (("blocked", pb), ("natural", pn)):
        a = D.DeviceArray.from_numpy(x)
        b = D.DeviceArray(x.nbytes)
        p.execute(a, b, D.Execute.FORWARD)
        outs[name] = b.download(shape_out, x.dtype).astype(np.float64)
    ref = R.execute_generic_r2r if p3 else R.execute_2d_r2r
    assert rel(outs["blocked"], ref(x.astype(np.float64), kinds, True)) < TOL[prec]
    assert rel(outs["blocked"], outs["natural"]) < TOL[prec] / 4
    y = rng.standard_normal(shape_out).astype(rdt(prec))
    for name, p in (("blocked", pb), ("natural", pn)):
        a = D.DeviceArray.from_numpy(y)
        b = D.DeviceArray(y.nbytes)
        p.execute(a, b, D.Execute.BACKWARD)
        outs[name] = b.download(shape_in, y.dtype).astype(np.float64)
    assert rel(outs["blocked"], ref(y.astype(np.float64), kinds, False)) < TOL[prec]
    assert rel(outs["blocked"], outs["natural"]) < TOL[prec] / 4
    if p3:
        a = D.DeviceArray.from_numpy(x)
        pb.execute(a, a, D.Execute.FORWARD)
        z = a.download(shape_out, x.dtype).astype(np.float64)
        assert rel(z, ref(x.astype(np.float64), kinds, True)) < TOL[prec]
