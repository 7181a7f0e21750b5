"""CPU checks (no GPU) that every kernel a plan picks fits gfx950's 160 KiB of LDS per
workgroup, for long axes where the twiddle table kept in LDS competes with the tile.

hipRTC refuses a kernel whose LDS exceeds the limit, so creating the plan in compile-only
mode (the kernels are compiled for gfx950, nothing is launched or allocated) is the check.
The shapes are those at which the LDS-twiddle form would not fit: 2048-point column-in
passes of 8 lines, the fused C2R of a 2048-point R2C axis, 8192-point rows, the fused
unpack+FFT of an 8192-point aligned axis and batched stages with a 2048-point axis."""
import pytest

D = pytest.importorskip("dtfft_amd")


@pytest.fixture(autouse=True)
def compile_only():
    if D.device_arch().startswith("gfx950"):
        pytest.skip("a GPU is present; compile-only behaviour is exercised on CPU hosts")
    yield
    D.set_config()


@pytest.mark.parametrize("twl", [1, 2])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dims", [(2048, 128, 128), (8192, 64, 32), (4096, 16, 16)])
def test_c2c_long_axes_fit(dims, fused, twl):
    D.set_config(enable_fused=fused, twl=twl)
    D.PlanC2C(list(dims), D.Precision.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
@pytest.mark.parametrize("twl", [1, 2])
@pytest.mark.parametrize("dims", [(2048, 256, 128), (4096, 8, 8)])
def test_r2c_long_axes_fit(dims, twl, prec):
    D.set_config(enable_fused=True, twl=twl)
    D.PlanR2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)


@pytest.mark.parametrize("twl", [1, 2])
def test_unpack_fft_long_axis_fits(twl):
    from dtfft_amd import pencil
    D.set_config(twl=twl)
    P, a, b, c = 4, 256, 8192, 32   # receive pencil (8192, 32, 64): 2048 lines of 8192 points
    _, recv, _, _ = pencil.forward_exchange(a, b, c, P, 1)
    D.UnpackFFT((b, c, a // P), recv, D.Precision.SINGLE)


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
def test_executor_forced_twl_fits(prec):
    D.set_config(twl=2)
    for n in (2048, 4096, 8192):
        D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, 4, [n])
        D.FFTExecutor(1, D.FFTType.R2C, prec, n, n // 2 + 1, 4, [n], [n], [n // 2 + 1])


@pytest.mark.parametrize("grids", [[[2048, 8, 8]], [[8, 8, 2048]], [[8, 2048, 8], [1024, 8, 8]]])
def test_cycle_long_axis_fits(grids):
    D.Cycle(grids, D.Precision.SINGLE)


@pytest.mark.parametrize("twl", [1, 2])
@pytest.mark.parametrize("prec,dims,dma", [
    (D.Precision.SINGLE, (1024, 512, 256), True), (D.Precision.DOUBLE, (1024, 512, 256), True),
    (D.Precision.SINGLE, (256, 12, 32), True), (D.Precision.SINGLE, (1024, 6, 16), True),
    (D.Precision.SINGLE, (256, 3, 48), True), (D.Precision.DOUBLE, (512, 6, 16), True),
    (D.Precision.DOUBLE, (1024, 4, 8), True), (D.Precision.DOUBLE, (512, 5, 24), True),
    (D.Precision.SINGLE, (256, 64), True), (D.Precision.SINGLE, (1024, 32), True),
    (D.Precision.DOUBLE, (512, 16), True), (D.Precision.DOUBLE, (1024, 24), True),
    (D.Precision.SINGLE, (64, 32, 16), False), (D.Precision.SINGLE, (1024, 6, 12), False),
])
def test_fused_c2r_lds_dma_kernel_choice(prec, dims, dma, twl, capfd):
    """The fused C2R takes the LDS-DMA kernel where the stage allows it (the GPU parity test
    of that kernel runs these shapes), and it fits the LDS with or without the LDS twiddles."""
    D.set_config(enable_fused=True, twl=twl)
    p = D.PlanR2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
    p.report()
    out = capfd.readouterr().out
    line = [ln for ln in out.splitlines() if "+c2r" in ln][0]
    assert ("fft_c2r_dma" in line) == dma, line
