"""CPU checks (no GPU) of the any-length transforms: every kernel a plan picks for lengths
with a prime factor above 13, for odd real lengths and for long radix-3 schedules compiles
for gfx950 (compile-only mode: hipRTC runs, nothing is launched or allocated), and a length
whose Bluestein work line (>= 2n - 1 points) does not fit the 160 KiB of LDS of one
workgroup is refused with DTFFT_HIP_ERROR_UNSUPPORTED_SIZE (602)."""
import pytest

D = pytest.importorskip("dtfft_amd")
P = D.Precision


@pytest.fixture(autouse=True)
def compile_only():
    if D.device_arch().startswith("gfx950"):
        pytest.skip("a GPU is present; compile-only behaviour is exercised on CPU hosts")
    yield
    D.set_config()


def has_large_prime(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n > 1


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [17, 1022, 4099, 6561])
def test_c2c_executor_compiles(n, prec):
    D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, 5, [n])


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
@pytest.mark.parametrize("n", [1, 35, 74, 1023])
def test_r2c_executor_compiles(n, prec):
    D.FFTExecutor(1, D.FFTType.R2C, prec, n, n // 2 + 1, 5, [n], [n], [n // 2 + 1])


@pytest.mark.parametrize("fused", [True, False])
def test_plans_compile(fused):
    D.set_config(enable_fused=fused)
    for prec in (P.SINGLE, P.DOUBLE):
        D.PlanC2C([37, 20, 17], prec, D.Effort.ESTIMATE, D.Executor.HIP)
        D.PlanR2C([74, 6, 8], prec, D.Effort.ESTIMATE, D.Executor.HIP)
        D.PlanR2C([35, 44], prec, D.Effort.ESTIMATE, D.Executor.HIP)
        D.PlanR2R([512, 12], [D.R2RKind.DCT_1, D.R2RKind.DST_1], prec, D.Effort.ESTIMATE, D.Executor.HIP)


@pytest.mark.parametrize("n,prec", [(12289, P.SINGLE), (6007, P.DOUBLE)])
def test_too_long_for_lds_is_refused(n, prec):
    assert has_large_prime(n)
    with pytest.raises(D.DtfftError) as e:
        D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, 1, [n])
    assert e.value.code == 602
