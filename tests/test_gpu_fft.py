"""GPU parity of the FFT executor plugin against the FP64 oracle.

Bars (stated per result): FP64 plans within 1e-12 relative L2 of the oracle; FP32
plans within 1e-6 relative L2 of the FP64 oracle evaluated on the same (FP32-rounded)
input.  The oracle itself is pinned by tests/test_oracle_fft.py."""
import os

import numpy as np
import pytest

import oracle as O
import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")

TOL = {D.Precision.DOUBLE: 1e-12, D.Precision.SINGLE: 1e-6}
LENGTHS = [2, 3, 4, 5, 7, 8, 11, 13, 16, 48, 60, 64, 84, 96, 105, 128, 160, 192, 210, 256, 320, 384, 512,
           1024, 2048]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdtype(prec):
    return np.complex128 if prec == D.Precision.DOUBLE else np.complex64


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
@pytest.mark.parametrize("n", LENGTHS)
def test_c2c_1d(n, prec):
    rng = np.random.default_rng(n)
    hm = max(3, 4096 // n)
    x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(cdtype(prec))
    ex = D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, hm, [n])
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    x64 = x.astype(np.complex128)
    for sign in (D.FFT_FORWARD, D.FFT_BACKWARD):
        ex.execute(a, b, sign)
        got = b.download(x.shape, x.dtype).astype(np.complex128)
        want = O.fft_c2c(x64, sign)
        assert rel(got, want) < TOL[prec], (n, sign, rel(got, want))
    # in place
    ex.execute(a, a, D.FFT_FORWARD)
    assert rel(a.download(x.shape, x.dtype).astype(np.complex128), O.fft_c2c(x64)) < TOL[prec]


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
@pytest.mark.parametrize("n", [2, 4, 6, 16, 60, 96, 210, 256, 512, 1024, 2048])
def test_r2c_c2r_1d(n, prec):
    rng = np.random.default_rng(n + 7)
    hm = max(3, 2048 // n)
    rdt = np.float64 if prec == D.Precision.DOUBLE else np.float32
    nc = n // 2 + 1
    x = rng.standard_normal((hm, n)).astype(rdt)
    ex = D.FFTExecutor(1, D.FFTType.R2C, prec, n, nc, hm, [n], [n], [nc])
    a = D.DeviceArray.from_numpy(x)
    c = D.DeviceArray(hm * nc * np.dtype(cdtype(prec)).itemsize)
    ex.execute(a, c, D.FFT_FORWARD)
    got = c.download((hm, nc), cdtype(prec)).astype(np.complex128)
    want = O.fft_r2c(x.astype(np.float64))
    assert rel(got, want) < TOL[prec]
    # C2R of a random half spectrum (imaginary parts of X_0 and X_{n/2} must be ignored)
    h = (rng.standard_normal((hm, nc)) + 1j * rng.standard_normal((hm, nc))).astype(cdtype(prec))
    c.upload(h)
    r = D.DeviceArray(x.nbytes)
    ex.execute(c, r, D.FFT_BACKWARD)
    got = r.download((hm, n), rdt).astype(np.float64)
    want = O.fft_c2r(h.astype(np.complex128), n)
    assert rel(got, want) < TOL[prec]


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
@pytest.mark.parametrize("shape", [(12, 10, 7), (3, 64, 48), (2, 105, 60)])
def test_c2c_2d(shape, prec):
    """fft_rank 2 (Z-slab executor): sizes (n_slow, n_fast) per batch."""
    hm, n0, n1 = shape
    rng = np.random.default_rng(sum(shape))
    x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(cdtype(prec))
    ex = D.FFTExecutor(2, D.FFTType.C2C, prec, n0 * n1, n0 * n1, hm, [n0, n1])
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    ex.execute(a, b, D.FFT_FORWARD)
    got = b.download(shape, x.dtype).astype(np.complex128)
    want = OP._fft_2d(x.astype(np.complex128), O.FFT_FORWARD)
    assert rel(got, want) < TOL[prec]


def test_autotuned_stage_is_correct():
    n, hm = 512, 2048
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(np.complex64)
    ex = D.FFTExecutor(1, D.FFTType.C2C, D.Precision.SINGLE, n, n, hm, [n], effort=D.Effort.MEASURE)
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    ex.execute(a, b, D.FFT_FORWARD)
    got = b.download(x.shape, x.dtype).astype(np.complex128)
    assert rel(got, O.fft_c2c(x.astype(np.complex128))) < 1e-6


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
def test_tuning_records_select_spread_lds_layout(tmp_path, prec):
    """Tuning records are honoured: records written for a row stage (pad 2) and an in-place
    column stage (pad 3) select the spread LDS layout (one spare slot per 16 elements); the
    1D and 2D executors must still match the oracle, both directions."""
    L = D.lib()
    es, c = (16, "d") if prec == D.Precision.DOUBLE else (8, "f")
    n, hm = 512, 64
    n0, n1, hm2 = 256, 128, 2
    recs = {
        f"fft|{n}|{c}|000|{hm}|1|{n},0,1|{n},0,1": "32 4 2 0",
        f"fft|{n0}|{c}|111|{n1}|{hm2}|1,{n0 * n1},{n1}|1,{n0 * n1},{n1}": "8 16 3 0",
    }
    (tmp_path / "tuning.txt").write_text("".join(f"{k}\t{v}\n" for k, v in recs.items()))
    default_dir = os.environ.get("DTFFT_HIP_CACHE_DIR") or os.path.join(os.path.dirname(D._lib.LIB_PATH), "kcache")
    assert L.dtfft_hip_set_cache_dir(str(tmp_path).encode()) == 0
    try:
        rng = np.random.default_rng(5)
        cdt = np.complex128 if es == 16 else np.complex64
        tol = 1e-12 if es == 16 else 1e-6
        x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(cdt)
        ex = D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, hm, [n])
        a = D.DeviceArray.from_numpy(x)
        b = D.DeviceArray(x.nbytes)
        ex.execute(a, b, D.FFT_FORWARD)
        assert rel(b.download(x.shape, cdt).astype(np.complex128), O.fft_c2c(x.astype(np.complex128))) < tol
        ex.execute(b, a, D.FFT_BACKWARD)
        assert rel(a.download(x.shape, cdt).astype(np.complex128) / n, x.astype(np.complex128)) < tol
        y = (rng.uniform(-1, 1, (hm2, n0, n1)) + 1j * rng.uniform(-1, 1, (hm2, n0, n1))).astype(cdt)
        e2 = D.FFTExecutor(2, D.FFTType.C2C, prec, n0 * n1, n0 * n1, hm2, [n0, n1])
        a2 = D.DeviceArray.from_numpy(y)
        b2 = D.DeviceArray(y.nbytes)
        e2.execute(a2, b2, D.FFT_FORWARD)
        want = OP._fft_2d(y.astype(np.complex128), D.FFT_FORWARD)
        assert rel(b2.download(y.shape, cdt).astype(np.complex128), want) < tol
    finally:
        L.dtfft_hip_set_cache_dir(default_dir.encode())


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
@pytest.mark.parametrize("sched,tpl", [("16,8,4", 32), ("4,8,16", 64), ("2,16,16", 16)])
def test_tuning_records_select_radix_schedule(tmp_path, prec, sched, tpl):
    """A tuning record's radix schedule (the autotuner's sixth field) is honoured: a 512-point
    row stage run as 16 x 8 x 4, 4 x 8 x 16 or 2 x 16 x 16 must match the oracle both ways."""
    L = D.lib()
    c = "d" if prec == D.Precision.DOUBLE else "f"
    n, hm = 512, 96
    key = f"fft|{n}|{c}|000|{hm}|1|{n},0,1|{n},0,1"
    (tmp_path / "tuning.txt").write_text(f"{key}\t{tpl} 2 0 0 0 {sched}\n")
    default_dir = os.environ.get("DTFFT_HIP_CACHE_DIR") or os.path.join(os.path.dirname(D._lib.LIB_PATH), "kcache")
    assert L.dtfft_hip_set_cache_dir(str(tmp_path).encode()) == 0
    try:
        rng = np.random.default_rng(11)
        cdt = np.complex128 if prec == D.Precision.DOUBLE else np.complex64
        tol = 1e-12 if prec == D.Precision.DOUBLE else 1e-6
        x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(cdt)
        ex = D.FFTExecutor(1, D.FFTType.C2C, prec, n, n, hm, [n])
        a = D.DeviceArray.from_numpy(x)
        b = D.DeviceArray(x.nbytes)
        for sign in (D.FFT_FORWARD, D.FFT_BACKWARD):
            ex.execute(a, b, sign)
            got = b.download(x.shape, cdt).astype(np.complex128)
            assert rel(got, O.fft_c2c(x.astype(np.complex128), sign)) < tol
    finally:
        L.dtfft_hip_set_cache_dir(default_dir.encode())


@pytest.mark.parametrize("n", [384, 320, 48])
def test_exhaustive_stage_with_schedule_candidates_is_correct(tmp_path, n):
    """At EXHAUSTIVE the autotuner also times other radix schedules (16 x 8 x 3, 3 x 16 x 8,
    ...); whichever wins must match the oracle, and its record names a schedule of n."""
    L = D.lib()
    default_dir = os.environ.get("DTFFT_HIP_CACHE_DIR") or os.path.join(os.path.dirname(D._lib.LIB_PATH), "kcache")
    assert L.dtfft_hip_set_cache_dir(str(tmp_path).encode()) == 0
    try:
        hm = 256
        rng = np.random.default_rng(n)
        x = (rng.uniform(-1, 1, (hm, n)) + 1j * rng.uniform(-1, 1, (hm, n))).astype(np.complex64)
        ex = D.FFTExecutor(1, D.FFTType.C2C, D.Precision.SINGLE, n, n, hm, [n], effort=D.Effort.EXHAUSTIVE)
        a = D.DeviceArray.from_numpy(x)
        b = D.DeviceArray(x.nbytes)
        ex.execute(a, b, D.FFT_FORWARD)
        assert rel(b.download(x.shape, x.dtype).astype(np.complex128), O.fft_c2c(x.astype(np.complex128))) < 1e-6
        rec = [ln for ln in (tmp_path / "tuning.txt").read_text().splitlines() if ln.startswith(f"fft|{n}|")]
        assert rec
        rads = [int(r) for r in rec[0].split("\t")[1].split()[5].split(",")]
        assert int(np.prod(rads)) == n
    finally:
        L.dtfft_hip_set_cache_dir(default_dir.encode())
