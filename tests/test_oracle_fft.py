"""Pins the FP64 FFT oracle (FFTW conventions) before it is trusted.

Checks, each chosen so that a plausible mistake fails one of them:
* numpy/pocketfft golden vectors (tests/golden/fft_golden.npz) -- sign, order,
  unnormalised backward, R2C n/2+1 layout, C2R semantics;
* an O(n^2) long-double DFT for every length of the batched-uneven size set
  and the large lengths used by the plans;
* closed forms: delta -> ones, pure exponential -> n*delta, linearity, Parseval,
  bwd(fwd(x)) = n*x;
* multi-axis composition of the plan sequences against numpy.fft.fftn.
"""
import os

import numpy as np
import pytest

import oracle as O
import oracle.plan as OP

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
LENGTHS = sorted({int(k.split("_")[-1]) for k in GOLD.files if k.startswith("c2c_in_")})


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n", LENGTHS)
def test_c2c_golden(n):
    x = GOLD[f"c2c_in_{n}"]
    assert rel(O.fft_c2c(x, O.FFT_FORWARD), GOLD[f"c2c_fwd_{n}"]) < 1e-14
    assert rel(O.fft_c2c(x, O.FFT_BACKWARD), GOLD[f"c2c_bwd_{n}"]) < 1e-14


@pytest.mark.parametrize("n", LENGTHS)
def test_r2c_c2r_golden(n):
    r = GOLD[f"r2c_in_{n}"]
    assert rel(O.fft_r2c(r), GOLD[f"r2c_out_{n}"]) < 1e-14
    if n >= 2:
        assert rel(O.fft_c2r(GOLD[f"c2r_in_{n}"], n), GOLD[f"c2r_out_{n}"]) < 1e-14


@pytest.mark.parametrize("n", [48, 60, 64, 84, 96, 105, 128, 160, 192, 210, 256, 320, 384,
                               512, 1024, 2048, 11, 13, 121, 169, 37, 101, 1022, 4099])
def test_against_naive_dft(n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    # a prime radix p is an O(p^2) sum in the oracle: its rounding grows with sqrt(p)
    m = n
    for p in (2, 3, 5, 7, 11, 13):
        while m % p == 0:
            m //= p
    tol = 2e-15 if m == 1 else 2e-15 * max(1.0, np.sqrt(m) / 8)
    for sign in (O.FFT_FORWARD, O.FFT_BACKWARD):
        assert rel(O.fft_c2c(x, sign), O.dft_naive(x, sign)) < tol


@pytest.mark.parametrize("n", [1, 7, 64, 105, 512])
def test_closed_forms(n):
    d = np.zeros(n, complex)
    d[0] = 1
    assert np.allclose(O.fft_c2c(d), np.ones(n), atol=1e-15, rtol=0)
    k0 = 3 % n
    e = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    want = np.zeros(n, complex)
    want[k0] = n
    assert np.abs(O.fft_c2c(e) - want).max() < 1e-12 * n
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a, b = 0.3 - 2j, 1.7
    assert rel(O.fft_c2c(a * x + b * y), a * O.fft_c2c(x) + b * O.fft_c2c(y)) < 1e-14
    fx = O.fft_c2c(x)
    assert abs(np.vdot(fx, fx).real / n - np.vdot(x, x).real) < 1e-12 * np.vdot(x, x).real
    assert rel(O.fft_c2c(fx, O.FFT_BACKWARD) / n, x) < 1e-14


def test_c2r_ignores_dc_and_nyquist_imag():
    """FFTW c2r drops Im(X[0]) and Im(X[n/2]); the oracle must too."""
    n = 16
    rng = np.random.default_rng(5)
    h = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    h2 = h.copy()
    h2[0] = h2[0].real
    h2[-1] = h2[-1].real
    assert rel(O.fft_c2r(h, n), O.fft_c2r(h2, n)) < 1e-15


def test_batched_lines():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((7, 3, 60)) + 1j * rng.standard_normal((7, 3, 60))
    assert rel(O.fft_c2c(x), np.fft.fft(x, axis=-1)) < 1e-14


def test_fp32_twin_is_close():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((4, 512)) + 1j * rng.standard_normal((4, 512))).astype(np.complex64)
    assert rel(O.fft_c2c(x).astype(np.complex128), O.fft_c2c(x.astype(np.complex128))) < 1e-6


def test_plan_generic_3d_against_fftn():
    """execute_generic forward = fftn, laid out as the Z-pencil [Z, X, Y]."""
    x = GOLD["c2c3d_in"]  # (nz, ny, nx)
    z = OP.execute_generic(x, True)  # (ny, nx, nz)
    want = np.transpose(GOLD["c2c3d_fwd"], (1, 2, 0))
    assert rel(z, want) < 1e-14
    back = OP.execute_generic(z, False) / x.size
    assert rel(back, x) < 1e-14


def test_plan_z_slab_3d_against_fftn():
    x = GOLD["c2c3d_in"]
    z = OP.execute_z_slab(x, True)
    assert rel(z, np.transpose(GOLD["c2c3d_fwd"], (1, 2, 0))) < 1e-14
    assert rel(OP.execute_z_slab(z, False) / x.size, x) < 1e-14


def test_plan_2d_against_fft2():
    x = GOLD["c2c2d_in"]  # (ny, nx)
    y = OP.execute_2d(x, True)  # (nx, ny)
    assert rel(y, GOLD["c2c2d_fwd"].T) < 1e-14
    assert rel(OP.execute_2d(y, False) / x.size, x) < 1e-14


def test_plan_r2c_3d_against_rfftn():
    r = GOLD["r2c3d_in"]
    z = OP.execute_generic_r2c(r)
    assert rel(z, np.transpose(GOLD["r2c3d_out"], (1, 2, 0))) < 1e-14
    back = OP.execute_generic_c2r(z, r.shape[-1]) / r.size
    assert rel(back, r) < 1e-14


@pytest.mark.parametrize("shape", [(10, 16), (21, 60), (8, 4)])
def test_plan_r2c_2d_against_rfft2(shape):
    """2D R2C on one rank ends in the Y-pencil (nx/2+1, ny): the transpose of numpy's
    rfft2 over (y, x); the unnormalised C2R inverse returns N x."""
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    z = OP.execute_2d_r2c(x)
    assert z.shape == (shape[1] // 2 + 1, shape[0])
    assert rel(z, np.fft.rfft2(x).T) < 1e-14
    back = OP.execute_2d_c2r(z, shape[1]) / x.size
    assert rel(back, x) < 1e-14


@pytest.mark.parametrize("n", [1, 3, 7, 35, 74, 1023])
def test_real_odd_and_prime_lengths(n):
    """R2C / C2R of odd lengths and of halves with a prime factor above 13, against the
    long-double DFT of the real line and of the Hermitian extension."""
    rng = np.random.default_rng(n + 3)
    r = rng.standard_normal(n)
    full = O.dft_naive(r.astype(np.complex128))
    assert rel(O.fft_r2c(r), full[: n // 2 + 1]) < 2e-15
    h = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    h[0] = h[0].real
    if n % 2 == 0:
        h[-1] = h[-1].real
    ext = np.concatenate([h, np.conj(h[1:(n + 1) // 2][::-1])])
    assert ext.size == n
    want = O.dft_naive(ext, O.FFT_BACKWARD)
    assert np.abs(want.imag).max() < 1e-12 * max(1.0, np.abs(want).max())
    assert rel(O.fft_c2r(h, n), want.real) < 2e-15
