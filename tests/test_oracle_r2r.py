"""Pins the R2R oracle (oracle/r2r.py): FFTW's REDFT/RODFT definitions against scipy's
golden vectors (generated in this repository, tests/golden/gen_golden.py), the inverse-pair
identities kind(inverse(kind) x) = N x with FFTW's logical sizes, and the permutation
structure of the 3D / 2D cycles (against separable numpy transposes)."""
import os

import numpy as np
import pytest

import oracle.r2r as R

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n", [2, 3, 5, 8, 16, 17, 23, 64, 100])
@pytest.mark.parametrize("kind", R.KINDS)
def test_r2r_golden(kind, n):
    x = GOLD[f"r2r_in_{n}"]
    assert rel(R.r2r(x, kind), GOLD[f"r2r_{kind}_{n}"]) < 1e-13


@pytest.mark.parametrize("n", [2, 3, 7, 16, 33])
@pytest.mark.parametrize("kind", R.KINDS)
def test_inverse_pairs(kind, n):
    rng = np.random.default_rng(n * 11 + kind)
    x = rng.standard_normal((4, n))
    back = R.r2r(R.r2r(x, kind), R.INVERSE[kind]) / R.logical_size(kind, n)
    assert rel(back, x) < 1e-13


def test_special_cases():
    """Closed forms: DCT-II of a constant is 2n at k = 0 only; DST-I of a unit impulse at
    j = 0 is 2 sin(pi (k+1) / (n+1)); DCT-IV is its own inverse up to 2n."""
    n = 12
    y = R.r2r(np.ones(n), R.DCT_2)
    assert abs(y[0] - 2 * n) < 1e-12 and np.max(np.abs(y[1:])) < 1e-12
    e = np.zeros(n); e[0] = 1
    assert np.max(np.abs(R.r2r(e, R.DST_1) - 2 * np.sin(np.pi * (np.arange(n) + 1) / (n + 1)))) < 1e-14
    x = np.random.default_rng(1).standard_normal(n)
    assert rel(R.r2r(R.r2r(x, R.DCT_4), R.DCT_4) / (2 * n), x) < 1e-13


@pytest.mark.parametrize("kinds", [(5, 6, 9), (3, 7, 10), (4, 8, 5)])
def test_generic_cycle(kinds):
    rng = np.random.default_rng(sum(kinds))
    nz, ny, nx = 5, 6, 7
    x = rng.standard_normal((nz, ny, nx))
    z = R.execute_generic_r2r(x, kinds, True)
    # separable: transform along each axis of the (z, y, x) array, then lay out as (y, x, z)
    want = R.r2r(x, kinds[0])
    want = np.moveaxis(R.r2r(np.moveaxis(want, 1, -1), kinds[1]), -1, 1)
    want = np.moveaxis(R.r2r(np.moveaxis(want, 0, -1), kinds[2]), -1, 0)
    assert rel(z, np.transpose(want, (1, 2, 0))) < 1e-13
    scale = np.prod([R.logical_size(k, n) for k, n in zip(kinds, (nx, ny, nz))])
    assert rel(R.execute_generic_r2r(z, kinds, False) / scale, x) < 1e-13


def test_2d_cycle():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 9))
    kinds = (R.DCT_2, R.DST_4)
    y = R.execute_2d_r2r(x, kinds, True)
    want = R.r2r(R.r2r(x, kinds[0]).T, kinds[1])
    assert rel(y, want) < 1e-14
    scale = R.logical_size(kinds[0], 9) * R.logical_size(kinds[1], 6)
    assert rel(R.execute_2d_r2r(y, kinds, False) / scale, x) < 1e-13
