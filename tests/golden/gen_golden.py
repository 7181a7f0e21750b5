"""Generate the FFT golden fixtures in this directory.

The vectors come from numpy.fft (pocketfft, FP64) and scipy.fft (DCT/DST,
FFTW REDFT/RODFT definitions) -- third-party FP64 implementations that are
independent of both the oracle and the HIP kernels.  The reference itself
stores no golden vectors (its tests are round trips and identities), so these
fixtures pin the FFT oracle's sign, ordering and normalisation conventions:
forward exp(-2*pi*i*jk/n), backward exp(+...), both unnormalised, R2C output
n/2+1, as FFTW does for dtFFT (src/interfaces/fft/fftw/dtfft_executor_fftw_m.F90).

Run:  python tests/golden/gen_golden.py   (writes fft_golden.npz)
"""
import os

import numpy as np
import scipy.fft as sfft

HERE = os.path.dirname(os.path.abspath(__file__))
LENGTHS = [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 15, 16, 17, 23, 25, 27, 32, 48, 49, 60, 64,
           84, 96, 105, 125, 128, 160, 192, 210, 243, 256, 320, 343, 384, 512, 1024, 2048]


def main():
    rng = np.random.default_rng(20250101)
    data = {}
    for n in LENGTHS:
        x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        data[f"c2c_in_{n}"] = x
        data[f"c2c_fwd_{n}"] = np.fft.fft(x)
        data[f"c2c_bwd_{n}"] = np.fft.ifft(x) * n  # unnormalised backward
        r = rng.standard_normal(n)
        data[f"r2c_in_{n}"] = r
        data[f"r2c_out_{n}"] = np.fft.rfft(r)
        if n >= 2:
            h = rng.uniform(-1, 1, n // 2 + 1) + 1j * rng.uniform(-1, 1, n // 2 + 1)
            data[f"c2r_in_{n}"] = h
            data[f"c2r_out_{n}"] = np.fft.irfft(h, n) * n  # unnormalised c2r
    # multi-axis: numpy C order (z, y, x), x fastest
    x3 = rng.uniform(-1, 1, (6, 10, 12)) + 1j * rng.uniform(-1, 1, (6, 10, 12))
    data["c2c3d_in"] = x3
    data["c2c3d_fwd"] = np.fft.fftn(x3)
    r3 = rng.standard_normal((6, 10, 12))
    data["r2c3d_in"] = r3
    data["r2c3d_out"] = np.fft.rfftn(r3)
    x2 = rng.uniform(-1, 1, (9, 14)) + 1j * rng.uniform(-1, 1, (9, 14))
    data["c2c2d_in"] = x2
    data["c2c2d_fwd"] = np.fft.fft2(x2)
    # R2R kinds: dtFFT kind id -> scipy (dct|dst, type); FFTW REDFT/RODFT unnormalised
    kinds = {3: ("dct", 1), 5: ("dct", 2), 4: ("dct", 3), 6: ("dct", 4),
             7: ("dst", 1), 9: ("dst", 2), 8: ("dst", 3), 10: ("dst", 4)}
    for n in [2, 3, 5, 8, 16, 17, 23, 64, 100]:
        r = rng.standard_normal(n)
        data[f"r2r_in_{n}"] = r
        for kid, (fam, t) in kinds.items():
            if fam == "dct" and t == 1 and n < 2:
                continue
            f = sfft.dct if fam == "dct" else sfft.dst
            data[f"r2r_{kid}_{n}"] = f(r, type=t, norm=None)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **data)


if __name__ == "__main__":
    main()
