"""The C-ABI driven from Fortran through the iso_c_binding module
(dtfft_amd/fortran/dtfft_hip_api.F90): the built driver runs a permute (bit-exact
against its own host loop), an executor round trip and per-peer packs on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dtfft_amd", "lib",
                   "test_dtfft_hip")


def test_fortran_driver():
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (no flang on the build host)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
