"""CPU checks of the C-ABI library: it loads, exports every symbol include/dtfft_hip.h
declares, validates arguments the way the reference does, and compiles kernels with
hipRTC for gfx950 without a GPU (compile-only mode).  No device work runs here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dtfft_hip.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dtfft_hip_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def D():
    import dtfft_amd
    return dtfft_amd


def test_library_exports_every_declared_symbol(D):
    L = D.lib()
    names = header_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header(D):
    from dtfft_amd._lib import SIGNATURES
    assert set(header_functions()) == set(SIGNATURES)


def test_error_strings(D):
    L = D.lib()
    assert L.dtfft_hip_get_error_string(0) == b"DTFFT_SUCCESS"
    assert b"dimension" in L.dtfft_hip_get_error_string(4)
    assert L.dtfft_hip_get_error_string(101).startswith(b"R2R")


def _i32(v):
    return (C.c_int32 * len(v))(*v)


def test_argument_validation(D):
    L = D.lib()
    h = C.c_void_p()
    # plan: invalid ndims / dims / precision / executor / R2C transpose-only
    assert L.dtfft_hip_create_plan_c2c(4, _i32([2, 2, 2, 2]), 0, 0, 5, C.byref(h)) == 3
    assert L.dtfft_hip_create_plan_c2c(3, _i32([2, 0, 2]), 0, 0, 5, C.byref(h)) == 4
    assert L.dtfft_hip_create_plan_c2c(3, _i32([2, 2, 2]), 7, 0, 5, C.byref(h)) == 6
    assert L.dtfft_hip_create_plan_c2c(3, _i32([2, 2, 2]), 0, 9, 5, C.byref(h)) == 7
    assert L.dtfft_hip_create_plan_c2c(3, _i32([2, 2, 2]), 0, 0, 3, C.byref(h)) == 8
    assert L.dtfft_hip_create_plan_r2c(3, _i32([2, 2, 2]), 0, 0, 0, C.byref(h)) == 13
    # kernels: 2-step permutes are 3D only; pipelined kernels need neighbours
    assert L.dtfft_hip_kernel_create(9, 2, _i32([4, 4]), 8, None, 0, 0, 0, C.byref(h)) == 603
    assert L.dtfft_hip_kernel_create(6, 3, _i32([4, 4, 4]), 8, None, 0, 0, 0, C.byref(h)) == 607
    assert L.dtfft_hip_kernel_create(7, 3, _i32([4, 4, 4]), 12, None, 0, 0, 0, C.byref(h)) == 607
    # more than 2^31-1 elements per process
    assert L.dtfft_hip_kernel_create(7, 3, _i32([2048, 2048, 1024]), 8, None, 0, 0, 0, C.byref(h)) == 605
    # executor R2R: kinds missing (DTFFT_ERROR_MISSING_R2R_KINDS) or not DCT/DST 1-4 (INVALID_R2R_KINDS)
    assert L.dtfft_hip_fft_create(1, 2, 0, 8, 8, 1, _i32([8]), _i32([8]), _i32([8]), None, 0, None,
                                  C.byref(h)) == 11
    assert L.dtfft_hip_fft_create(1, 2, 0, 8, 8, 1, _i32([8]), _i32([8]), _i32([8]), _i32([11]), 0, None,
                                  C.byref(h)) == 12
    # plans: R2R with an executor needs kinds
    assert L.dtfft_hip_create_plan_r2r(3, _i32([4, 4, 4]), None, 0, 0, 5, C.byref(h)) == 11
    assert L.dtfft_hip_create_plan_r2r(3, _i32([4, 4, 4]), _i32([5, 5, 2]), 0, 0, 5, C.byref(h)) == 12
    # executor: a prime length whose Bluestein work line (>= 2n - 1 points) exceeds the LDS of
    # one workgroup (FP64: 10240 points); lengths with a prime factor > 13 are otherwise taken
    assert L.dtfft_hip_fft_create(1, 0, 1, 6007, 6007, 1, _i32([6007]), _i32([6007]), _i32([6007]), None, 0,
                                  None, C.byref(h)) == 602
    # DST-1 of 5200 needs a 10402-point FFT (2 x 7 x 743), beyond that in FP64 as well
    assert L.dtfft_hip_create_plan_r2r(2, _i32([4, 5200]), _i32([5, 7]), 1, 0, 5, C.byref(h)) == 602
    # destroying nothing
    assert L.dtfft_hip_destroy(C.byref(C.c_void_p())) == 1


def test_dummy_kernel_on_zero_dims(D):
    """Zero dims make a dummy kernel that does nothing (src/dtfft_abstract_kernel.F90:233-237)."""
    L = D.lib()
    h = C.c_void_p()
    assert L.dtfft_hip_kernel_create(7, 3, _i32([4, 0, 4]), 8, None, 0, 0, 0, C.byref(h)) == 0
    assert L.dtfft_hip_kernel_execute(h, None, None, None, 0, 0) == 0
    assert L.dtfft_hip_kernel_destroy(h) == 0


def test_compile_only_plan_and_kernels(D, tmp_path):
    """hipRTC compiles the kernels of a plan for gfx950 with no GPU present."""
    L = D.lib()
    if L.dtfft_hip_get_device_arch is None:  # pragma: no cover
        pytest.skip()
    arch = D.device_arch()
    if arch.startswith("gfx950"):
        pytest.skip("a GPU is present; compile-only behaviour is exercised on CPU hosts")
    p = D.PlanC2C([12, 10, 6], D.Precision.DOUBLE, D.Effort.ESTIMATE, D.Executor.HIP)
    ins, inc, outs, outc, alloc = p.local_sizes
    assert inc == [12, 10, 6]
    assert outc == [6, 12, 10]  # Z-pencil counts (nz, nx, ny), natural Fortran order
    assert alloc == 720
    assert p.element_size == 16
    r = D.PlanR2C([16, 10, 6], D.Precision.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)
    assert r.local_sizes[3] == [6, 9, 10]
    assert r.alloc_size == max(16 * 60, 2 * 9 * 60)
    assert r.element_size == 4
    # execute needs a device
    with pytest.raises(D.DtfftError) as e:
        p.execute(1, 2, D.Execute.FORWARD)
    assert e.value.code == 606


def test_force_effort_estimate_ignores_tuning_records(D, tmp_path, capfd, monkeypatch):
    """A tuning record for a permute is honoured at ESTIMATE, and ignored when the effort is
    forced (get_kernel, src/dtfft_kernel_device.F90:326-335): the default shape is taken."""
    L = D.lib()
    if D.device_arch().startswith("gfx950"):
        pytest.skip("a GPU is present; compile-only behaviour is exercised on CPU hosts")
    dims = [24, 20, 12]
    key = "mover|T|8|24x20x12|24,240"   # PERMUTE_FORWARD of 8-byte elements: is1 = nx, os0 = ny*nz
    (tmp_path / "tuning.txt").write_text(f"{key}\t16 4 0\n")
    default_dir = os.environ.get("DTFFT_HIP_CACHE_DIR") or os.path.join(os.path.dirname(D._lib.LIB_PATH), "kcache")
    monkeypatch.setenv("DTFFT_HIP_DEBUG", "1")
    assert L.dtfft_hip_set_cache_dir(str(tmp_path).encode()) == 0
    try:
        capfd.readouterr()
        D.Kernel(D.KernelType.PERMUTE_FORWARD, dims, 8)
        err = capfd.readouterr().err
        assert f"{key}: tile 16 rows 4 pad 0 vec 1 (record)" in err, err
        D.Kernel(D.KernelType.PERMUTE_FORWARD, dims, 8, force_effort=True)
        err = capfd.readouterr().err
        assert f"{key}: tile 64 rows 16 pad 1 vec 2 (default)" in err, err
    finally:
        L.dtfft_hip_set_cache_dir(default_dir.encode())


def test_compile_only_r2c_padded_pair_layout(D, capfd):
    """The fused 3D R2C forward builds the padded chunk-blocked pair where the padding is at
    most 1/16 of the lines (nc = 257 -> 272 here) and reports the larger buffers; where the
    padding would cost more (nc = 9 -> 16) or fusion is off, the natural layout and size stay."""
    L = D.lib()
    if L.dtfft_hip_get_device_arch is None:  # pragma: no cover
        pytest.skip()
    if D.device_arch().startswith("gfx950"):
        pytest.skip("a GPU is present; compile-only behaviour is exercised on CPU hosts")

    def plan(dims, fused):
        D.set_config(enable_fused=fused)
        try:
            p = D.PlanR2C(dims, D.Precision.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)
        finally:
            D.set_config()
        capfd.readouterr()
        p.report()
        return p, [ln.split()[0] for ln in capfd.readouterr().out.splitlines()[1:] if ln.strip()]

    nx, ny, nz = 512, 32, 16
    p, names = plan([nx, ny, nz], True)
    assert "fwd_rows" in names and "fwd_cols" in names, names
    assert p.alloc_size == 2 * 272 * ny * nz
    q, names = plan([nx, ny, nz], False)
    assert "fwd_rows" not in names
    assert q.alloc_size == max(nx * ny * nz, 2 * 257 * ny * nz)
    r, names = plan([16, 32, 16], True)
    assert "fwd_rows" not in names
    assert r.alloc_size == max(16 * 32 * 16, 2 * 9 * 32 * 16)


def test_compile_only_r2r_blocked_pairs(D, capfd):
    """The fused R2R sequence builds its chunk-blocked row -> column pairs where the row stage's
    tiles stay within one plane (128 x 64 x 64 FP32: 32-line forward row tiles over 64 lines
    per plane, 64-line backward ones over 128) and keeps the natural pair where they would
    straddle planes (64 x 32 x 96: 64-line row tiles over 32 lines per plane) or where
    `blocked` is off."""
    L = D.lib()
    if L.dtfft_hip_get_device_arch is None:  # pragma: no cover
        pytest.skip()
    if D.device_arch().startswith("gfx950"):
        pytest.skip("a GPU is present; compile-only behaviour is exercised on CPU hosts")

    def names(dims, blocked):
        D.set_config(enable_fused=True, blocked=blocked)
        try:
            p = D.PlanR2R(dims, [5] * len(dims), D.Precision.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)
        finally:
            D.set_config()
        capfd.readouterr()
        p.report()
        return [ln.split()[0] for ln in capfd.readouterr().out.splitlines()[1:] if ln.strip()]

    n = names([128, 64, 64], True)
    assert {"r2r_frow", "r2r_fcol", "r2r_brow", "r2r_bcol"} <= set(n), n
    assert not {"r2r_frow", "r2r_brow"} & set(names([128, 64, 64], False))
    assert not {"r2r_frow", "r2r_brow"} & set(names([64, 32, 96], True))
