"""GPU parity at the benchmark workloads' full sizes, against the FP64 oracle.

* 512^3 C2C FP32 (the headline metric), fused and plain: forward output within 1e-6
  relative L2 of the FP64 oracle (src/dtfft_plan.F90:1057-1101 restated), and the round trip;
* R2C 1024 x 512 x 256, FP32 and its FP64 twin: forward within 1e-6 / 1e-12 of the FP64
  oracle, round trip within 1e-6 / 1e-13 -- at this size the default plan runs the
  LDS-twiddle column-in pass (513 lines per row) and the LDS-twiddle fused Z->X C2R;
* one-GPU share of a 4 x 4 pencil decomposition of 2048^3 complex64 (local pencil
  2048 x 512 x 512, 4 GiB): the per-peer pack bit-exact against the host kernels, the
  unpack bit-exact, and the 2048-point FFT (the LDS-twiddle rows, fused with the unpack or
  after it) within 1e-6 of the FP64 oracle on a seeded subsample of 4096 lines;
* 64 uneven grids of the batched cycle: every grid's forward output against the oracle.

Each test prints its case before the GPU work so that a failure names it."""
import numpy as np
import pytest

import oracle as O
import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def uniform_c64(n, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(n, dtype=np.complex64)
    x.real = rng.uniform(-1, 1, n).astype(np.float32)
    x.imag = rng.uniform(-1, 1, n).astype(np.float32)
    return x


@pytest.mark.parametrize("fused", [True, False])
def test_metric_512_fp32_forward_against_fp64_oracle(fused):
    n = 512
    print(f"case: 512^3 C2C FP32 fused={fused}", flush=True)
    x = uniform_c64(n ** 3, 20250512)
    D.set_config(enable_fused=fused)
    try:
        plan = D.PlanC2C((n, n, n), P.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    c = D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    # the plain backward (execute_generic) overwrites its input, as the reference's does
    got = b.download((n, n, n), np.complex64)
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download(x.size, np.complex64)
    rt = rel((back / np.float32(n ** 3)).astype(np.complex128), x.astype(np.complex128))
    del back, a, b, c
    want = OP.execute_generic(x.reshape(n, n, n).astype(np.complex128), True)
    err = rel(got.astype(np.complex128), want)
    print(f"forward rel L2 {err:.3e}, round trip {rt:.3e}", flush=True)
    assert err < 1e-6, err
    assert rt < 1e-6, rt


@pytest.mark.parametrize("prec", [P.SINGLE, P.DOUBLE])
def test_config3_r2c_1024x512x256(prec):
    nx, ny, nz = 1024, 512, 256
    nc = nx // 2 + 1
    print(f"case: R2C 1024x512x256 {prec.name} fused", flush=True)
    rd, cd = (np.float32, np.complex64) if prec == P.SINGLE else (np.float64, np.complex128)
    x = np.random.default_rng(3).standard_normal((nz, ny, nx)).astype(rd)
    plan = D.PlanR2C((nx, ny, nz), prec, D.Effort.ESTIMATE, D.Executor.HIP)
    a = D.DeviceArray.from_numpy(x, plan.alloc_bytes)
    b = D.DeviceArray(plan.alloc_bytes)
    c = D.DeviceArray(plan.alloc_bytes)
    plan.execute(a, b, D.Execute.FORWARD)
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), rd).astype(np.float64) / x.size
    rt = rel(back, x.astype(np.float64))
    del back
    got = b.download((ny, nc, nz), cd).astype(np.complex128)
    want = OP.execute_generic_r2c(x.astype(np.float64))
    err = rel(got, want)
    print(f"forward rel L2 {err:.3e}, round trip {rt:.3e}", flush=True)
    assert err < (1e-6 if prec == P.SINGLE else 1e-12), err
    assert rt < (1e-6 if prec == P.SINGLE else 1e-13), rt


@pytest.mark.parametrize("fused", [True, False])
def test_config5_pencil_share_2048x512x512(fused):
    Pn, me = 4, 1
    a = b = 2048
    c = 512
    print(f"case: pencil share 2048x512x512 c64 P=4 fused={fused}", flush=True)
    sh = D.ExchangeShare(a, b, c, Pn, me, P.SINGLE, fused=fused)
    n_a = a * (b // Pn) * c
    A = uniform_c64(n_a, 5)
    raw = A.view(O.elem_dtype(8))
    dA = D.DeviceArray.from_numpy(A)
    send, recv = D.DeviceArray(sh.bytes), D.DeviceArray(sh.bytes)
    n_b = int(np.prod(sh.b_dims))
    B = D.DeviceArray(n_b * 8)
    tmp = None if fused else D.DeviceArray(n_b * 8)
    sh.run(dA, send, recv, B, D.FFT_FORWARD, tmp=tmp)
    D.synchronize()
    del dA
    want_send = np.zeros(sh.bytes // 8, dtype=raw.dtype)
    for i in range(1, Pn + 1):
        O.host_kernel(O.KERNEL_PACK_FORWARD, sh.a_dims, raw, want_send, sh.send_rows, i)
    assert send.download(want_send.size, want_send.dtype).tobytes() == want_send.tobytes()
    del send, recv
    want_B = np.zeros(n_b, dtype=raw.dtype)
    O.host_kernel(O.KERNEL_UNPACK, sh.b_dims, want_send, want_B, sh.recv_rows)
    del want_send
    if tmp is not None:
        assert tmp.download(n_b, want_B.dtype).tobytes() == want_B.tobytes()
        del tmp
    lines = np.sort(np.random.default_rng(55).choice(n_b // b, 4096, replace=False))
    got = B.download(n_b, np.complex64).reshape(-1, b)[lines].astype(np.complex128)
    want = O.fft_c2c(want_B.view(np.complex64).reshape(-1, b)[lines].astype(np.complex128))
    err = rel(got, want)
    print(f"FFT rel L2 on 4096 lines {err:.3e}", flush=True)
    assert err < 1e-6, err


def test_config4_64_uneven_grids_forward_every_grid():
    sizes = [48, 60, 64, 84, 96, 105, 128, 160, 192, 210, 256, 320, 384]
    rng = np.random.default_rng(20250104)
    grids = [list(map(int, rng.choice(sizes, 3))) for _ in range(64)]
    print("case: 64 uneven grids forward, every grid against the oracle", flush=True)
    cy = D.Cycle(grids, P.SINGLE)
    offs, total = cy.layout()
    buf = np.zeros(total, dtype=np.complex64)
    buf.view(np.float32)[:] = rng.uniform(-1, 1, 2 * total).astype(np.float32)
    a = D.DeviceArray.from_numpy(buf)
    b, aux = D.DeviceArray(total * 8), D.DeviceArray(total * 8)
    cy.execute(a, b, aux, D.Execute.FORWARD)
    out = b.download(total, np.complex64)
    worst = 0.0
    for g, o in zip(grids, offs):
        nx, ny, nz = g
        x = buf[o:o + nx * ny * nz].reshape(nz, ny, nx).astype(np.complex128)
        z = out[o:o + x.size].reshape(ny, nx, nz).astype(np.complex128)
        worst = max(worst, rel(z, OP.execute_generic(x, True)))
    print(f"worst forward rel L2 {worst:.3e}", flush=True)
    assert worst < 1e-6, worst


@pytest.mark.parametrize("dims", [(512, 64, 32), (64, 512, 32), (32, 64, 512)])
def test_default_512_point_column_kernel(dims):
    """The metric's dominant kernel (single-buffer LDS-DMA column-in / row-out pass of 512
    points, 16 lines of 8 bytes per 128-byte tile row) at a small shape, against the oracle."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + ny + nz)
    x = (rng.uniform(-1, 1, (nz, ny, nx)) + 1j * rng.uniform(-1, 1, (nz, ny, nx))).astype(np.complex64)
    plan = D.PlanC2C(dims, P.SINGLE, D.Effort.ESTIMATE, D.Executor.HIP)
    a = D.DeviceArray.from_numpy(x)
    b, c = D.DeviceArray(x.nbytes), D.DeviceArray(x.nbytes)
    plan.execute(a, b, D.Execute.FORWARD)
    z = b.download((ny, nx, nz), np.complex64).astype(np.complex128)
    assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < 1e-6
    plan.execute(b, c, D.Execute.BACKWARD)
    back = c.download((nz, ny, nx), np.complex64).astype(np.complex128)
    assert rel(back, OP.execute_generic(z, False)) < 1e-6
