"""GPU parity of the one-GPU share of a P > 1 forward transpose (pack per peer ->
loopback -> unpack [+ FFT along the aligned axis]) against the oracle.

Pack and unpack are bit-exact against the oracle's host kernels; the FFT after the
fused unpack is within 1e-6 (FP32) / 1e-12 (FP64) relative L2 of the FP64 oracle.
The 256^3-global share is the subsample on which the FFT error of the 2048^3 share
is stated."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_share(sh, A):
    """pack per peer -> loopback -> unpack on the oracle, plus the FP64 FFT along b."""
    es = sh.es
    raw = A.view(O.elem_dtype(es))
    send = np.zeros(sh.bytes // es, dtype=raw.dtype)
    for i in range(1, sh.P + 1):
        O.host_kernel(O.KERNEL_PACK_FORWARD, sh.a_dims, raw, send, sh.send_rows, i)
    recv = send.copy()
    B = np.zeros(int(np.prod(sh.b_dims)), dtype=raw.dtype)
    O.host_kernel(O.KERNEL_UNPACK, sh.b_dims, recv, B, sh.recv_rows)
    return send, B.view(A.dtype)


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("glob,P", [(64, 4), (96, 4), (256, 4)])
def test_exchange_share(glob, P, fused, prec):
    a = b = glob
    c = glob // P
    me = 1
    sh = D.ExchangeShare(a, b, c, P, me, prec, fused=fused)
    dt = np.complex64 if prec == D.Precision.SINGLE else np.complex128
    rng = np.random.default_rng(glob + P)
    n_a = a * (b // P) * c
    A = (rng.uniform(-1, 1, n_a) + 1j * rng.uniform(-1, 1, n_a)).astype(dt)
    want_send, want_B = oracle_share(sh, A)
    dA = D.DeviceArray.from_numpy(A)
    send = D.DeviceArray(sh.bytes)
    recv = D.DeviceArray(sh.bytes)
    n_b = int(np.prod(sh.b_dims))
    B = D.DeviceArray(n_b * sh.es)
    tmp = D.DeviceArray(n_b * sh.es)
    sh.run(dA, send, recv, B, D.FFT_FORWARD, tmp=tmp)
    D.synchronize()
    # pack into per-peer chunks: bit-exact
    assert send.download(want_send.size, want_send.dtype).tobytes() == want_send.tobytes()
    if not fused:
        assert tmp.download(n_b, dt).tobytes() == want_B.tobytes()  # unpack: bit-exact
    got = B.download(n_b, dt).astype(np.complex128).reshape(-1, b)
    want = O.fft_c2c(want_B.astype(np.complex128).reshape(-1, b))
    tol = 1e-6 if prec == D.Precision.SINGLE else 1e-12
    assert rel(got, want) < tol, rel(got, want)
