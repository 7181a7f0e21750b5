"""GPU check of dtfft_hip_execute_no_alloc (no counterpart in the reference: for callers that
cannot allocate), called through the C-ABI: its output must equal dtfft_hip_execute's bit for
bit, forward and backward, on two calls in a row with the second plan smaller than the first
(the scratch keeps the first call's contents), for the fused C2C, plain C2C, R2C and R2R
sequences; a plan whose aux exceeds the scratch is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
P = D.Precision


def no_alloc(plan, a, b, et):
    return D.lib().dtfft_hip_execute_no_alloc(plan.h, C.c_void_p(a.ptr), C.c_void_p(b.ptr), int(et))


def plan_of(kind, dims, prec, fused):
    D.set_config(enable_fused=fused)
    try:
        if kind == "c2c":
            return D.PlanC2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
        if kind == "r2c":
            return D.PlanR2C(list(dims), prec, D.Effort.ESTIMATE, D.Executor.HIP)
        kinds = [D.R2RKind.DCT_2] * len(dims)
        return D.PlanR2R(list(dims), kinds, prec, D.Effort.ESTIMATE, D.Executor.HIP)
    finally:
        D.set_config()


def one_call(kind, dims, prec, fused, seed):
    print(f"case: no-alloc {kind} {dims} {prec.name} fused={fused}", flush=True)
    plan = plan_of(kind, dims, prec, fused)
    nbytes = plan.alloc_bytes
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, nbytes // 4).astype(np.float32).view(np.uint8)
    a = D.DeviceArray.from_numpy(x)
    # outputs start equal (zero): an R2C backward writes only the real part of the buffer
    zero = np.zeros(nbytes, np.uint8)
    want_f, got_f, want_b, got_b = (D.DeviceArray.from_numpy(zero) for _ in range(4))
    # reference entry point (library-allocated aux)
    plan.execute(a, want_f, D.Execute.FORWARD)
    ref_f = want_f.download(nbytes, np.uint8)
    plan.execute(want_f, want_b, D.Execute.BACKWARD)
    # allocation-free entry point
    assert no_alloc(plan, a, got_f, D.Execute.FORWARD) == 0
    assert got_f.download(nbytes, np.uint8).tobytes() == ref_f.tobytes()
    assert no_alloc(plan, got_f, got_b, D.Execute.BACKWARD) == 0
    D.synchronize()
    assert got_b.download(nbytes, np.uint8).tobytes() == want_b.download(nbytes, np.uint8).tobytes()


@pytest.mark.parametrize("kind,big,small,prec,fused", [
    ("c2c", (64, 48, 40), (24, 20, 12), P.SINGLE, True),
    ("c2c", (64, 48, 40), (24, 20, 12), P.DOUBLE, False),
    ("r2c", (64, 32, 16), (16, 10, 6), P.SINGLE, True),
    ("r2r", (32, 24, 16), (12, 8, 6), P.DOUBLE, False),
])
def test_no_alloc_matches_execute_twice(kind, big, small, prec, fused):
    one_call(kind, big, prec, fused, 1)
    one_call(kind, small, prec, fused, 2)


def test_no_alloc_refuses_aux_beyond_scratch():
    # 512 x 512 x 160 complex64: 320 MiB of aux against 128 MiB of scratch (plan creation
    # allocates only twiddle tables; the refusal comes before any launch)
    plan = plan_of("c2c", (512, 512, 160), P.SINGLE, True)
    assert plan.aux_bytes > 128 << 20
    a = D.DeviceArray(16)
    rc = D.lib().dtfft_hip_execute_no_alloc(plan.h, C.c_void_p(a.ptr), C.c_void_p(a.ptr + 0),
                                            int(D.Execute.FORWARD))
    assert rc == 608
    assert b"no-alloc scratch" in D.lib().dtfft_hip_get_last_error_message()
