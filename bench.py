#!/usr/bin/env python3
"""Benchmark of the single-GPU pencil-cycle hot path on one MI355X.

Workload: 512^3 C2C FP32 (complex64), one forward + one backward pass of the 3D
cycle per step, input i.i.d. uniform[-1, 1] (re and im) from a fixed seed, resident
in HBM before timing.  Two forms are timed:
  * fused   -- each transpose fused into the FFT pass that follows it (3 launches per
               direction); this is `value`;
  * plain   -- the reference's sequence, separate FFT and permute launches (5 per
               direction).
Throughput = algorithmic bytes / time, with 12 * N * 8 bytes per fwd+bwd cycle
(3 axis passes x read+write x 2 directions).  The roofline entry is the dominant
kernel (the fused transpose+FFT pass): 2 * N * 8 algorithmic bytes per launch over
its average duration, timed with HIP events on the stream it runs on.
The CPU baseline is the oracle's C restatement of dtFFT's host path (host permutes +
a mixed-radix FFT, OpenMP) on the same grid, timed on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--effort", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


class Timer:
    """HIP events on the default stream (the plans' stream)."""

    def __init__(self, D):
        self.D = D
        self.a, self.b = C.c_void_p(), C.c_void_p()
        D.call("dtfft_hip_event_create", C.byref(self.a))
        D.call("dtfft_hip_event_create", C.byref(self.b))

    def start(self):
        self.D.call("dtfft_hip_event_record", self.a, None)

    def stop_ms(self) -> float:
        self.D.call("dtfft_hip_event_record", self.b, None)
        ms = C.c_float()
        self.D.call("dtfft_hip_event_elapsed_ms", self.a, self.b, C.byref(ms))
        return float(ms.value)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU exchange is out of scope here)", flush=True)
        sys.exit(2)
    import dtfft_amd as D

    n = args.n
    N = n ** 3
    es = 8
    algo_cycle = 12 * N * es
    algo_pass = 2 * N * es
    log(f"[bench] device {D.device_arch()}  grid {n}^3 c64")

    rng = np.random.default_rng(20250512)
    x = np.empty(N, dtype=np.complex64)
    x.real = rng.uniform(-1, 1, N).astype(np.float32)
    x.imag = rng.uniform(-1, 1, N).astype(np.float32)

    def make(fused):
        D.set_config(enable_fused=fused)
        try:
            return D.PlanC2C([n, n, n], D.Precision.SINGLE, D.Effort(args.effort), D.Executor.HIP)
        finally:
            D.set_config()

    t0 = time.time()
    fused, plain = make(True), make(False)
    log(f"[bench] plans created in {time.time() - t0:.1f} s")
    a = D.DeviceArray.from_numpy(x)
    b = D.DeviceArray(x.nbytes)
    c = D.DeviceArray(x.nbytes)
    aux = D.DeviceArray(x.nbytes)
    timer = Timer(D)

    def cycle(plan):
        plan.execute(a, b, D.Execute.FORWARD, aux)
        plan.execute(b, c, D.Execute.BACKWARD, aux)

    def time_cycles(plan, steps, warmup):
        for _ in range(warmup):
            cycle(plan)
        D.synchronize()
        timer.start()
        for _ in range(steps):
            cycle(plan)
        ms = timer.stop_ms()
        D.synchronize()
        return ms / steps

    ms_fused = time_cycles(fused, args.steps, args.warmup)
    log(f"[bench] fused cycle {ms_fused:.3f} ms")
    back = c.download(N, np.complex64)
    rt_err = float(np.linalg.norm((back / np.float32(N)).astype(np.complex128) - x) / np.linalg.norm(x))
    del back
    ms_plain = time_cycles(plain, args.steps, args.warmup)
    log(f"[bench] plain cycle {ms_plain:.3f} ms")

    # per-launch durations of the fused sequence (dominant kernel for the roofline)
    stage_ms = {}
    for et, name in ((D.Execute.FORWARD, "fwd"), (D.Execute.BACKWARD, "bwd")):
        src, dst = (a, b) if et == D.Execute.FORWARD else (b, c)
        nst = fused.execute_stage(src, dst, et, -1, aux)
        for s in range(nst):
            for _ in range(2):
                fused.execute_stage(src, dst, et, s, aux)
            D.synchronize()
            timer.start()
            for _ in range(args.steps):
                fused.execute_stage(src, dst, et, s, aux)
            stage_ms[f"{name}{s}"] = timer.stop_ms() / args.steps
    # stage 0 of each direction is the contiguous FFT; stages 1-2 the fused transpose+FFT
    fused_keys = [k for k in stage_ms if not k.endswith("0")]
    dom_ms = float(np.mean([stage_ms[k] for k in fused_keys]))
    achieved = algo_pass / (dom_ms * 1e-3) / 1e9
    log(f"[bench] stages (ms): {json.dumps({k: round(v, 4) for k, v in stage_ms.items()})}")

    value = algo_cycle / (ms_fused * 1e-3) / 1e9
    plain_gbs = algo_cycle / (ms_plain * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, n, args.cpu_threads, algo_cycle)

    line = {
        "metric": "fwd+bwd cycle GB/s & %HBM peak, 512^3 C2C fp32; fused-batched vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_fused, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 butterflies, fp32 hi+lo twiddles)",
        "data": "synthetic uniform[-1,1] re/im, seed 20250512",
        "config": {"workload": f"3D C2C {n}x{n}x{n} complex64, one forward + one backward pass, "
                               "transposes fused into the following FFT pass",
                   "grid": [n, n, n], "algorithmic_bytes_per_cycle": algo_cycle},
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "cycles_per_s": round(1000.0 / ms_fused, 2),
        "plain": {"value": round(plain_gbs, 2), "unit": "GB/s", "ms_per_step": round(ms_plain, 4),
                  "hbm_frac": round(plain_gbs / HBM_PEAK_GBS, 4)},
        "fused_over_plain": round(ms_plain / ms_fused, 3),
        "roundtrip_rel_l2": rt_err,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "fused transpose+FFT pass (fft_c2c_kernel, column-in / row-out)",
                     "algorithmic_bytes_per_launch": algo_pass},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(x, n, threads, algo_cycle):
    """dtFFT's CPU path restated (oracle): host permutes + FFT, one fwd+bwd cycle of the
    same grid, FP32, OpenMP over this host's cores."""
    import oracle
    import oracle.plan as OP
    nthreads = threads or min(os.cpu_count() or 1, 16)
    oracle.set_threads(nthreads)
    xs = x.reshape(n, n, n)
    t0 = time.time()
    z = OP.execute_generic(xs, True)
    OP.execute_generic(z, False)
    dt = time.time() - t0
    log(f"[bench] cpu baseline cycle {dt:.2f} s on {nthreads} threads")
    return {"value": round(algo_cycle / dt / 1e9, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": f"one full fwd+bwd cycle of the {n}^3 complex64 grid (oracle C restatement of "
                      "kernel_host permutes + mixed-radix FFT, FP32, OpenMP)",
            "seconds": round(dt, 3)}


if __name__ == "__main__":
    main()
