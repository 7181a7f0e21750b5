#!/usr/bin/env python3
"""Benchmark of the single-GPU pencil-cycle hot path on one MI355X.

Workload: 512^3 C2C FP32 (complex64), one forward + one backward pass of the 3D
cycle per step, input i.i.d. uniform[-1, 1] (re and im) from a fixed seed, resident
in HBM before timing.  Three forms are timed:
  * fused, graph -- each transpose fused into the FFT pass that follows it (3 launches per
                    direction), the fwd+bwd cycle captured once as a hipGraph and replayed;
                    this is `value`;
  * fused, eager -- the same launches issued one by one;
  * plain        -- the reference's sequence, separate FFT and permute launches (5 per
                    direction).
Throughput = algorithmic bytes / time, with 12 * N * 8 bytes per fwd+bwd cycle
(3 axis passes x read+write x 2 directions).  The roofline entry is the dominant
kernel (the fused transpose+FFT pass): 2 * N * 8 algorithmic bytes per launch over
its average duration, timed with HIP events on the stream it runs on.
The CPU leg times the oracle's C restatement of dtFFT's host path (host permutes with the
variant picked by timing, as dtFFT does at DTFFT_EXHAUSTIVE, + a mixed-radix FFT, OpenMP)
on the same grid on this host's cores, and -- as the checker -- states the GPU forward
output's relative L2 error against the FP64 oracle (`forward_rel_l2`).  `--config N` adds
the other workloads, each with its error against the FP64 oracle.

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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--effort", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--config", type=int, action="append", default=[],
                    help="also run another workload: 1 2D c128, 2 transpose c128, 3 R2C, 4 batched, 5 pencil share, "
                         "6 R2R, 7 the metric's FP64 twin")
    ap.add_argument("--only-configs", action="store_true", help="skip the headline metric")
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


class StepTimer:
    """One HIP event per step boundary on the default stream: the bracketed total over K
    steps and each step's own duration (for the median and the minimum)."""

    def __init__(self, D, steps):
        self.D = D
        self.ev = [C.c_void_p() for _ in range(steps + 1)]
        for e in self.ev:
            D.call("dtfft_hip_event_create", C.byref(e))

    def run(self, fn, steps, warmup):
        for _ in range(warmup):
            fn()
        self.D.synchronize()
        self.D.call("dtfft_hip_event_record", self.ev[0], None)
        for i in range(steps):
            fn()
            self.D.call("dtfft_hip_event_record", self.ev[i + 1], None)
        self.D.synchronize()
        per = []
        for i in range(steps):
            ms = C.c_float()
            self.D.call("dtfft_hip_event_elapsed_ms", self.ev[i], self.ev[i + 1], C.byref(ms))
            per.append(float(ms.value))
        tot = C.c_float()
        self.D.call("dtfft_hip_event_elapsed_ms", self.ev[0], self.ev[steps], C.byref(tot))
        return float(tot.value) / steps, float(np.median(per)), float(np.min(per))

    def close(self):
        for e in self.ev:
            self.D.call("dtfft_hip_event_destroy", e)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU exchange is out of scope here)", flush=True)
        sys.exit(2)
    import dtfft_amd as D

    if args.config:
        timer = Timer(D)
        for cfg in args.config:
            log(f"[bench] config {cfg}")
            {1: config1, 2: config2, 3: config3, 4: config4, 5: config5, 6: config6, 7: config7}[cfg](D, args, timer)
        if args.only_configs:
            return

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

    steps_timer = StepTimer(D, args.steps)

    def time_cycles(fn, steps, warmup):
        return steps_timer.run(fn, steps, warmup)

    # the fused cycle captured once as a hipGraph and replayed (the headline value), and the
    # same launches issued one by one
    ms_fused, med_fused, min_fused = time_cycles(lambda: fused.execute_cycle_graph(a, b, c, aux), args.steps,
                                                 args.warmup)
    log(f"[bench] fused cycle, graph replays {ms_fused:.3f} ms (median {med_fused:.3f}, min {min_fused:.3f})")
    back = c.download(N, np.complex64)
    rt_err = float(np.linalg.norm((back / np.float32(N)).astype(np.complex128) - x) / np.linalg.norm(x))
    del back
    fwd_out = b.download(N, np.complex64) if not args.no_cpu else None
    ms_eager, med_eager, min_eager = time_cycles(lambda: cycle(fused), args.steps, args.warmup)
    log(f"[bench] fused cycle, eager launches {ms_eager:.3f} ms (median {med_eager:.3f}, min {min_eager:.3f})")
    ms_plain, med_plain, min_plain = time_cycles(lambda: cycle(plain), args.steps, args.warmup)
    log(f"[bench] plain cycle {ms_plain:.3f} ms (median {med_plain:.3f}, min {min_plain:.3f})")

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
    # Stage 0 of each direction is the FFT along contiguous rows, written chunk-blocked.  fwd1
    # (X->Y) and bwd1 (Z->X) are the fused transpose+FFT passes that read those chunks as
    # contiguous column tiles and write rows: one kernel, the largest share of the cycle (the
    # dominant kernel).  fwd2 (Y->Z) reads gathered 128-byte column pieces; bwd2 is the FFT
    # along columns in place.
    dom_keys = ["fwd1", "bwd1"]
    dom_ms = float(np.mean([stage_ms[k] for k in dom_keys]))
    achieved = algo_pass / (dom_ms * 1e-3) / 1e9
    log(f"[bench] stages (ms): {json.dumps({k: round(v, 4) for k, v in stage_ms.items()})}")

    value = algo_cycle / (ms_fused * 1e-3) / 1e9
    traffic = _profiled_traffic() if n == 512 else None
    plain_gbs = algo_cycle / (ms_plain * 1e-3) / 1e9

    # measured device-copy bandwidth: the box_copy kernel moving the same 1 GiB buffer
    kc = D.Kernel(D.KernelType.PACK, [n, n, n], es, [[n, n, n, 0, 0]])
    ms_copy = _time(D, timer, lambda: kc.execute(a, c, neighbor=0), args.steps, 2)
    copy_gbs = algo_pass / (ms_copy * 1e-3) / 1e9
    log(f"[bench] copy kernel {ms_copy:.3f} ms = {copy_gbs:.0f} GB/s")

    cpu, fwd_err = None, None
    if not args.no_cpu:
        cpu, fwd_err = cpu_baseline(x, n, args.cpu_threads, algo_cycle, fwd_out)
        del fwd_out

    line = {
        "metric": "fwd+bwd cycle GB/s & %HBM peak, 512^3 C2C fp32; fused-batched vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_fused, 4),
        "ms_median": round(med_fused, 4),
        "ms_min": round(min_fused, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 butterflies, fp32 hi+lo twiddles)",
        "data": "synthetic uniform[-1,1] re/im, seed 20250512",
        "config": {"workload": f"3D C2C {n}x{n}x{n} complex64, one forward + one backward pass, "
                               "transposes fused into the following FFT pass, the cycle replayed as one "
                               "hipGraph",
                   "grid": [n, n, n], "algorithmic_bytes_per_cycle": algo_cycle},
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "cycles_per_s": round(1000.0 / ms_fused, 2),
        "copy_kernel_gbs": round(copy_gbs, 1),
        "frac_of_copy": round(value / copy_gbs, 4),
        "eager": {"value": round(algo_cycle / (ms_eager * 1e-3) / 1e9, 2), "unit": "GB/s",
                  "ms_per_step": round(ms_eager, 4), "ms_median": round(med_eager, 4), "ms_min": round(min_eager, 4),
                  "form": "the same fused launches issued one by one (6 per cycle)"},
        "plain": {"value": round(plain_gbs, 2), "unit": "GB/s", "ms_per_step": round(ms_plain, 4),
                  "ms_median": round(med_plain, 4), "ms_min": round(min_plain, 4),
                  "hbm_frac": round(plain_gbs / HBM_PEAK_GBS, 4),
                  "form": "the reference's sequence, separate FFT and permute launches (10 per cycle)"},
        "fused_over_plain": round(ms_plain / ms_fused, 3),
        "forward_rel_l2": fwd_err,
        "forward_rel_l2_against": "FP64 oracle (host permutes + FP64 Stockham FFT) on the same input",
        "roundtrip_rel_l2": rt_err,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": TRAFFIC_FILE if traffic else None,
                     "kernel": "fused transpose+FFT pass, chunk-blocked column tiles in / rows out "
                               "(fft_cols_dma_kernel: two LDS tile buffers filled by "
                               "global_load_lds_dwordx4, persistent blocks), stages fwd1 and bwd1",
                     "algorithmic_bytes_per_launch": algo_pass},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


TRAFFIC_FILE = "profiles/fft_512c64_pmc_traffic.md"


def _profiled_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 summary of the
    L2-to-memory request counters (each request weighed by its size); None when the summary is
    absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_FILE)
    try:
        for ln in open(path):
            if ln.startswith("traffic_bytes_per_launch"):
                return int(ln.split("=")[1])
    except OSError:
        return None
    return None


def _time(D, timer, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    D.synchronize()
    timer.start()
    for _ in range(steps):
        fn()
    ms = timer.stop_ms()
    D.synchronize()
    return ms / steps


def _line(metric, value, unit, ms, args, workload, **extra):
    d = {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
         "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "data": "synthetic", "config": {"workload": workload},
         "hbm_frac": round(value / HBM_PEAK_GBS, 4) if unit == "GB/s" else None}
    d.update(extra)
    print(json.dumps(d), flush=True)


def config1(D, args, timer):
    """2D 256x256 C2C complex128 on one rank (the reference's execute_2d order):
    FFT along X, X->Y permute, FFT along Y, and back; a launch-bound correctness case."""
    n0 = n1 = 256
    N = n0 * n1
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, N) + 1j * rng.uniform(-1, 1, N)).astype(np.complex128)
    algo = 2 * 2 * 2 * N * 16
    for fused in (True, False):
        D.set_config(enable_fused=fused)
        plan = D.PlanC2C([n0, n1], D.Precision.DOUBLE, D.Effort(args.effort), D.Executor.HIP)
        D.set_config()
        a = D.DeviceArray.from_numpy(x)
        b, c, aux = D.DeviceArray(N * 16), D.DeviceArray(N * 16), D.DeviceArray(N * 16)

        def cyc():
            plan.execute(a, b, D.Execute.FORWARD, aux)
            plan.execute(b, c, D.Execute.BACKWARD, aux)
        ms = _time(D, timer, cyc, args.steps, args.warmup)
        back = c.download(N, np.complex128) / N
        err = float(np.linalg.norm(back - x) / np.linalg.norm(x))
        _line(f"2D C2C fwd+bwd cycle GB/s, 256x256 c128 ({'fused' if fused else 'plain'})",
              algo / (ms * 1e-3) / 1e9, "GB/s", ms, args,
              "2D C2C 256x256 complex128, one rank, fwd+bwd", dtype="c128", algorithmic_bytes=algo,
              roundtrip_rel_l2=err, cycles_per_s=round(1000.0 / ms, 1))


def config7(D, args, timer):
    """3D C2C 512^3 complex128 (the FP64 twin of the metric): the fused cycle against the
    plain (reference) sequence, fwd+bwd, with the round trip bwd(fwd(x)) / N against x."""
    n = 512
    N = n ** 3
    rng = np.random.default_rng(7)
    x = np.empty(N, dtype=np.complex128)
    x.real = rng.uniform(-1, 1, N)
    x.imag = rng.uniform(-1, 1, N)
    a = D.DeviceArray.from_numpy(x)
    b, c, aux = D.DeviceArray(N * 16), D.DeviceArray(N * 16), D.DeviceArray(N * 16)
    algo = 12 * N * 16
    res = {}
    for fused in (True, False):
        D.set_config(enable_fused=fused)
        try:
            plan = D.PlanC2C([n, n, n], D.Precision.DOUBLE, D.Effort(args.effort), D.Executor.HIP)
        finally:
            D.set_config()

        def cyc():
            plan.execute(a, b, D.Execute.FORWARD, aux)
            plan.execute(b, c, D.Execute.BACKWARD, aux)
        res[fused] = _time(D, timer, cyc, max(3, args.steps // 2), 2)
        if fused:
            back = c.download(N, np.complex128)
            rt = _rel(back / N, x)
            del back
        del plan
    _line("fwd+bwd cycle GB/s, 512^3 C2C fp64", algo / (res[True] * 1e-3) / 1e9, "GB/s", res[True], args,
          "3D C2C 512x512x512 complex128, one forward + one backward pass, transposes fused into the "
          "following FFT pass", dtype="c128", algorithmic_bytes=algo,
          plain={"value": round(algo / (res[False] * 1e-3) / 1e9, 2), "ms_per_step": round(res[False], 4),
                 "form": "the reference's sequence, separate FFT and permute launches"},
          fused_over_plain=round(res[False] / res[True], 3), roundtrip_rel_l2=rt)


def config6(D, args, timer):
    """3D R2R 512^3 FP32: DCT-2 along every axis forward, DCT-3 (its inverse) backward.  The
    fused sequence folds each transpose into the R2R stage that follows it (3 launches per
    direction); the plain one is the reference's sequence, R2R stages with permutes between
    them (5 launches per direction).  Every R2R stage is one kernel: pre-process, half-length
    C2C and post-process through LDS."""
    n = 512
    N = n ** 3
    plans = {}
    for fused in (True, False):
        D.set_config(enable_fused=fused)
        try:
            plans[fused] = D.PlanR2R([n, n, n], [5, 5, 5], D.Precision.SINGLE, D.Effort(args.effort),
                                     D.Executor.HIP)
        finally:
            D.set_config()
    a, b, c, aux = (D.DeviceArray(N * 4) for _ in range(4))
    a.upload(np.random.default_rng(6).standard_normal(N).astype(np.float32))

    def cyc(plan):
        plan.execute(a, b, D.Execute.FORWARD, aux)
        plan.execute(b, c, D.Execute.BACKWARD, aux)
    ms = _time(D, timer, lambda: cyc(plans[True]), args.steps, args.warmup)
    ms_p = _time(D, timer, lambda: cyc(plans[False]), max(3, args.steps // 2), 2)
    algo = 12 * N * 4
    err = None
    if not args.no_cpu:
        plans[True].execute(a, b, D.Execute.FORWARD, aux)
        err = check_r2r(a.download((n, n, n), np.float32), b.download((n, n, n), np.float32))
    _line("R2R fwd+bwd cycle GB/s, 512^3 fp32 DCT-2/DCT-3", algo / (ms * 1e-3) / 1e9, "GB/s", ms, args,
          "3D R2R 512^3 float32, DCT-2 on X, Y, Z forward, DCT-3 backward, transposes fused into the "
          "R2R stages that follow them",
          dtype="f32", algorithmic_bytes=algo,
          plain={"value": round(algo / (ms_p * 1e-3) / 1e9, 2), "ms_per_step": round(ms_p, 4),
                 "form": "the reference's sequence: R2R stages with permutes between them"},
          fused_over_plain=round(ms_p / ms, 3), forward_rel_l2=err,
          forward_rel_l2_against="FP64 oracle (FFTW's REDFT10 as a dense matrix) along X, Y, Z")


def config2(D, args, timer):
    """512^3 complex128 transpose-only XYZ -> YZX -> ZXY -> XYZ (three permute launches)."""
    n = 512
    N = n ** 3
    a = D.DeviceArray(N * 16)
    b = D.DeviceArray(N * 16)
    ks = [D.Kernel(D.KernelType.PERMUTE_FORWARD, d, 16, effort=D.Effort(args.effort))
          for d in ((n, n, n), (n, n, n), (n, n, n))]

    def cyc():
        ks[0].execute(a, b)
        ks[1].execute(b, a)
        ks[2].execute(a, b)
    ms = _time(D, timer, cyc, args.steps, args.warmup)
    _line("transpose-only cycle GB/s, 512^3 c128", 6 * N * 16 / (ms * 1e-3) / 1e9, "GB/s", ms, args,
          "512^3 complex128 XYZ->YZX->ZXY->XYZ, 3 permute launches", dtype="c128 (exact moves)")


def config3(D, args, timer):
    """R2C 1024x512x256 (X fastest) fwd+bwd, FP32 and the FP64 twin; fused (value) and the
    reference's plain sequence."""
    nx, ny, nz = 1024, 512, 256
    nc = nx // 2 + 1
    errs = {}
    for prec, rs in ((D.Precision.SINGLE, 4), (D.Precision.DOUBLE, 8)):
        res = {}
        for fused in (True, False):
            D.set_config(enable_fused=fused)
            try:
                plan = D.PlanR2C([nx, ny, nz], prec, D.Effort(args.effort), D.Executor.HIP)
            finally:
                D.set_config()
            a = D.DeviceArray(plan.alloc_bytes)
            b = D.DeviceArray(plan.alloc_bytes)
            c = D.DeviceArray(plan.alloc_bytes)
            aux = D.DeviceArray(plan.alloc_bytes)
            rng = np.random.default_rng(3)
            xr = rng.standard_normal(nx * ny * nz).astype(np.float32 if rs == 4 else np.float64)
            a.upload(xr)

            def cyc():
                plan.execute(a, b, D.Execute.FORWARD, aux)
                plan.execute(b, c, D.Execute.BACKWARD, aux)
            res[fused] = _time(D, timer, cyc, args.steps, args.warmup)
            if fused and not args.no_cpu:
                plan.execute(a, b, D.Execute.FORWARD, aux)
                plan.execute(b, c, D.Execute.BACKWARD, aux)
                got = b.download((ny, nc, nz), np.complex64 if rs == 4 else np.complex128)
                back = c.download((nz, ny, nx), xr.dtype)
                errs[rs] = check_r2c(xr.reshape(nz, ny, nx), got, back)
            del a, b, c, aux, plan
        R = nx * ny * nz * rs
        Cb = nc * ny * nz * 2 * rs
        algo = 2 * (R + 5 * Cb)
        ms = res[True]
        _line(f"R2C fwd+bwd cycle GB/s, 1024x512x256 {'fp32' if rs == 4 else 'fp64'}", algo / (ms * 1e-3) / 1e9,
              "GB/s", ms, args, "3D R2C 1024x512x256: R2C X, X->Y+FFT_Y, Y->Z+FFT_Z; back Z, Y in place, "
              "Z->X fused with C2R X", dtype="f32" if rs == 4 else "f64", algorithmic_bytes=algo,
              plain={"value": round(algo / (res[False] * 1e-3) / 1e9, 2), "ms_per_step": round(res[False], 4),
                     "form": "reference sequence: R2C X, permute, FFT Y, permute, FFT Z"},
              fused_over_plain=round(res[False] / ms, 3),
              forward_rel_l2=errs.get(rs, (None, None))[0], roundtrip_rel_l2=errs.get(rs, (None, None))[1],
              forward_rel_l2_against="FP64 oracle R2C cycle on the same input")


def config4(D, args, timer):
    """64 uneven 3D C2C complex64 grids: one persistent launch per stage, fwd+bwd as a
    hipGraph, against 64 sequential plain per-grid plans."""
    sizes = [48, 60, 64, 84, 96, 105, 128, 160, 192, 210, 256, 320, 384]
    rng = np.random.default_rng(20250104)
    grids = [list(map(int, rng.choice(sizes, 3))) for _ in range(64)]
    # the autotuned cycle (tile shapes and radix schedules timed per length group, at least
    # MEASURE) is the value; the default shapes (ESTIMATE) are reported beside it
    t0 = time.time()
    cy0 = D.Cycle(grids, D.Precision.SINGLE, D.Effort.ESTIMATE)
    t1 = time.time()
    tuned_effort = D.Effort(max(int(args.effort), int(D.Effort.MEASURE)))
    cy = D.Cycle(grids, D.Precision.SINGLE, tuned_effort)
    t2 = time.time()
    log(f"[bench] cycle with 64 grids created in {t1 - t0:.1f} s (default), {t2 - t1:.1f} s ({tuned_effort.name})")
    offs, total = cy.layout()
    a, b, aux, c = (D.DeviceArray(total * 8) for _ in range(4))
    a.upload(rng.uniform(-1, 1, 2 * total).astype(np.float32))
    algo = sum(12 * g[0] * g[1] * g[2] * 8 for g in grids)
    ms_d = _time(D, timer, lambda: cy0.execute_graph(a, b, aux, c), args.steps, args.warmup)
    ms_b = _time(D, timer, lambda: cy.execute_graph(a, b, aux, c), args.steps, args.warmup)
    log(f"[bench] batched cycle {ms_b:.3f} ms tuned, {ms_d:.3f} ms default shapes")
    fwd_err = None
    if not args.no_cpu:
        fwd_err = check_grids(grids, offs, a.download(total, np.complex64), b.download(total, np.complex64))
    D.set_config(enable_fused=False)
    plans = [D.PlanC2C(g, D.Precision.SINGLE, D.Effort(args.effort), D.Executor.HIP) for g in grids]
    D.set_config()

    def plain():
        for p, o in zip(plans, offs):
            ob = o * 8
            p.execute(a.ptr + ob, b.ptr + ob, D.Execute.FORWARD, aux.ptr + ob)
            p.execute(b.ptr + ob, c.ptr + ob, D.Execute.BACKWARD, aux.ptr + ob)
    ms_p = _time(D, timer, plain, max(3, args.steps // 4), 2)
    _line("batched uneven fwd+bwd GB/s, 64 C2C complex64 grids", algo / (ms_b * 1e-3) / 1e9, "GB/s", ms_b, args,
          "64 grids, axes drawn (seed 20250104) from {48..384}, persistent launch per stage + hipGraph, "
          f"tile shapes and radix schedules from the autotuner ({tuned_effort.name})",
          dtype="c64", algorithmic_bytes=algo,
          default_shapes={"value": round(algo / (ms_d * 1e-3) / 1e9, 2), "ms_per_step": round(ms_d, 4),
                          "form": "the same cycle with the default tile shapes and schedules (ESTIMATE)"},
          tuning_seconds=round(t2 - t1, 1),
          plain={"value": round(algo / (ms_p * 1e-3) / 1e9, 2), "ms_per_step": round(ms_p, 4),
                 "form": "64 sequential plain plans (FFT + permute launches)"},
          batched_over_plain=round(ms_p / ms_b, 3), forward_rel_l2_worst_grid=fwd_err,
          forward_rel_l2_against="FP64 oracle generic cycle, every grid")


def config5(D, args, timer):
    """One-GPU share of a 4x4 pencil decomposition of 2048^3 complex64: Y-pencil
    [2048, 512, 512] -> pack into 4 per-peer chunks -> loopback -> unpack + FFT along Z."""
    P = 4
    a = b = 2048
    c = 512
    res = {}
    chk = (None, None)
    for fused in (True, False):
        sh = D.ExchangeShare(a, b, c, P, 1, D.Precision.SINGLE, fused=fused, effort=D.Effort(args.effort))
        S = sh.bytes
        A = D.DeviceArray(S)
        send = D.DeviceArray(S)
        recv = D.DeviceArray(S)
        B = D.DeviceArray(S)
        tmp = None if fused else D.DeviceArray(S)
        xa = np.random.default_rng(5).uniform(-1, 1, S // 4).astype(np.float32)
        A.upload(xa)
        res[fused] = _time(D, timer, lambda: sh.run(A, send, recv, B, D.FFT_FORWARD, tmp=tmp),
                           max(3, args.steps // 2), 2)
        if fused and not args.no_cpu:
            chk = check_share(sh, xa.view(np.complex64), send.download(S // 8, np.uint64),
                              B.download(S // 8, np.complex64))
        del A, send, recv, B, tmp, xa
    S = 2048 * 512 * 512 * 8
    _line("pencil-share Y->Z stage GB/s, 2048x512x512 c64 (4 peers)", 6 * S / (res[True] * 1e-3) / 1e9, "GB/s",
          res[True], args, "pack_forward x4 -> loopback copy -> unpack fused with FFT_Z (n=2048)",
          dtype="c64", algorithmic_bytes=6 * S,
          plain={"value": round(6 * S / (res[False] * 1e-3) / 1e9, 2), "ms_per_step": round(res[False], 4),
                 "form": "pack x4, loopback, unpack x4, separate FFT_Z"},
          pack_bit_exact=chk[0], forward_rel_l2=chk[1],
          forward_rel_l2_against="FP64 oracle FFT of 4096 seeded lines of the oracle's unpacked pencil")


def check_r2c(x, got, back):
    """CPU leg (checker) of the R2C workload: the GPU's forward output against the FP64
    oracle's R2C cycle on the same input, and the round trip."""
    import oracle.plan as OP
    want = OP.execute_generic_r2c(x.astype(np.float64))
    err = _rel(got.astype(np.complex128), want)
    del want
    rt = _rel(back.astype(np.float64) / x.size, x.astype(np.float64))
    log(f"[bench] R2C forward rel L2 {err:.3e}, round trip {rt:.3e}")
    return err, rt


def check_grids(grids, offs, inp, out):
    """CPU leg (checker) of the batched workload: every grid's forward output against the
    FP64 oracle's generic cycle.  Returns the worst relative L2."""
    import oracle.plan as OP
    worst = 0.0
    for g, o in zip(grids, offs):
        nx, ny, nz = g
        x = inp[o:o + nx * ny * nz].reshape(nz, ny, nx).astype(np.complex128)
        z = out[o:o + x.size].reshape(ny, nx, nz).astype(np.complex128)
        worst = max(worst, _rel(z, OP.execute_generic(x, True)))
    log(f"[bench] batched forward, worst grid rel L2 {worst:.3e}")
    return worst


def check_share(sh, A, send, B):
    """CPU leg (checker) of the pencil share: the per-peer pack against the host kernels
    (bit-exact), and the fused unpack+FFT on 4096 seeded lines against the FP64 oracle FFT
    of the oracle's unpacked pencil."""
    import oracle as O
    raw = A.view(np.uint64)
    want_send = np.zeros(send.size, dtype=np.uint64)
    for i in range(1, sh.P + 1):
        O.host_kernel(O.KERNEL_PACK_FORWARD, sh.a_dims, raw, want_send, sh.send_rows, i)
    exact = bool(np.array_equal(send, want_send))
    want_B = np.zeros(B.size, dtype=np.uint64)
    O.host_kernel(O.KERNEL_UNPACK, sh.b_dims, want_send, want_B, sh.recv_rows)
    del want_send
    b = sh.b_dims[0]
    lines = np.sort(np.random.default_rng(55).choice(B.size // b, 4096, replace=False))
    want = O.fft_c2c(want_B.view(np.complex64).reshape(-1, b)[lines].astype(np.complex128))
    err = _rel(B.reshape(-1, b)[lines].astype(np.complex128), want)
    log(f"[bench] share pack bit-exact {exact}, FFT rel L2 {err:.3e}")
    return exact, err


def check_r2r(x, got):
    """CPU leg (checker) of the R2R workload: DCT-2 along X, Y and Z by the FP64 oracle
    (FFTW's REDFT10 definition as a dense matrix) against the GPU's forward output."""
    import oracle.r2r as R
    want = R.execute_generic_r2r(x.astype(np.float64), (R.DCT_2, R.DCT_2, R.DCT_2), True)
    err = _rel(got.astype(np.float64), want)
    log(f"[bench] R2R forward rel L2 {err:.3e}")
    return err


def _host():
    """Host CPU description: logical CPUs, the CPUs this process may run on, the model."""
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "allowed": allowed, "model": model}


def _cpu_threads(threads):
    """Threads of the CPU baseline: --cpu-threads, else OMP_NUM_THREADS (this process's CPU
    share where the launcher sets it), else every CPU this process may run on."""
    if threads:
        return threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return _host()["allowed"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(float(np.linalg.norm(b)), 1e-300))


def cpu_baseline(x, n, threads, algo_cycle, gpu_forward=None):
    """CPU leg: dtFFT's CPU path restated (the oracle: host permutes, picked by the restated
    host-kernel autotune, + FFT), one fwd+bwd cycle of the same grid in FP32 over this host's
    cores -- one warm-up cycle, then the median of 3 -- and, as the checker, the FP64 oracle's
    forward of the same input against the GPU's forward output."""
    import oracle
    import oracle.plan as OP
    nthreads = _cpu_threads(threads)
    oracle.set_threads(nthreads)
    xs = x.reshape(n, n, n)
    tuned = OP.autotune_generic((n, n, n), 8)
    times = []
    for i in range(4):
        t0 = time.perf_counter()
        z = OP.execute_generic(xs, True, kernels=tuned)
        OP.execute_generic(z, False, kernels=tuned)
        dt = time.perf_counter() - t0
        del z
        if i > 0:
            times.append(dt)
    dt = float(np.median(times))
    host = _host()
    log(f"[bench] cpu baseline cycle {dt:.2f} s (median of {len(times)}) on {nthreads} threads, "
        f"{host['model']} ({host['allowed']} of {host['nproc']} CPUs); host kernels {tuned}")
    fwd_err = None
    if gpu_forward is not None:
        want = OP.execute_generic(xs.astype(np.complex128), True)
        fwd_err = _rel(gpu_forward.reshape(want.shape).astype(np.complex128), want)
        del want
        log(f"[bench] forward rel L2 against the FP64 oracle {fwd_err:.3e}")
    base = {"value": round(algo_cycle / dt / 1e9, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": f"full fwd+bwd cycles of the {n}^3 complex64 grid: oracle C restatement of dtFFT's "
                      "host path (kernel_host permutes with the BASE/BLOCK and WRITE/READ variant picked by "
                      "timing as at DTFFT_EXHAUSTIVE, + a mixed-radix FFT standing in for FFTW), FP32, OpenMP; "
                      "1 warm-up, median of 3",
            "seconds": round(dt, 3), "seconds_all": [round(t, 3) for t in times],
            "host_kernels": tuned, "host": host}
    return base, fwd_err


if __name__ == "__main__":
    main()
