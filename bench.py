#!/usr/bin/env python3
"""Benchmark of the free-streaming + Landau-matching step on one MI355X.

Workload (default): the 2-D lumpy event -- 261 x 261 cells, DX = DY = 0.06 fm,
200 Gaussian hotspots, DIM_PHIP = 200, TAU0 = 0.1, DTAU = 1.0 fm/c, conformal
EoS, viscous matching -- streamed and Landau-matched on the GPU with the
event already resident in HBM.  One step = one batch of `--batch` independent
events of that shape (different seeds).  `value` is events/s over the timed
steps.  The line also carries single-event latency (p50/p99), the 131^2 x 101
3-D event rate, the dominant kernel's roofline position (HIP events on the
launch stream, summed over the timed region) and the reference-algorithm CPU
baseline (the oracle's float path, OpenMP on this host's cores).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "freestream-milne_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_SAMPLE_2D = 52  # 16 spline taps * 2 + 10 quadrature rows * 2
FLOP_PER_SAMPLE_3D = 36  # 8 trilinear taps * 2 + 10 * 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--latency-samples", type=int, default=200)
    ap.add_argument("--c2-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(events shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import fsm_amd
    import fsm_amd.workloads as W

    dev = torch.device("cuda:0")
    ctx = fsm_amd.Context(0)
    stream = torch.cuda.current_stream()

    # ---- C1 batch ---------------------------------------------------------
    B = args.batch
    p1, e1 = W.config("C1")
    q1 = fsm_amd.params_finalize(p1, e1)
    D = q1.DIM
    evs = [e1] + [W.config("C1", seed=100 + k)[1] for k in range(B - 1)]
    d_in = torch.from_numpy(np.stack(evs).astype(np.float32)).to(dev)
    d_out = torch.empty((B, 17, D), dtype=torch.float64, device=dev)

    def step():
        for b in range(B):
            ctx.run_device(q1, d_in[b].data_ptr(), d_out[b].data_ptr(), stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.enable_timing(True)
    ctx.kernel_time_totals()  # reset
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ktot, ncalls = ctx.kernel_time_totals()
    ctx.enable_timing(False)
    elapsed = t1 - t0
    events = B * args.steps
    value = events / elapsed
    ms_per_step = elapsed * 1e3 / args.steps

    # dominant kernel: stream + quadrature (stage index 1)
    k_ms = ktot[1] / max(ncalls, 1)
    samples = D * q1.DIM_PHIP * q1.DIM_RAP
    flops = FLOP_PER_SAMPLE_2D * samples
    achieved = flops / (k_ms * 1e-3) / 1e12
    roofline = {
        "kernel": "k_stream2d_one",
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
        "traffic": None,
        "avg_kernel_ms": round(k_ms, 5),
        "flop_per_launch": flops,
        "note": "FP32 VALU kernel: 52 algorithmic FLOP per (cell, phi) sample; 157.3 TFLOP/s is "
                "MI355X's FP32 peak",
    }
    stage_ms = {"spline": ktot[0] / max(ncalls, 1), "stream": k_ms, "landau": ktot[2] / max(ncalls, 1)}

    # ---- single-event latency ---------------------------------------------
    lat = []
    for i in range(args.latency_samples):
        torch.cuda.synchronize()
        a = time.perf_counter()
        ctx.run_device(q1, d_in[i % B].data_ptr(), d_out[i % B].data_ptr(), stream)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat)

    # ---- 3-D event ----------------------------------------------------------
    c2 = None
    if not args.no_3d:
        p2, e2 = W.config("C2")
        q2 = fsm_amd.params_finalize(p2, e2)
        d2 = torch.from_numpy(e2).to(dev)
        o2 = torch.empty(17 * q2.DIM, dtype=torch.float64, device=dev)
        ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream)
        torch.cuda.synchronize()
        ctx.enable_timing(True)
        ctx.kernel_time_totals()
        a = time.perf_counter()
        for _ in range(args.c2_steps):
            ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream)
        torch.cuda.synchronize()
        b = time.perf_counter()
        kt2, n2 = ctx.kernel_time_totals()
        ctx.enable_timing(False)
        c2_ms = (b - a) * 1e3 / args.c2_steps
        c2 = {"workload": "C2 3-D 131x131x101, DIM_RAP=21, DIM_PHIP=51, DTAU=0.8",
              "events_per_s": round(1e3 / c2_ms, 2), "ms_per_event": round(c2_ms, 4),
              "stage_ms": {"src": kt2[0] / n2, "stream": kt2[1] / n2, "landau": kt2[2] / n2}}
        del d2, o2

    # ---- CPU baseline (oracle float path = reference algorithm) ------------
    cpu = None
    if not args.no_cpu:
        import oracle as O

        qo = O.params_finalize(p1, e1)
        n = 0
        a = time.perf_counter()
        while True:
            O.run(qo, e1, "float", want_T=False, want_extra=False)
            n += 1
            if time.perf_counter() - a >= args.cpu_seconds:
                break
        cs = time.perf_counter() - a
        cpu = {"value": round(n / cs, 4), "unit": "events/s", "cores": O.num_threads(), "kind": "port",
               "sample": "%d C1 event(s) through the oracle's reference float path (unfused shiftedDensity, "
                         "10-pass quadrature, per-cell eigen-solve), OpenMP, %.1f s" % (n, cs)}

    line = {
        "metric": "events/s + p99 latency, 261^2 2D & 131^2x101 3D, mixed vs FP32 @1e-4 rel",
        "value": round(value, 3),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "C1 2-D lumpy event 261x261, DX=0.06, 200 hotspots, DIM_PHIP=200, DTAU=1.0; "
                               "batch of %d events per step, device-resident" % B,
                   "batch": B, "path": "fp32 per-event launches"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4)},
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "c2_3d": c2,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
