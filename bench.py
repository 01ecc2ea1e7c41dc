#!/usr/bin/env python3
"""Benchmark of the free-streaming + Landau-matching step on one MI355X.

Workload (the line's `value`): the 2-D lumpy event -- 261 x 261 cells,
DX = DY = 0.06 fm, 200 Gaussian hotspots, DIM_PHIP = 200, TAU0 = 0.1,
DTAU = 1.0 fm/c, conformal EoS, viscous matching -- with the events already
resident in HBM.  One step = one batch of `--batch` independent events of that
shape (different seeds) run with one launch per stage, replayed from a HIP
graph, on the mixed-precision path (2-D quadrature on FP16 MFMA with hi+lo
split, FP32/FP64 Landau).  `value` is events/s over the timed steps.

The line also carries: the same batch on the FP32 path; the plain path (one
launch sequence per event, no graph) on both paths; single-event latency
(p50/p99, device-resident); the 131^2 x 101 3-D event; a short serving run
(Poisson arrivals through the request loop, host inputs, results left in HBM)
with p50/p99 latency; the stream kernel's roofline position (HIP events on the
launch stream); and the reference-algorithm CPU baseline (the oracle's float
path, OpenMP on this host's cores).

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

FP32_PEAK_TFLOPS = 157.3
FP16_MFMA_PEAK_TFLOPS = 2500.0


def load_traffic():
    """HBM bytes per launch of the stream kernels, measured with rocprofv3
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size) and
    kept in profiles/traffic_stream2d_c1_batch64.txt as `kernel bytes` lines."""
    path = os.path.join(ROOT, "profiles", "traffic_stream2d_c1_batch64.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 2 and not line.startswith("#"):
                try:
                    out[parts[0]] = int(float(parts[1]))
                except ValueError:
                    pass
    return out


TRAFFIC = load_traffic()
FLOP_PER_SAMPLE_2D = 52  # 16 spline taps * 2 + 10 quadrature rows * 2
FLOP_PER_SAMPLE_3D = 36  # 8 trilinear taps * 2 + 10 * 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--latency-samples", type=int, default=200)
    ap.add_argument("--c2-steps", type=int, default=20)
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-rates", type=str, default="50,200,1000,3000,5000",
                    help="offered request rates (per second) of the serving sweep")
    ap.add_argument("--serve-outstanding", type=int, default=256,
                    help="requests kept outstanding in the closed-loop (capacity) serving run")
    ap.add_argument("--serve-pool", type=int, default=40, help="distinct requests in the serving pool")
    ap.add_argument("--serve-rapidity", choices=("gauss_hermite", "tan"), default="gauss_hermite",
                    help="rapidity rule of the pool's 3-D requests: the 21-point Gauss-Hermite rule, or tan = "
                         "the reference's own grid")
    ap.add_argument("--c3-events", type=int, default=512)
    ap.add_argument("--c3-steps", type=int, default=6)
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only launch what a counter profile needs: a 256 MiB fill and a 256 MiB elementwise "
                         "copy (the byte-counter calibration), then one eager C1 batch on each path")
    ap.add_argument("--profile-3d", action="store_true",
                    help="only launch what a 3-D counter profile needs: the byte-counter calibration kernels, then "
                         "the 131^2 x 101 event once on the FP32 path under the reference's rapidity grid and once "
                         "under the 21-point Gauss-Hermite rule")
    return ap.parse_args()


def profile_pass(args):
    """Short run for rocprofv3 --pmc: plain kernels that move a known number of
    bytes, then one batch of C1 events on the FP32 and on the mixed path
    (eager launches, so each stage is one dispatch)."""
    import torch
    import fsm_amd
    import fsm_amd.workloads as W

    n = 64 << 20  # floats: 256 MiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")  # fill kernel: writes 256 MiB
    b = torch.empty_like(a)
    for _ in range(2):
        torch.mul(a, 2.0, out=b)  # elementwise kernel: reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    ctx = fsm_amd.Context(0)
    p1, e1 = W.config("C1")
    q1 = fsm_amd.params_finalize(p1, e1)
    B = args.batch
    d_in = torch.from_numpy(np.stack([e1] + [W.config("C1", seed=100 + k)[1] for k in range(B - 1)])).cuda()
    d_out = torch.empty((B, 17, q1.DIM), dtype=torch.float64, device="cuda")
    for path in (fsm_amd.PATH_FP32, fsm_amd.PATH_MIXED):
        ctx.set_path(path)
        bt = fsm_amd.Batch(ctx)
        bt.set_graph(False)
        for k in range(B):
            bt.add(q1, d_in[k].data_ptr(), d_out[k].data_ptr())
        bt.run(torch.cuda.current_stream())
        torch.cuda.synchronize()
        bt.close()
    ctx.close()
    print(json.dumps({"profile_pass": True, "copy_bytes": 4 * n, "batch": B}), flush=True)


def profile_3d():
    """Short run for rocprofv3 --pmc on the 3-D kernels: the calibration
    kernels of profile_pass, then the 131^2 x 101 event (one-event launches,
    FP32 path) under each rapidity rule."""
    import torch
    import fsm_amd
    import fsm_amd.workloads as W

    n = 64 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    ctx = fsm_amd.Context(0)
    ctx.set_path(fsm_amd.PATH_FP32)
    p2, e2 = W.config("C2")
    q2 = fsm_amd.params_finalize(p2, e2)
    d_e = torch.from_numpy(e2).cuda()
    d_out = torch.empty(17 * q2.DIM, dtype=torch.float64, device="cuda")
    for rule in (fsm_amd.RAPIDITY_TAN, fsm_amd.RAPIDITY_GAUSS_HERMITE):
        ctx.set_rapidity_rule(rule)
        ctx.run_device(q2, d_e.data_ptr(), d_out.data_ptr(), torch.cuda.current_stream())
        torch.cuda.synchronize()
    ctx.close()
    print(json.dumps({"profile_3d": True, "copy_bytes": 4 * n}), flush=True)


def timed(fn, steps, sync):
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return time.perf_counter() - t0


def _req_class(p, nb):
    return ("3d" if p.DIM_ETA > 1 else "2d") + ("_baryon" if nb is not None else "")


def _pcts(v):
    v = np.array(v) if len(v) else np.zeros(1)
    return {"p50": round(float(np.percentile(v, 50)), 4), "p99": round(float(np.percentile(v, 99)), 4),
            "n": int(len(v))}


def serve_run(fsm_amd, W, ctx, seconds, rate, pool, dev_pool=None, seed=5, closed_loop=0):
    """Poisson arrivals at `rate` per second for `seconds` drawn from the C4
    request pool (2-D and 3-D events, some with n_B) through the server;
    latency = submit to results in HBM, also per request class.  With
    dev_pool (the pool's inputs already in HBM) requests are submitted with
    fsm_server_submit_device, otherwise from host memory.  closed_loop > 0:
    instead of Poisson arrivals, keep that many requests outstanding (a new one
    is submitted as each completes): the server's capacity."""
    rng = np.random.default_rng(seed)
    srv = fsm_amd.Server(ctx)  # defaults: 5 launch slots (three for 2-D work only), 1 ms of work per launch

    def submit(k):
        p, e, nb = pool[k]
        if dev_pool is None:
            return srv.submit(p, e, nb)
        de, dn = dev_pool[k]
        return srv.submit_device(p, de.data_ptr(), e.size, None if dn is None else dn.data_ptr())

    # every shape once outside the timed run: plans, tile tuning, memory pools
    for k in range(len(pool)):
        rid = submit(k)
        while srv.poll(100000) != rid:
            pass
        srv.release(rid)
    l0 = srv.stats()["launches"]
    lat, done, busy = [], 0, 0
    by_class = {}
    cls = {}
    t0 = time.perf_counter()
    t_next = t0
    submitted = []

    def one():
        k = int(rng.integers(len(pool)))
        rid = submit(k)
        if rid is not None:
            submitted.append(rid)
            cls[rid] = _req_class(pool[k][0], pool[k][2])
        return rid

    if closed_loop:
        for _ in range(closed_loop):
            one()
    while True:
        now = time.perf_counter()
        if not closed_loop and now - t0 < seconds and now >= t_next:
            if one() is None:
                busy += 1
            t_next += rng.exponential(1.0 / rate)
            continue
        rid = srv.poll(50)
        if rid is not None:
            st, l_us, _ = srv.result(rid)
            lat.append(l_us / 1e3)
            by_class.setdefault(cls.pop(rid), []).append(l_us / 1e3)
            srv.release(rid)
            done += 1
            if closed_loop and now - t0 < seconds:
                one()
        elif now - t0 >= seconds and done >= len(submitted):
            break
        if now - t0 > seconds + 30:
            break
    el = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    out = {"offered_per_s": rate if not closed_loop else None, "inputs": "host" if dev_pool is None else "device",
           "completed": done, "refused_busy": busy,
           "events_per_s": round(done / el, 2),
           "latency_ms": _pcts(lat),
           "latency_ms_by_class": {k: _pcts(v) for k, v in sorted(by_class.items())},
           "launches": stats["launches"] - l0}
    if closed_loop:
        out["closed_loop_outstanding"] = closed_loop
    return out


def plain_pool_run(fsm_amd, ctx, pool, dev_pool, stream, reps=2):
    """The plain per-event path (one launch sequence per event, no graph, no
    batching) on the context's path over every request of the pool, inputs
    in HBM: events/s, and the device time per request class."""
    import torch

    outs = [torch.empty((fsm_amd.NUM_OUTPUTS_BARYON if nb is not None else fsm_amd.NUM_OUTPUTS) * p.DIM,
                        dtype=torch.float64, device="cuda") for p, _, nb in pool]
    qs = [fsm_amd.params_finalize(p, e) for p, e, _ in pool]

    def run_all():
        for q, (de, dn), o, (_, _, nb) in zip(qs, dev_pool, outs, pool):
            if nb is None:
                ctx.run_device(q, de.data_ptr(), o.data_ptr(), stream)
            else:
                ctx.run_baryon_device(q, de.data_ptr(), dn.data_ptr(), o.data_ptr(), stream)

    run_all()
    torch.cuda.synchronize()
    t = timed(run_all, reps, torch.cuda.synchronize)
    per_class = {}
    for k, (q, (de, dn), o, (p, _, nb)) in enumerate(zip(qs, dev_pool, outs, pool)):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(2):
            if nb is None:
                ctx.run_device(q, de.data_ptr(), o.data_ptr(), stream)
            else:
                ctx.run_baryon_device(q, de.data_ptr(), dn.data_ptr(), o.data_ptr(), stream)
        torch.cuda.synchronize()
        per_class.setdefault(_req_class(p, nb), []).append((time.perf_counter() - a) * 1e3 / 2)
    del outs
    return {"events_per_s": round(len(pool) * reps / t, 2),
            "ms_per_event_by_class": {k: round(float(np.mean(v)), 4) for k, v in sorted(per_class.items())}}


def c3_run(fsm_amd, W, ctx, stream, steps, n_events, dev):
    """C3: n_events independent 2-D events of mixed shapes (each its own
    plan) as one batch, graph-replayed, on the mixed and on the FP32 path;
    and the same events through the plain per-event FP32 path."""
    import torch

    evs = W.batch_mix(n_events, seed=7)
    qs = [fsm_amd.params_finalize(p, e) for p, e in evs]
    d_in = [torch.from_numpy(e).to(dev) for _, e in evs]
    d_out = [torch.empty(17 * q.DIM, dtype=torch.float64, device=dev) for q in qs]
    sync = torch.cuda.synchronize
    res = {}
    for name, path in (("mixed", fsm_amd.PATH_MIXED), ("fp32", fsm_amd.PATH_FP32)):
        ctx.set_path(path)
        bt = fsm_amd.Batch(ctx)
        for q, a, o in zip(qs, d_in, d_out):
            bt.add(q, a.data_ptr(), o.data_ptr())
        for _ in range(2):
            bt.run(stream)
        res[name] = n_events * steps / timed(lambda: bt.run(stream), steps, sync)
        bt.close()

    def plain():
        for q, a, o in zip(qs, d_in, d_out):
            ctx.run_device(q, a.data_ptr(), o.data_ptr(), stream)

    plain()
    ps = max(1, steps // 2)
    res["plain_fp32"] = n_events * ps / timed(plain, ps, sync)
    cells = sum(q.DIM for q in qs)
    samples = sum(q.DIM * q.DIM_PHIP for q in qs)
    return {"workload": "%d 2-D events in one batch: side in {101..301}, DX 0.06-0.1, 20-300 hotspots, "
                        "DIM_PHIP {48,100,200} (after the 2 pi DTAU/DX rule: mean %.0f), DTAU 0.4-1.2; "
                        "device-resident, graph replay" % (n_events, samples / cells),
            "events": n_events, "cells": cells, "samples": samples,
            "mixed_batched_events_per_s": round(res["mixed"], 2),
            "fp32_batched_events_per_s": round(res["fp32"], 2),
            "plain_fp32_events_per_s": round(res["plain_fp32"], 2),
            "mixed_batched_over_plain_fp32": round(res["mixed"] / res["plain_fp32"], 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path here "
              "(events would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)
    if args.profile_3d:
        profile_3d()
        return
    if args.profile_pass:
        profile_pass(args)
        return
    import torch
    import fsm_amd
    import fsm_amd.workloads as W

    dev = torch.device("cuda:0")
    ctx = fsm_amd.Context(0)
    stream = torch.cuda.Stream()
    sync = torch.cuda.synchronize

    # ---- C1 batch, FP32 and mixed-precision paths ----------------------------
    B = args.batch
    p1, e1 = W.config("C1")
    q1 = fsm_amd.params_finalize(p1, e1)
    D = q1.DIM
    evs = [e1] + [W.config("C1", seed=100 + k)[1] for k in range(B - 1)]
    d_in = torch.from_numpy(np.stack(evs).astype(np.float32)).to(dev)
    d_out = torch.empty((B, 17, D), dtype=torch.float64, device=dev)

    def measure(path):
        """events/s of the graph-replayed batch and of the plain per-event
        path, plus per-stage kernel times (HIP events on the launch stream)."""
        ctx.set_path(path)
        batch = fsm_amd.Batch(ctx)
        for b in range(B):
            batch.add(q1, d_in[b].data_ptr(), d_out[b].data_ptr())

        def step_batch():
            batch.run(stream)

        def step_plain():
            for b in range(B):
                ctx.run_device(q1, d_in[b].data_ptr(), d_out[b].data_ptr(), stream)

        for _ in range(args.warmup):
            step_batch()
            step_plain()
        el = timed(step_batch, args.steps, sync)
        el_plain = timed(step_plain, max(1, args.steps // 2), sync)
        batch.set_graph(False)
        ctx.enable_timing(True)
        ctx.kernel_time_totals()
        timed(step_batch, args.steps, sync)
        ktot, ncalls = ctx.kernel_time_totals()
        ctx.enable_timing(False)
        batch.close()
        n = max(ncalls, 1)
        return {"events_per_s": B * args.steps / el, "ms_per_step": el * 1e3 / args.steps,
                "plain_events_per_s": B * max(1, args.steps // 2) / el_plain,
                "stream_launch_ms": ktot[1] / n,
                "stage_ms_per_event": {"spline": ktot[0] / n / B, "stream": ktot[1] / n / B,
                                       "landau": ktot[2] / n / B}}

    fp32 = measure(fsm_amd.PATH_FP32)
    out_fp32 = d_out[:4].cpu().numpy()
    mixed = measure(fsm_amd.PATH_MIXED)
    out_mixed = d_out[:4].cpu().numpy()
    # same events, both paths: largest e difference relative to max e
    mixed_vs_fp32_e = float((np.abs(out_mixed[:, 0] - out_fp32[:, 0]).max(axis=1) /
                             np.abs(out_fp32[:, 0]).max(axis=1)).max())
    value = mixed["events_per_s"]
    ms_per_step = mixed["ms_per_step"]
    info = ctx.plan_info(q1)  # context is on the mixed path here
    # the drop-in call on host buffers (fsm_run: input copied in, 17 double
    # outputs copied back each call), one event at a time: the PCIe-inclusive rate
    ctx.run(q1, e1)
    host_n = max(4, args.steps // 2)
    t_host = timed(lambda: ctx.run(q1, e1), host_n, sync)
    host_buffers = {"events_per_s": round(host_n / t_host, 2), "ms_per_event": round(t_host * 1e3 / host_n, 4),
                    "bytes_per_event": D * (4 + 17 * 8), "path": "mixed",
                    "note": "fsm_run on pageable host buffers, one C1 event per call (4 B in, 136 B out per cell)"}
    samples = D * q1.DIM_PHIP * q1.DIM_RAP * B
    k_ms = mixed["stream_launch_ms"]
    # algorithmic work: the reference's 52 FLOP per (cell, phi) sample (16
    # spline taps and 10 quadrature rows, 2 FLOP each) over every sample of
    # the launch, against the peak of the unit the kernel runs on (FP16 MFMA)
    alg_flops = FLOP_PER_SAMPLE_2D * samples
    achieved = alg_flops / (k_ms * 1e-3) / 1e12
    # what the MFMA pipe executes: v_mfma_f32_16x16x32_f16 instructions (the
    # plan mirrors the kernel's per-wave choice), 16*16*32*2 FLOP each
    mfma_flops = B * info["mfma_batch"] * 16384
    executed = mfma_flops / (k_ms * 1e-3) / 1e12
    roofline = {
        "kernel": "k_stream2d_mfma_batch",
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": FP16_MFMA_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP16_MFMA_PEAK_TFLOPS, 4),
        "traffic": TRAFFIC.get("k_stream2d_mfma_batch"),
        "avg_kernel_ms": round(k_ms, 5),
        "flop_per_launch": alg_flops,
        "samples_per_launch": samples,
        "samples_kept_per_launch": int(info["samples_kept"] * B),
        "achieved_vs_fp32_peak": round(achieved / FP32_PEAK_TFLOPS, 4),
        "executed_mfma_tflops": round(executed, 3),
        "executed_mfma_frac": round(executed / FP16_MFMA_PEAK_TFLOPS, 4),
        "executed_mfma_flop_per_launch": mfma_flops,
        "mfma_per_event": info["mfma_batch"],
        "algorithmic_hbm_bytes_per_launch": B * q1.DIM * (16 + 136),
        "note": "achieved = the reference's 52 FLOP per (cell, phi) sample x %d samples per launch (%d events) / "
                "the launch's average time (HIP events on the launch stream), against MI355X's dense FP16 MFMA "
                "peak; the launch also runs the fused 3x3 Landau stage in its tiles' epilogue and writes the 17 "
                "output planes (16 B of knots in, 136 B out per cell: algorithmic_hbm_bytes_per_launch); "
                "executed_mfma counts the v_mfma_f32_16x16x32_f16 instructions the kernel issues (%d per "
                "event, 16384 FLOP each: knots and weights both split hi+lo, all four products, 12 of 16 columns used), %.1fx the algorithmic "
                "work; traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/)"
                % (samples, B, info["mfma_batch"], mfma_flops / alg_flops),
    }
    kf = fp32["stream_launch_ms"]
    roofline_fp32 = {
        "kernel": "k_stream2d_batch", "bound": "valu", "traffic": TRAFFIC.get("k_stream2d_batch"),
        "achieved": round(FLOP_PER_SAMPLE_2D * samples / (kf * 1e-3) / 1e12, 3),
        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(FLOP_PER_SAMPLE_2D * samples / (kf * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
        "avg_kernel_ms": round(kf, 5),
    }

    # ---- single-event latency (device-resident) ------------------------------
    lat = []
    for i in range(args.latency_samples):
        sync()
        a = time.perf_counter()
        ctx.run_device(q1, d_in[i % B].data_ptr(), d_out[i % B].data_ptr(), stream)
        stream.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat)
    del d_out

    # ---- 3-D event, both paths ------------------------------------------------------
    c2 = None
    if not args.no_3d:
        p2, e2 = W.config("C2")
        q2 = fsm_amd.params_finalize(p2, e2)
        d2 = torch.from_numpy(e2).to(dev)
        o2 = torch.empty(17 * q2.DIM, dtype=torch.float64, device=dev)
        samples3 = q2.DIM * q2.DIM_RAP * q2.DIM_PHIP
        info3 = ctx.plan_info(q2)
        c2 = {"workload": "3-D 131x131x101, DIM_RAP=21, DIM_PHIP=48->51, DTAU=0.8, SIGMA=1; one event per call, "
                          "device-resident"}
        e_by_path = {}
        for path, name in ((fsm_amd.PATH_MIXED, "mixed"), (fsm_amd.PATH_FP32, "fp32")):
            ctx.set_path(path)
            ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream)
            sync()
            lat2 = []
            for _ in range(args.c2_steps):
                a = time.perf_counter()
                ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream)
                stream.synchronize()
                lat2.append((time.perf_counter() - a) * 1e3)
            ctx.enable_timing(True)
            ctx.kernel_time_totals()
            timed(lambda: ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream), args.c2_steps, sync)
            kt2, n2 = ctx.kernel_time_totals()
            ctx.enable_timing(False)
            e_by_path[name] = o2[:q2.DIM].cpu().numpy()
            lat2 = np.array(lat2)
            c2_ms = float(lat2.mean())
            c2[name] = {"events_per_s": round(1e3 / c2_ms, 2), "ms_per_event": round(c2_ms, 4),
                        "latency_ms": {"p50": round(float(np.percentile(lat2, 50)), 4),
                                       "p99": round(float(np.percentile(lat2, 99)), 4)},
                        "stage_ms": {"src": round(kt2[0] / n2, 4), "stream": round(kt2[1] / n2, 4),
                                     "landau": round(kt2[2] / n2, 4)},
                        # useful work: 36 FLOP per (cell, rapidity, phi) sample the masks keep
                        "stream_tflops_kept_samples":
                            round(FLOP_PER_SAMPLE_3D * info3["samples_kept"] / (kt2[1] / n2 * 1e-3) / 1e12, 3),
                        "stream_frac_of_fp32_peak_kept_samples":
                            round(FLOP_PER_SAMPLE_3D * info3["samples_kept"] / (kt2[1] / n2 * 1e-3) / 1e12 /
                                  FP32_PEAK_TFLOPS, 4),
                        "landau_hbm_tb_per_s":
                            round(q2.DIM * (40 + 136) / (kt2[2] / n2 * 1e-3) / 1e12, 3)}
        c2["mixed_over_fp32"] = round(c2["mixed"]["events_per_s"] / c2["fp32"]["events_per_s"], 3)
        c2["mixed_vs_fp32_max_e_diff_rel"] = float(np.abs(e_by_path["mixed"] - e_by_path["fp32"]).max() /
                                                   np.abs(e_by_path["fp32"]).max())
        c2["samples"] = {"nominal": samples3, "kept": int(info3["samples_kept"]),
                         "evaluated": int(info3["samples_evaluated"])}
        c2["note"] = ("3-D mixed path = the FP32 stream kernel + FP32/FP64 Landau (an FP16-MFMA form of the 3-D "
                      "quadrature measured slower and is not built); stream_tflops counts the reference's 36 FLOP "
                      "per sample the masks keep (kept of nominal DIM*DIM_RAP*DIM_PHIP; evaluated = what the "
                      "kernel computes); landau_hbm = 40 B in + 136 B out per cell over the Landau stage time")
        # the same event under the 21-point Gauss-Hermite rapidity rule (the reference's own rule, above, is
        # the tan grid), mixed path
        ctx.set_path(fsm_amd.PATH_MIXED)
        ctx.set_rapidity_rule(fsm_amd.RAPIDITY_GAUSS_HERMITE)
        info_gh = ctx.plan_info(q2)
        ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream)
        sync()
        lat_gh = []
        for _ in range(args.c2_steps):
            a = time.perf_counter()
            ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream)
            stream.synchronize()
            lat_gh.append((time.perf_counter() - a) * 1e3)
        ctx.enable_timing(True)
        ctx.kernel_time_totals()
        timed(lambda: ctx.run_device(q2, d2.data_ptr(), o2.data_ptr(), stream), args.c2_steps, sync)
        kt_gh, n_gh = ctx.kernel_time_totals()
        ctx.enable_timing(False)
        ctx.set_rapidity_rule(fsm_amd.RAPIDITY_TAN)
        lat_gh = np.array(lat_gh)
        c2["gauss_hermite_mixed"] = {
            "events_per_s": round(1e3 / float(lat_gh.mean()), 2), "ms_per_event": round(float(lat_gh.mean()), 4),
            "latency_ms": {"p50": round(float(np.percentile(lat_gh, 50)), 4),
                           "p99": round(float(np.percentile(lat_gh, 99)), 4)},
            "stage_ms": {"src": round(kt_gh[0] / n_gh, 4), "stream": round(kt_gh[1] / n_gh, 4),
                         "landau": round(kt_gh[2] / n_gh, 4)},
            "samples": {"kept": int(info_gh["samples_kept"]), "evaluated": int(info_gh["samples_evaluated"])},
            "note": "rapidity nodes eta + sqrt(2) SIGMA x_k, weights sqrt(2) SIGMA W_k of the 21-point "
                    "Gauss-Hermite rule (fsm_set_rapidity_rule); parity against the oracle under the same rule "
                    "in tests/test_gpu_gauss_hermite.py"}
        del d2, o2

    # ---- C3: a batch of 512 mixed 2-D events --------------------------------------
    c3 = None
    if not args.no_c3:
        ctx.set_path(fsm_amd.PATH_MIXED)
        c3 = c3_run(fsm_amd, W, ctx, stream, args.c3_steps, args.c3_events, dev)

    # ---- serving (C4 mix) ----------------------------------------------------------
    serve = None
    if not args.no_serve:
        ctx.set_path(fsm_amd.PATH_MIXED)
        ctx.set_rapidity_rule(fsm_amd.RAPIDITY_GAUSS_HERMITE if args.serve_rapidity == "gauss_hermite"
                              else fsm_amd.RAPIDITY_TAN)
        pool = W.serving_mix(args.serve_pool, seed=11)
        n3 = sum(1 for p, _, _ in pool if p.DIM_ETA > 1)
        nb = sum(1 for _, _, b in pool if b is not None)
        dev_pool = [(torch.from_numpy(e).to(dev), None if b is None else torch.from_numpy(b).to(dev))
                    for _, e, b in pool]
        sync()
        rates = [float(r) for r in args.serve_rates.split(",")]
        sweep = [serve_run(fsm_amd, W, ctx, args.serve_seconds, r, pool, dev_pool) for r in rates]
        capacity = serve_run(fsm_amd, W, ctx, args.serve_seconds, 0.0, pool, dev_pool,
                             closed_loop=args.serve_outstanding)
        host_point = serve_run(fsm_amd, W, ctx, args.serve_seconds, 1000.0, pool)
        # the plain FP32 path (one event per launch sequence, no batching) over the same pool
        ctx.set_path(fsm_amd.PATH_FP32)
        plain = plain_pool_run(fsm_amd, ctx, pool, dev_pool, stream)
        ctx.set_path(fsm_amd.PATH_MIXED)
        plain_mixed = plain_pool_run(fsm_amd, ctx, pool, dev_pool, stream)
        ctx.set_rapidity_rule(fsm_amd.RAPIDITY_TAN)
        del dev_pool
        serve = {"workload": "Poisson arrivals for %.1f s per rate from a pool of %d requests: %d 3-D (131^2 x "
                             "DIM_ETA 41-101, DIM_RAP 21, %s rapidity rule), the rest C3-style 2-D; %d carry n_B "
                             "(BARYON=1, SIGMA_B=0.8); profiles resident in HBM (submit_device), one rate also from "
                             "host memory; results kept in HBM; mixed path"
                             % (args.serve_seconds, len(pool), n3, args.serve_rapidity.replace("_", "-"), nb),
                 "rapidity_rule_3d": args.serve_rapidity,
                 "sweep": sweep, "capacity_closed_loop": capacity, "host_inputs": host_point,
                 "plain_fp32_same_pool": plain, "plain_mixed_same_pool": plain_mixed,
                 "serving_over_plain_fp32": round(capacity["events_per_s"] / plain["events_per_s"], 3),
                 "note": "serving_over_plain_fp32 = the server's closed-loop capacity on the mixed path over the "
                         "plain per-event FP32 path on the same pool; the pool's device time is dominated by its "
                         "3-D events, whose stream kernel is FP32 on both paths"}

    # ---- CPU baseline (oracle float path = the reference algorithm) --------------
    cpu = None
    if not args.no_cpu:
        import oracle as O

        qo = O.params_finalize(p1, e1)
        n = 0
        stage = np.zeros(4)
        a = time.perf_counter()
        while True:
            r = O.run(qo, e1, "float", want_T=False, want_extra=False)
            stage += r["stage_seconds"]
            n += 1
            if time.perf_counter() - a >= args.cpu_seconds:
                break
        cs = time.perf_counter() - a
        model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(n / cs, 4), "unit": "events/s", "cores": O.num_threads(), "kind": "port",
               "cpu_model": model, "host_cpus": os.cpu_count(),
               # per event, at the reference's own timer points (src/FreestreamMilne.cpp:375-393 freeStream,
               # :416-424 stress tensor, :475-482 eigen-solve; the output pack :677-718)
               "stage_seconds_per_event": {"convert_and_freestream": round(stage[0] / n, 5),
                                           "hypertrig_and_stress_tensor": round(stage[1] / n, 5),
                                           "eigen_eos_bulk_shear": round(stage[2] / n, 5),
                                           "output_pack": round(stage[3] / n, 5)},
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
        "dtype": "fp16 hi+lo MFMA, fp32 accumulate; fp64 Landau frame",
        "data": "synthetic",
        "config": {"workload": "2-D lumpy event 261x261, DX=0.06, 200 hotspots, DIM_PHIP=200, DTAU=1.0; "
                               "%d events per step, device-resident, batched launch replayed from a HIP graph, "
                               "mixed-precision path" % B,
                   "batch": B, "path": "mixed"},
        "fp32_batched_events_per_s": round(fp32["events_per_s"], 3),
        "plain_fp32_events_per_s": round(fp32["plain_events_per_s"], 3),
        "plain_mixed_events_per_s": round(mixed["plain_events_per_s"], 3),
        "mixed_batched_over_plain_fp32": round(value / fp32["plain_events_per_s"], 3),
        "mixed_vs_fp32_max_e_diff_rel": mixed_vs_fp32_e,
        "host_buffers_c1": host_buffers,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4), "path": "mixed"},
        "stage_ms_per_event": {"mixed": {k: round(v, 5) for k, v in mixed["stage_ms_per_event"].items()},
                               "fp32": {k: round(v, 5) for k, v in fp32["stage_ms_per_event"].items()}},
        "plan": info,
        "c2_3d": c2,
        "c3_batch512": c3,
        "serving": serve,
        "roofline": roofline,
        "roofline_fp32": roofline_fp32,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
