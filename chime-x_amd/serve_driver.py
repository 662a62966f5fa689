"""Drives the streaming engine with timed frame arrivals (chx_serve).

Frames arrive by a Poisson process at a chosen rate; each is submitted at
its arrival time (or, when every frame slot is busy, as soon as the oldest
frame in flight completes and frees one).  A frame's latency is counted from
its arrival: the wait for a slot plus the engine's submit-to-done time, so
overload shows as queueing delay.

    python chime-x_amd/serve_driver.py            # streaming workload, JSON summary
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from collections import deque

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def _submit_blocking(serve, sid, ptr, device: bool, inflight: deque):
    """Submits, waiting for the oldest frame in flight while slots are full."""
    submit = serve.submit_device if device else serve.submit
    while True:
        fid = submit(sid, ptr)
        if fid is not None:
            inflight.append(fid)
            return fid
        if inflight:
            serve.wait_frame(inflight.popleft())
        else:  # other streams hold every slot
            time.sleep(1e-4)


def run_arrivals(serve, sid, frames, arrivals, device: bool) -> dict:
    """Submits frames[i] (addresses) at arrivals[i] seconds after the start.

    Returns the per-frame latency from arrival (ms), the offered and achieved
    frame rates and the engine's batch statistics."""
    serve.reset_stats()
    inflight: deque = deque()
    t_sub = {}
    t0 = time.perf_counter()
    for i, ptr in enumerate(frames):
        target = t0 + arrivals[i]
        now = time.perf_counter()
        if target > now:
            time.sleep(target - now)
        fid = _submit_blocking(serve, sid, ptr, device, inflight)
        t_sub[fid] = (time.perf_counter() - t0) - arrivals[i]  # wait for a slot
    while inflight:
        serve.wait_frame(inflight.popleft())
    t_end = time.perf_counter() - t0
    st = serve.stats()
    ids, ms = serve.latencies()
    lat = np.array([ms[k] + 1e3 * t_sub[int(ids[k])] for k in range(len(ids))])
    return dict(frames=len(frames), seconds=t_end,
                offered_rate=len(frames) / max(arrivals[-1], 1e-9) if len(frames) > 1 else 0.0,
                achieved_rate=len(frames) / t_end,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                max_ms=float(lat.max()), mean_ms=float(lat.mean()),
                engine_p50_ms=st["latency_p50_ms"], engine_p99_ms=st["latency_p99_ms"],
                batches=st["batches"], mean_batch_frames=st["mean_batch_frames"],
                tcmacs=st["cmacs"] / max(st["wall_seconds"], 1e-12) / 1e12)


def poisson_arrivals(n: int, rate_hz: float, seed: int) -> np.ndarray:
    gaps = np.random.default_rng(seed).exponential(1.0 / rate_hz, size=n)
    gaps[0] = 0.0
    return np.cumsum(gaps)


def saturated(serve, sid, frames, device: bool) -> dict:
    """All frames at time 0: the sustainable rate is the achieved one."""
    return run_arrivals(serve, sid, frames, np.zeros(len(frames)), device)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--elements", type=int, default=2048)
    ap.add_argument("--freq", type=int, default=8)
    ap.add_argument("--frame-samples", type=int, default=1024)
    ap.add_argument("--frames", type=int, default=1024, help="frames per run (1024 = 2^20 samples)")
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--loads", default="0.5,1,2", help="arrival rates as multiples of sustainable")
    ap.add_argument("--host-pool", type=int, default=32, help="pinned frames cycled by host runs")
    ap.add_argument("--max-item-stages", type=int, default=0,
                    help="cut GEMM work items into pieces of this many 64-sample stages (0 = whole)")
    ap.add_argument("--modes", default="device_resident,host_pcie")
    ap.add_argument("--method", choices=["fp6", "int8"], default="fp6",
                    help="matrix-core operands (bit-identical results)")
    ap.add_argument("--out", default="")
    a = ap.parse_args()

    import torch

    import chx
    import workloads as W

    if not torch.cuda.is_available():
        print("serve_driver needs a GPU", file=sys.stderr)
        return 2
    N, F, T, n = a.elements, a.freq, a.frame_samples, a.frames
    fb = T * F * N
    data = W.streaming_frame(torch, n * T, F, N, seed=4)  # [n*T][F][N] on the device
    torch.cuda.synchronize()
    base = data.data_ptr()
    dev_frames = [base + i * fb for i in range(n)]
    pool = torch.empty((a.host_pool, T, F, N), dtype=torch.uint8, pin_memory=True)
    pool.copy_(data[:a.host_pool * T].view(a.host_pool, T, F, N))
    host_frames = [pool[i % a.host_pool].data_ptr() for i in range(n)]
    loads = [float(x) for x in a.loads.split(",")]
    res = dict(workload=f"streaming: N={N}, F={F}, frames of {T} samples, {n} frames "
                        f"({n * T} samples), 5% of elements held at (7,-8)", runs={})
    method = chx.FP6_MFMA if a.method == "fp6" else chx.INT8_MFMA
    res["method"] = a.method
    with chx.Serve(frame_slots=a.slots, slot_bytes=fb,
                   max_item_stages=a.max_item_stages, method=method) as sv:
        for mode, frames, device in (("device_resident", dev_frames, True),
                                     ("host_pcie", host_frames, False)):
            if mode not in a.modes.split(","):
                continue
            sid = sv.open(N, F, T)
            sat = saturated(sv, sid, frames, device)
            rate = sat["achieved_rate"]
            runs = dict(saturated=sat)
            for k, load in enumerate(loads):
                runs[f"poisson_{load:g}x"] = run_arrivals(
                    sv, sid, frames, poisson_arrivals(n, load * rate, seed=10 + k), device)
            sv.close_stream(sid)
            res["runs"][mode] = runs
            print(f"{mode}: sustainable {rate:.0f} frames/s "
                  f"({sat['tcmacs']:.1f} TCMAC/s)", flush=True)
            for name, r in runs.items():
                print(f"  {name:14s} offered {r['offered_rate']:8.0f}/s  achieved "
                      f"{r['achieved_rate']:8.0f}/s  p50 {r['p50_ms']:8.3f} ms  p99 "
                      f"{r['p99_ms']:8.3f} ms  batch {r['mean_batch_frames']:.1f}", flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
