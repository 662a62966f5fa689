#!/usr/bin/env python3
"""Benchmark of the X-engine correlator on one MI355X.

Workload (the metric's configuration): array-scale correlation, N=2048
elements, F=16 channels, T=65536 samples of uniform 4-bit complex data from
a seed, device-resident input ([T][F][N] bytes, 2 GiB) and block-order int32
output (16 x 2080 blocks, 260 MiB).  One step = one chx_correlate call (the
whole integration of all 16 channels) through the C ABI.

Prints ONE JSON line.  Both matrix-core methods are timed: int8 operands
with int32 accumulation, and FP6 E2M3 operands (every 4-bit value v exact as v/8)
with FP32 accumulation kept below 2^24 (exact); their outputs are compared
bit for bit.  value = complex MACs per second (Tera) of the faster one,
algorithmic work F*T*N(N+1)/2 per step; its roofline is priced at its own
dense peak (int8 5.0 POP/s, FP6 10.1 PFLOP/s).  The plain packed-integer
paths (conference, highly packed) are timed beside them, and the conference
output is compared bit for bit too.

Three more workloads are reported beside the line (not in `value`): the
uneven batch of 64 jobs (N in {256..2048}, 1-8 channels, 4096-131072
samples) as one captured graph replay with its cached tuning, the streaming
engine (N=2048, F=8, 1024-sample frames, device-resident) saturated and at
a Poisson load equal to the saturated rate, and the large/long integration
(N=4096, F=32, 2^20 samples) with both matrix-core methods.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "chime-x_amd"))

N_ELEM, N_FREQ, T_SAMPLES = 2048, 16, 65536
INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 matrix peak, 5.0 POP/s
FP6_PEAK_TOPS = 10100.0  # MI355X dense FP6 matrix peak, 10.1 PFLOP/s
MATRIX = {  # method name -> (kernel_batch, GEMM kernel, dtype, peak, plane bytes per value)
    "int8_mfma": (2, "corr_mfma_kernel", "int8 operands, int32 accumulation", INT8_PEAK_TOPS,
                  1.0),
    "fp6_mfma": (3, "corr_f6_kernel", "fp6-e2m3 operands, fp32 accumulation (exact)",
                 FP6_PEAK_TOPS, 0.75),
}


def cmac(F, T, N):
    return F * T * N * (N + 1) // 2


def cpu_info() -> dict:
    """Host cores: online count, the ones this process may use, the model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
              else (os.cpu_count() or 1))
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def cpu_baseline(budget_s: float = 12.0):
    """The reference CPU correlator (cpu_corr_test.c:156-203, generator
    included) on channels of the workload, T scaled so a run takes about
    budget_s / 2 seconds: once on one thread, then one channel per thread on
    every usable host core (up to 16, the channels being independent).
    Falls back to the C restatement when the reference is not built."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    ref = os.path.join(HERE, "oracle", "_ref", "libchxref.so")
    N = N_ELEM
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        fn = L.cpu_data_generate_and_correlate_upper_triangle_only
        fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] + [ctypes.c_int] * 9
        kind = "reference"

        def call(T, out):
            return fn(T, 1, N, out, 4, 42, 0, 0, 0, 0, -1, 1, 0)
    else:
        import oracle as O
        L = O.lib()
        kind = "port"

        def call(T, out):
            return L.chxo_generate_and_correlate_triangle(T, 1, N, out, 4, 42, 0, 0, 0, 0,
                                                          -1, 1, 1)
    out = (ctypes.c_int32 * (N * (N + 1)))()
    T = 8
    while True:
        t0 = time.perf_counter()
        rc = call(T, out)
        dt = time.perf_counter() - t0
        if rc != 0:
            return None
        if dt * 4 > budget_s or T >= 4096:
            break
        T = min(4096, max(T * 2, int(T * budget_s / 2 / max(dt, 1e-3) / 1.5) // 8 * 8))
    one = cmac(1, T, N) / dt / 1e12
    # one channel per thread (ctypes releases the GIL during the call)
    import threading
    info = cpu_info()
    cores = max(1, min(16, info["usable"]))
    outs = [(ctypes.c_int32 * (N * (N + 1)))() for _ in range(cores)]
    th = [threading.Thread(target=call, args=(T, outs[k])) for k in range(cores)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    dtm = time.perf_counter() - t0
    return {
        "value": cores * cmac(1, T, N) / dtm / 1e12,
        "unit": "TCMAC/s",
        "cores": cores,
        "kind": kind,
        "single_thread_value": one,
        "host": info,
        "sample": f"N={N}, one channel of the workload per thread, T={T} continuous seeded "
                  f"random, generator included: 1 thread {dt:.2f} s, {cores} threads "
                  f"{dtm:.2f} s wall",
        "note": "threads capped at 16 (the host share of one GPU); the reference loop "
                "(t outer, cpu_corr_test.c:176-198) sweeps the whole N(N+1)/2 int32 "
                "triangle (16.8 MB at N=2048) once per sample on every thread, so the "
                "threads share the memory bandwidth and scale sub-linearly",
    }


def from_profile(kernel: str, key: str):
    """A figure recorded in the kernel's rocprofv3 summary under profiles/:
    traffic_bytes_per_launch (bytes per GEMM launch between the L2s and
    memory: the L2's 128-byte read requests and 64-byte write requests to
    the fabric, PMC, checked against a device copy of known size),
    mfma_busy_pct (MfmaUtil: SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE x
    SIMDs, steady launches) or clock_ghz (GRBM_GUI_ACTIVE per XCD over the
    kernel time)."""
    import re
    path = os.path.join(HERE, "profiles", f"{kernel.replace('_kernel', '')}_N2048_F16_T65536.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(key + r":\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def copy_calibration(torch, gib: int = 1, reps: int = 5):
    """A plain device copy of known size, for calibrating PMC byte counters."""
    n = gib << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "reps": reps,
                      "expected_fetch_bytes": n, "expected_write_bytes": n}))


def batch_uneven64(torch, chx, reps: int = 5) -> dict:
    """The 64-job uneven batch as one graph replay, with its cached tuning."""
    import workloads as W
    jobs = W.uneven_batch_jobs()
    ins, outs, descs = [], [], []
    for j in jobs:
        N, F, T = j["num_elements"], j["num_freq"], j["time_steps"]
        ins.append(W.device_uniform(torch, (T, F, N), j["seed"]))
        outs.append(torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32,
                                device="cuda"))
        descs.append(dict(j, d_input=ins[-1].data_ptr(), d_vis=outs[-1].data_ptr()))
    total = sum(W.cmacs(j["num_elements"], j["num_freq"], j["time_steps"]) for j in jobs)
    tuning = chx.cached_tuning(jobs)
    stream = torch.cuda.current_stream()
    with chx.Batch(descs, tuning) as b:
        b.run(stream=stream.cuda_stream)  # capture + first replay
        b.run(stream=stream.cuda_stream)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            b.run(stream=stream.cuda_stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        info = b.info()
    del ins, outs
    torch.cuda.empty_cache()
    ms = sorted(ts)[len(ts) // 2]
    return {"workload": "64 jobs, N in {256,512,1024,2048}, 1-8 channels, 4096-131072 "
                        "samples (seed 3), one captured graph replay",
            "value": round(total / ms / 1e9, 2), "unit": "TCMAC/s", "ms_per_replay": round(ms, 3),
            "replays_timed": reps, "cmacs": total, "tuning": tuning,
            "rounds": info["rounds"], "items": info["items"]}


def serve_streaming(torch, chx, frames: int = 1024) -> dict:
    """Streaming engine, N=2048, F=8, 1024-sample frames from device memory
    (5% of the elements held at (7,-8)): the saturated rate, then Poisson
    arrivals at 0.5, 1 and 2 times that rate with p50/p99 frame-to-visibility
    latency."""
    import serve_driver as D
    import workloads as W
    N, F, T = 2048, 8, 1024
    fb = T * F * N
    data = W.streaming_frame(torch, frames * T, F, N, seed=4)
    torch.cuda.synchronize()
    ptrs = [data.data_ptr() + i * fb for i in range(frames)]
    with chx.Serve(frame_slots=64, slot_bytes=fb) as sv:
        sid = sv.open(N, F, T)
        sat = D.saturated(sv, sid, ptrs, True)
        loads = {}
        for x, seed in ((1.0, 11), (0.5, 12), (2.0, 13)):
            loads[x] = D.run_arrivals(
                sv, sid, ptrs, D.poisson_arrivals(frames, x * sat["achieved_rate"], seed=seed),
                True)
        sv.close_stream(sid)
    del data
    torch.cuda.empty_cache()
    return {"workload": f"N={N}, F={F}, {frames} frames of {T} samples (2^20 samples), "
                        "device-resident frames, 5% of elements held at (7,-8)",
            "saturated_frames_per_s": round(sat["achieved_rate"], 1),
            "saturated_tcmacs": round(sat["tcmacs"], 2),
            **{f"poisson_{x:g}x": {"offered_frames_per_s": round(r["offered_rate"], 1),
                                   "achieved_frames_per_s": round(r["achieved_rate"], 1),
                                   "p50_ms": round(r["p50_ms"], 3),
                                   "p99_ms": round(r["p99_ms"], 3),
                                   "mean_batch_frames": round(r["mean_batch_frames"], 2)}
               for x, r in loads.items()}}


def large_long(torch, chx) -> dict:
    """N=4096, F=32, 2^20 samples integrated in int32 (128 GiB of input
    resident, many plane chunks): one warm-up and one timed integration per
    matrix-core method, outputs compared bit for bit."""
    import workloads as W
    N, F, T = 4096, 32, 1 << 20
    x = W.device_uniform(torch, (T, F, N), 5)
    outs, res = [], {}
    for name, method in (("fp6_mfma", chx.FP6_MFMA), ("int8_mfma", chx.INT8_MFMA)):
        out = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device="cuda")
        with chx.Plan(N, F, T, 256, 1, method) as p:
            p.correlate_device(x.data_ptr(), out.data_ptr())
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            p.correlate_device(x.data_ptr(), out.data_ptr())
            e1.record()
            torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        res[name] = {"value": round(cmac(F, T, N) / (ms * 1e-3) / 1e12, 2), "unit": "TCMAC/s",
                     "ms_per_integration": round(ms, 2)}
        outs.append(out)
    res["bit_exact_fp6_vs_int8"] = bool(torch.equal(outs[0], outs[1]))
    del x, outs
    torch.cuda.empty_cache()
    return {"workload": f"N={N}, F={F}, T=2^20, uniform random bytes (seed 5), device-resident",
            **res}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # The GPU clock ramps up over the first ~20 ms of work after idle; ten
    # untimed steps let it settle before the timed ones.
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-packed", action="store_true",
                    help="skip timing the plain packed-integer paths")
    ap.add_argument("--only", choices=["int8_mfma", "fp6_mfma", "conference", "highly_packed"],
                    help="time one path only (for profiling)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the uneven-batch and streaming-engine figures")
    ap.add_argument("--extra", choices=["batch_uneven64", "serve_streaming", "large_long"],
                    help="only run this one of the extra workloads (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(frequency channels shard across GPUs with no exchange; see the design notes)",
              file=sys.stderr)
        return 2

    import torch
    import chx

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 2
    if args.copy_calibration:
        copy_calibration(torch)
        return 0
    if args.extra:
        fn = {"batch_uneven64": batch_uneven64, "serve_streaming": serve_streaming,
              "large_long": large_long}[args.extra]
        print(json.dumps({args.extra: fn(torch, chx)}), flush=True)
        return 0
    dev = torch.device("cuda:0")
    F, T, N = N_FREQ, T_SAMPLES, N_ELEM
    g = torch.Generator(device=dev)
    g.manual_seed(42)
    d_in = torch.randint(0, 256, (T, F, N), dtype=torch.uint8, device=dev, generator=g)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    work = cmac(F, T, N)

    def time_path(method, ta, steps, warmup, out):
        with chx.Plan(N, F, T, ta, 1, method) as p:
            for _ in range(warmup):
                p.correlate_device(d_in.data_ptr(), out.data_ptr(), s)
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
            t0 = time.perf_counter()
            ev[0].record(stream)
            for i in range(steps):
                p.correlate_device(d_in.data_ptr(), out.data_ptr(), s)
                ev[i + 1].record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
        return sum(per) / len(per), wall

    def kernel_times(method, steps, out):
        """Per-launch unpack and GEMM durations from HIP events the library
        records on its stream around each kernel (a separate, untimed pass)."""
        with chx.Plan(N, F, T, 256, 1, method) as p:
            p.correlate_device(d_in.data_ptr(), out.data_ptr(), s)
            torch.cuda.synchronize()
            p.profile(True)
            for _ in range(steps):
                p.correlate_device(d_in.data_ptr(), out.data_ptr(), s)
            u, g, n = p.kernel_times()
        return u / n, g / n, n

    only = args.only
    mat = {}
    outs = {}
    for name, (method, kernel, dtype, peak, pbytes) in MATRIX.items():
        if only not in (None, name):
            continue
        out = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device=dev)
        ms, wall = time_path(method, 256, args.steps, args.warmup, out)
        unpack_ms, gemm_ms, launches = kernel_times(method, max(3, min(args.steps, 10)), out)
        outs[name] = out
        achieved = 8 * work / (gemm_ms * 1e-3) / 1e12  # GEMM kernel alone
        # unpack: reads the input once, writes 2 planes (N and T are tile
        # multiples here, so no padding)
        ubytes = T * F * N + 2 * pbytes * T * F * N
        mat[name] = {
            "value": round(work / (ms * 1e-3) / 1e12, 3), "ms_per_step": round(ms, 4),
            "wall_s_timed": round(wall, 4), "dtype": dtype,
            "pct_of_own_peak": round(100 * 8 * work / (ms * 1e-3) / 1e12 / peak, 2),
            "pct_int8_peak": round(100 * 8 * work / (ms * 1e-3) / 1e12 / INT8_PEAK_TOPS, 2),
            "roofline": {
                "bound": "mfma", "kernel": kernel, "achieved": round(achieved, 2),
                "peak": peak, "unit": "TFLOP/s",
                "ops": "8 * F*T*N(N+1)/2 per launch (1 MAC = 2 ops, 4 real MACs per complex "
                       "MAC), over the kernel's HIP-event duration",
                "frac": round(achieved / peak, 4), "kernel_ms": round(gemm_ms, 4),
                "launches_timed": launches,
                "traffic": from_profile(kernel, "traffic_bytes_per_launch"),
                "mfma_busy_pct": from_profile(kernel, "mfma_busy_pct"),
                "clock_ghz": from_profile(kernel, "clock_ghz"),
            },
            "unpack_kernel": {
                "bound": "hbm", "ms": round(unpack_ms, 4),
                "achieved": round(ubytes / (unpack_ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                "unit": "GB/s", "frac": round(ubytes / (unpack_ms * 1e-3) / 8e12, 4),
                "bytes": f"input T*F*N read + 2 planes of {pbytes:g} byte per value written",
            },
        }
    if len(outs) == 2:
        same = bool(torch.equal(outs["int8_mfma"], outs["fp6_mfma"]))
        for name in mat:
            mat[name]["bit_exact_vs_other_matrix_path"] = same
    ref_out = next(iter(outs.values()), None)
    packed = {}
    if not args.no_packed and only in (None, "conference", "highly_packed"):
        out_p = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device=dev)
        # The packed paths are ~20x slower: a few steps are enough.
        k = max(1, min(3, args.steps))
        for name, method, ta in (("conference", chx.CONFERENCE, 256),
                                 ("highly_packed", chx.HIGHLY_PACKED, 2048)):
            if only not in (None, name):
                continue
            ms, _ = time_path(method, ta, k, 1, out_p)
            packed[name] = {"value": work / (ms * 1e-3) / 1e12, "ms_per_step": ms,
                            "time_accum": ta}
            if ref_out is not None:
                packed[name]["bit_exact_vs_matrix_paths"] = bool(torch.equal(out_p, ref_out))
        del out_p
    if not mat:
        print(json.dumps({"only": only, "packed": packed}))
        return 0

    extras = {}
    if not args.no_extras and only is None:
        del outs, ref_out, d_in
        torch.cuda.empty_cache()
        extras["batch_uneven64"] = batch_uneven64(torch, chx)
        extras["serve_streaming"] = serve_streaming(torch, chx)
        extras["large_long"] = large_long(torch, chx)

    best = max(mat, key=lambda n: mat[n]["value"])
    b = mat[best]
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    line = {
        "metric": "complex MAC/s (Tera) at N=2048, 16 freq; % INT8 MFMA peak; bit-exact",
        "value": b["value"],
        "unit": "TCMAC/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": b["dtype"],
        "data": "synthetic: uniform random bytes (4-bit complex in [-8,7]) from seed 42, "
                "device-resident",
        "config": {"workload": "array-scale N=2048 F=16 T=65536", "N": N, "F": F, "T": T,
                   "method": best, "time_accum": 256,
                   "output": "block order, 16 x 2080 x 32 x 32 x (re,im) int32"},
        # value (8 int8-equivalent ops per complex MAC) against the int8
        # MFMA peak, and the int8 path's own whole step against that peak
        "pct_int8_peak": b["pct_int8_peak"],
        "int8_path_pct_int8_peak": (mat["int8_mfma"]["pct_int8_peak"]
                                    if "int8_mfma" in mat else None),
        "wall_s_timed": b["wall_s_timed"],
        "roofline": b["roofline"],
        "unpack_kernel": b["unpack_kernel"],
        "matrix_paths": mat,
        "packed_paths": packed,
        "speedup_vs_conference": (round(b["value"] / packed["conference"]["value"], 2)
                                  if "conference" in packed else None),
        "pct_int8_peak_path": best,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
