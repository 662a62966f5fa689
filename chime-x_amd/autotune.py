"""Autotuner of the batch launch shape (chx_batch_tuning).

For a job list it times every candidate shape -- the matrix-core method
(FP6 or int8 operands; bit-identical results), the work-item time extent
(max_item_stages: whole chunks, or pieces of 128..1024 stages whose partial
sums add atomically), the output tile shape (FP6: 128 x 128 tiles for arrays
of up to 256 or 512 elements, 128 x 256 otherwise), the persistent
workgroup count (one per compute unit,
or fewer, which lowers the power the matrix cores draw), and the workspace
budget (which sets the number of rounds) -- on the GPU with the captured
graph, and keeps the fastest.  Results are cached in a JSON file keyed by
the GPU target and the job list's shapes, so a later batch of the same
shapes on the same target loads its tuning without timing anything.

    python chime-x_amd/autotune.py --out chime-x_amd/tuning.json
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import torch  # noqa: E402  (before libchx: one HIP runtime in the process)

import chx  # noqa: E402
import workloads as W  # noqa: E402

DEFAULT_CACHE = chx.TUNING_CACHE


def signature(jobs) -> str:
    return chx.job_signature(jobs, chx.device_arch(0))


def candidates(cus: int, methods=(3, 2)):
    """Launch shapes to time; methods: 3 = FP6, 2 = int8 matrix cores
    (bit-identical results)."""
    stages = (0, 1024, 512, 256, 128)  # whole chunk (FP6: <= 2048 stages), or pieces
    wgs = (cus, cus * 7 // 8 // 8 * 8, cus // 2 // 8 * 8)
    budgets = (16 << 30, 64 << 30)
    for m, st, wg, ws in itertools.product(methods, stages, wgs, budgets):
        # tile shape (FP6): 128 x 128 tiles for arrays up to this many
        # elements, 128 x 256 above
        for small in ((0, 256, 512) if m == 3 else (0,)):
            yield dict(method=m, max_item_stages=st, workgroups=wg, workspace_bytes=ws,
                       small_tile_elements=small)


def time_batch(descs, tuning, reps: int = 3) -> float:
    """Milliseconds per graph replay (median of reps)."""
    with chx.Batch(descs, tuning) as b:
        s = torch.cuda.current_stream().cuda_stream
        b.run(stream=s)  # capture + first replay
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            b.run(stream=s)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
    ts.sort()
    return ts[len(ts) // 2]


def tune(jobs, descs, reps: int = 3, log=print, methods=(3, 2)) -> dict:
    cus = torch.cuda.get_device_properties(0).multi_processor_count // 8 * 8
    total = sum(W.cmacs(j["num_elements"], j["num_freq"], j["time_steps"]) for j in jobs)
    table = []
    for c in candidates(cus, methods):
        try:
            ms = time_batch(descs, c, reps)
        except chx.ChxError as e:  # e.g. a budget the device cannot hold
            log(f"  {c}: {e}")
            continue
        table.append(dict(tuning=c, ms=ms, tcmacs=total / ms / 1e9))
        log(f"  {c}: {ms:.3f} ms  {total / ms / 1e9:.1f} TCMAC/s")
    best = min(table, key=lambda r: r["ms"])
    return dict(signature=signature(jobs), best=best, table=table)


def load(jobs, path: str = DEFAULT_CACHE) -> dict | None:
    """The cached tuning of this job list's shapes, or None."""
    return chx.cached_tuning(jobs, path)


def save(result: dict, path: str = DEFAULT_CACHE) -> None:
    cache = {}
    if os.path.exists(path):
        with open(path) as f:
            cache = json.load(f)
    cache[result["signature"]] = result
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--njobs", type=int, default=64)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--sizes", default="256,512,1024,2048",
                    help="array sizes the job list draws from")
    ap.add_argument("--methods", default="3,2",
                    help="matrix-core methods to time (3 = FP6, 2 = int8)")
    ap.add_argument("--replay", type=int, default=0,
                    help="no tuning: replay the cached shape this many times and report it")
    a = ap.parse_args()
    if not torch.cuda.is_available():
        print("autotune needs a GPU", file=sys.stderr)
        return 2
    jobs = W.uneven_batch_jobs(a.njobs, a.seed,
                               sizes=tuple(int(x) for x in a.sizes.split(",")))
    ins, outs, descs = [], [], []
    for j in jobs:
        N, F, T = j["num_elements"], j["num_freq"], j["time_steps"]
        ins.append(W.device_uniform(torch, (T, F, N), j["seed"]))
        outs.append(torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32,
                                device="cuda"))
        descs.append(dict(j, d_input=ins[-1].data_ptr(), d_vis=outs[-1].data_ptr()))
    print(f"uneven batch: {len(jobs)} jobs, signature {signature(jobs)}", flush=True)
    if a.replay:
        tuning = load(jobs, a.out)
        ms = time_batch(descs, tuning, a.replay)
        total = sum(W.cmacs(j["num_elements"], j["num_freq"], j["time_steps"]) for j in jobs)
        print(json.dumps(dict(tuning=tuning, ms=ms, tcmacs=total / ms / 1e9)), flush=True)
        return 0
    res = tune(jobs, descs, a.reps, log=lambda m: print(m, flush=True),
               methods=tuple(int(x) for x in a.methods.split(",")))
    save(res, a.out)
    print(json.dumps(res["best"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
