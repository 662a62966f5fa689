"""A seeded random sweep of shapes and launch shapes, bit for bit against
the oracle: plans of every method and convention over random element
counts (multiples of 32 up to 1280), channel counts (1-4) and lengths
(multiples of 32 up to 3072), some at the all-(-8,-8) saturation input;
and uneven batches of 1-6 jobs (any length up to 4000 samples) under random
tunings (item length, workgroups, tile shape, workspace budget, eager or
graph)."""
import sys

import numpy as np
import pytest

import chx
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _torch():
    import torch
    return torch


@pytest.mark.parametrize("block", range(4))
def test_random_plans_against_oracle(block):
    rng = np.random.default_rng(1000 + block)
    for case in range(40):
        method = int(rng.integers(0, 4))
        U = int(rng.integers(0, 2))
        N = 32 * int(rng.integers(1, 41))
        F = int(rng.integers(1, 5))
        T = 32 * int(rng.integers(1, 97))
        while T * F * N * N > 1.0e9 and N > 32:
            N -= 32
        data = rng.integers(0, 256, size=(T, F, N), dtype=np.uint8)
        if case % 13 == 0:
            data[:] = 0
        say(f"  plan block={block} case={case} method={method} U={U} N={N} F={F} T={T}")
        with chx.Plan(N, F, T, 32, U, method) as p:
            got = p.correlate(data)
        np.testing.assert_array_equal(got, O.correlate_blocks(data, U),
                                      err_msg=f"method={method} U={U} N={N} F={F} T={T}")


@pytest.mark.parametrize("block", range(3))
def test_random_batches_against_oracle(block):
    torch = _torch()
    rng = np.random.default_rng(2000 + block)
    for case in range(8):
        method = chx.FP6_MFMA if rng.random() < 0.6 else chx.INT8_MFMA
        jobs, datas, outs = [], [], []
        for j in range(int(rng.integers(1, 7))):
            N = 32 * int(rng.integers(1, 25))
            F = int(rng.integers(1, 4))
            T = int(rng.integers(1, 4000))
            U = int(rng.integers(0, 2))
            g = torch.Generator(device="cuda").manual_seed(block * 1000 + case * 10 + j)
            d = torch.randint(0, 256, (T, F, N), dtype=torch.uint8, device="cuda", generator=g)
            o = torch.full((F, chx.num_blocks(N), 32, 32, 2), -5, dtype=torch.int32,
                           device="cuda")
            jobs.append(dict(num_elements=N, num_freq=F, time_steps=T,
                             upper_triangle_convention=U, d_input=d.data_ptr(),
                             d_vis=o.data_ptr()))
            datas.append(d)
            outs.append(o)
        tuning = dict(method=method)
        if rng.random() < 0.4:
            tuning["max_item_stages"] = int(rng.choice([4, 8, 16, 64]))
        if rng.random() < 0.4:
            tuning["workgroups"] = int(rng.choice([8, 16, 64, 128]))
        if method == chx.FP6_MFMA and rng.random() < 0.4:
            tuning["small_tile_elements"] = int(rng.choice([128, 256, 512]))
        if rng.random() < 0.3:
            per_round = sum(1.5 * j["num_freq"] * -(-j["num_elements"] // 128) * 128 * 1024
                            for j in jobs)
            tuning["workspace_bytes"] = int(max(2 * per_round, 1 << 22))
        eager = bool(rng.random() < 0.5)
        say(f"  batch block={block} case={case} jobs={[(j['num_elements'], j['num_freq'], j['time_steps']) for j in jobs]} tuning={tuning} eager={eager}")
        with chx.Batch(jobs, tuning) as b:
            b.run(eager=eager)
            torch.cuda.synchronize()
        for j, (job, d, o) in enumerate(zip(jobs, datas, outs)):
            np.testing.assert_array_equal(
                o.cpu().numpy(), O.correlate_blocks(d.cpu().numpy(),
                                                    job["upper_triangle_convention"]),
                err_msg=f"case {case} job {j} tuning {tuning}")
