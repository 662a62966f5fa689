"""Batches of uneven correlations (chx_batch_*) on the GPU, bit-exact.

Every job of a batch must equal the oracle and what the job gives on its
own (a single-job Plan), for every launch shape the autotuner can pick
(time pieces with atomic adds, few workgroups, many rounds), and every
replay of the captured graph must equal the eager launch sequence.
"""
import sys

import numpy as np
import pytest

import chx
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


MIXED = [  # (N, F, T, U): ragged T, N not a multiple of 128, several flush chunks
    (32, 1, 100, 1),
    (96, 3, 1000, 0),
    (160, 2, 4096, 1),
    (256, 1, 33000, 1),
    (384, 2, 512, 0),
    (64, 4, 2560, 1),
]
MIXED_PER_SAMPLE = 5120  # sum of 2 * F * (N rounded up to 128)


def _mixed(torch, seed=0):
    data, ins, outs, jobs = [], [], [], []
    for k, (N, F, T, U) in enumerate(MIXED):
        d = np.random.default_rng(seed + k).integers(0, 256, size=(T, F, N), dtype=np.uint8)
        di = torch.from_numpy(d).cuda()
        do = torch.full((F, chx.num_blocks(N), 32, 32, 2), 0x5A5A5A5A, dtype=torch.int32,
                         device="cuda")
        data.append(d)
        ins.append(di)
        outs.append(do)
        jobs.append(dict(num_elements=N, num_freq=F, time_steps=T, upper_triangle_convention=U,
                         d_input=di.data_ptr(), d_vis=do.data_ptr()))
    return data, ins, outs, jobs


@pytest.fixture(scope="module")
def mixed_oracle():
    want = []
    for k, (N, F, T, U) in enumerate(MIXED):
        d = np.random.default_rng(k).integers(0, 256, size=(T, F, N), dtype=np.uint8)
        want.append(O.correlate_blocks(d, U))
    return want


@pytest.mark.parametrize("tuning", [
    None,
    dict(max_item_stages=4),
    dict(max_item_stages=8, workgroups=16),
    dict(workspace_bytes=MIXED_PER_SAMPLE * 1024),
    dict(workspace_bytes=MIXED_PER_SAMPLE * 512, max_item_stages=4, workgroups=8),
    dict(small_tile_elements=4096),
    dict(small_tile_elements=160, max_item_stages=4, workgroups=16),
])
@pytest.mark.parametrize("eager", [True, False])
@pytest.mark.parametrize("method", [chx.INT8_MFMA, chx.FP6_MFMA])
def test_mixed_batch_against_oracle(mixed_oracle, tuning, eager, method):
    torch = _torch()
    data, ins, outs, jobs = _mixed(torch)
    with chx.Batch(jobs, dict(tuning or {}, method=method)) as b:
        say(f"  batch tuning={tuning} eager={eager} info={b.info()}")
        b.run(eager=eager)
        torch.cuda.synchronize()
    for k, (N, F, T, U) in enumerate(MIXED):
        np.testing.assert_array_equal(outs[k].cpu().numpy(), mixed_oracle[k],
                                      err_msg=f"job {k} N={N} F={F} T={T} U={U}")


@pytest.mark.parametrize("method", [chx.INT8_MFMA, chx.FP6_MFMA])
def test_rounds_follow_the_budget(method):
    torch = _torch()
    _, _, _, jobs = _mixed(torch)
    # one ring of every job: int8 256 samples x 2 bytes, fp6 256 x 1.5 bytes
    int8 = method == chx.INT8_MFMA
    unit = 256
    per_unit = MIXED_PER_SAMPLE * unit if int8 else MIXED_PER_SAMPLE * unit * 3 // 4
    with chx.Batch(jobs, dict(workspace_bytes=MIXED_PER_SAMPLE * 1024, method=method)) as b:
        inf = b.info()
    chunk = MIXED_PER_SAMPLE * 1024 // per_unit * unit
    assert inf["chunk_samples"] == chunk
    assert inf["rounds"] == -(-33000 // chunk)
    assert inf["cmacs"] == sum(W.cmacs(N, F, T) for N, F, T, _ in MIXED)
    with pytest.raises(chx.ChxError) as e:
        chx.Batch(jobs, dict(workspace_bytes=per_unit - 1, method=method))
    assert e.value.status == chx.ERR_BUDGET


@pytest.mark.parametrize("method", [chx.INT8_MFMA, chx.FP6_MFMA])
def test_graph_replays_match_eager_and_reread_inputs(method):
    torch = _torch()
    data, ins, outs, jobs = _mixed(torch, seed=50)
    with chx.Batch(jobs, dict(max_item_stages=8, method=method)) as b:
        b.run(eager=True)
        torch.cuda.synchronize()
        eager = [o.clone() for o in outs]
        for rep in range(3):
            for o in outs:  # the program's own eager work between replays
                o.fill_(-1)
            torch.cuda.synchronize()
            b.run(stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            for k in range(len(outs)):
                assert torch.equal(outs[k], eager[k]), f"replay {rep} job {k}"
        # new data in the same input buffers: the next replay sees it
        for k, (N, F, T, U) in enumerate(MIXED):
            d = np.random.default_rng(900 + k).integers(0, 256, size=(T, F, N), dtype=np.uint8)
            ins[k].copy_(torch.from_numpy(d))
            data[k] = d
        b.run()
        torch.cuda.synchronize()
    for k, (N, F, T, U) in enumerate(MIXED):
        np.testing.assert_array_equal(outs[k].cpu().numpy(), O.correlate_blocks(data[k], U),
                                      err_msg=f"job {k}")


def _pairs_check(torch, d_in, d_vis, N, U, rng, nsub=16):
    """Oracle on a sorted element subset: its triangle equals the batch's
    entries at those elements' pairs (rows y <= columns x keep their order)."""
    E = np.sort(rng.choice(N, size=nsub, replace=False))
    sub = d_in[:, :, torch.as_tensor(E, device="cuda")].contiguous().cpu().numpy()
    want = O.correlate_triangle(sub, U)  # [F][nsub(nsub+1)/2][2]
    vis = d_vis
    nb = N // 32
    a = 0
    ys, xs = [], []
    for i in range(nsub):
        for j in range(i, nsub):
            ys.append(E[i])
            xs.append(E[j])
            a += 1
    ys, xs = np.array(ys), np.array(xs)
    by, bx = ys // 32, xs // 32
    b = by * nb - by * (by - 1) // 2 + (bx - by)
    idx = torch.as_tensor(b, device="cuda"), torch.as_tensor(ys % 32, device="cuda"), \
        torch.as_tensor(xs % 32, device="cuda")
    got = vis[:, idx[0], idx[1], idx[2], :].cpu().numpy()
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("method,small", [(chx.INT8_MFMA, 0), (chx.FP6_MFMA, 0),
                                          (chx.FP6_MFMA, 512)])
def test_uneven_batch_64_jobs_matches_alone_and_oracle(method, small):
    """The 64-job uneven batch: each job's output from the one graph launch
    equals the job correlated alone (its own Plan) bit for bit, and the
    oracle on a 16-element subset of every job; FP6 also with the arrays of
    up to 512 elements on 128 x 128 tiles."""
    torch = _torch()
    jobs = W.uneven_batch_jobs()
    ins, outs, descs = [], [], []
    for j in jobs:
        N, F, T = j["num_elements"], j["num_freq"], j["time_steps"]
        ins.append(W.device_uniform(torch, (T, F, N), j["seed"]))
        outs.append(torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32,
                                device="cuda"))
        descs.append(dict(j, d_input=ins[-1].data_ptr(), d_vis=outs[-1].data_ptr()))
    total = sum(W.cmacs(j["num_elements"], j["num_freq"], j["time_steps"]) for j in jobs)
    with chx.Batch(descs, dict(method=method, small_tile_elements=small)) as b:
        say(f"  uneven batch: {b.info()}")
        b.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.run(stream=torch.cuda.current_stream().cuda_stream)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        say(f"  uneven batch: {total / ms / 1e9:.1f} TCMAC/s ({ms:.2f} ms, {total:.3e} CMAC)")
    rng = np.random.default_rng(5)
    for k, j in enumerate(jobs):
        N, F, T, U = j["num_elements"], j["num_freq"], j["time_steps"], 1
        with chx.Plan(N, F, T, 256, U, method) as p:
            alone = torch.empty_like(outs[k])
            p.correlate_device(ins[k].data_ptr(), alone.data_ptr(),
                               torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
        assert torch.equal(alone, outs[k]), f"job {k} {j}"
        _pairs_check(torch, ins[k], outs[k], N, U, rng)


@pytest.mark.parametrize("byte, re_per_sample", [(0xFF, 98), (0x00, 128)])
def test_fp6_items_at_the_code_bound(byte, re_per_sample):
    """FP6 items of 1024 stages (the most one holds) with every sample (7,7)
    (codes 15: every accumulator climbs 450 per sample from -2^24 to
    -2^24 + 450 * 65536, the largest value it holds) or (-8,-8) (codes 0:
    the accumulators stay at -2^24 and the offset corrections carry the
    whole result), 2052 stages in pieces of 1024 + 1024 + 4 that add
    atomically: re = 98 T or 128 T, im = 0."""
    torch = _torch()
    T, F, N = 2048 * 64 + 256, 2, 128
    x = torch.full((T, F, N), byte, dtype=torch.uint8, device="cuda")
    vis = torch.full((F, chx.num_blocks(N), 32, 32, 2), 3, dtype=torch.int32, device="cuda")
    job = dict(num_elements=N, num_freq=F, time_steps=T, upper_triangle_convention=1,
               d_input=x.data_ptr(), d_vis=vis.data_ptr())
    with chx.Batch([job], dict(method=chx.FP6_MFMA, workspace_bytes=1 << 30)) as b:
        inf = b.info()
        say(f"  bound batch: {inf}")
        assert inf["rounds"] == 1 and inf["items"] == F * 1 * 3
        b.run()
        torch.cuda.synchronize()
    v = vis.cpu().numpy()
    assert (v[..., 0] == re_per_sample * T).all() and (v[..., 1] == 0).all()


def test_fp6_item_near_the_bound_with_mixed_signs():
    """Visibilities near 2^24 with products of both signs: for the first
    2047 stages, elements of a flip set alternate per sample between
    (-8,-8) and (7,7) (products +128 / -112 against the others, +128 / +98
    among themselves) and the rest hold (-8,-8) (+128 per sample); the last
    stage and the one-ring rest are uniform random.  FP6 (items of 1024
    stages on offset-binary codes, pieces added atomically) equals the int8
    method (one item of 2048 stages) and the oracle."""
    torch = _torch()
    T, F, N = 2048 * 64 + 256, 1, 128
    t_const = 2047 * 64
    data = np.random.default_rng(24).integers(0, 256, size=(T, F, N), dtype=np.uint8)
    data[:t_const] = 0x00
    flip = np.random.default_rng(25).choice(N, size=N // 3, replace=False)
    odd = np.arange(t_const) % 2 == 1
    data[np.ix_(np.nonzero(odd)[0], [0], flip)] = 0xFF
    want = O.correlate_blocks(data, 1)
    assert np.abs(want[..., 0]).max() > (1 << 24) - 8192  # near the cap
    x = torch.from_numpy(data).cuda()
    got = {}
    for method in (chx.FP6_MFMA, chx.INT8_MFMA):
        vis = torch.full((F, chx.num_blocks(N), 32, 32, 2), 3, dtype=torch.int32,
                         device="cuda")
        job = dict(num_elements=N, num_freq=F, time_steps=T, upper_triangle_convention=1,
                   d_input=x.data_ptr(), d_vis=vis.data_ptr())
        with chx.Batch([job], dict(method=method, workspace_bytes=1 << 30)) as b:
            inf = b.info()
            say(f"  near-bound batch {chx.METHOD_NAMES[method]}: {inf}")
            assert inf["rounds"] == 1 and inf["items"] == F * 1 * (3 if method == chx.FP6_MFMA
                                                                 else 2)
            b.run()
            torch.cuda.synchronize()
        got[method] = vis.cpu().numpy()
    np.testing.assert_array_equal(got[chx.FP6_MFMA], got[chx.INT8_MFMA])
    np.testing.assert_array_equal(got[chx.FP6_MFMA], want)
