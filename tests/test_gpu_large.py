"""The large/long configuration on the GPU: N=4096, F=32, 2^20 samples
integrated in int32 (128 GiB of input resident in HBM, many plane chunks),
for both matrix-core methods.

Checks that hold at full size without a full CPU correlation:
  * saturation: every byte 0x00 = (-8,-8) gives re = 128 * 2^20 =
    134,217,728 and im = 0 for every pair (the int32 worst case);
  * repeat mode (the reference generator's default reseeds every sample,
    input_generator.c:52-54, so every sample is the same): V = T * V_1, with
    V_1 the oracle's visibilities of one sample, on sampled blocks;
  * uniform random: the oracle on a sorted 16-element subset over all 2^20
    samples for three channels equals the corresponding entries, and
    linearity in time (first half + second half = whole, mod 2^32).
"""
import sys

import numpy as np
import pytest

import chx
import oracle as O

pytestmark = pytest.mark.gpu

N, F, T = 4096, 32, 1 << 20


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def buffers():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    x = torch.empty((T, F, N), dtype=torch.uint8, device="cuda")
    out = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device="cuda")
    yield torch, x, out
    del x, out
    torch.cuda.empty_cache()


@pytest.fixture(scope="module", params=[chx.INT8_MFMA, chx.FP6_MFMA],
                ids=["int8_mfma", "fp6_mfma"])
def big(buffers, request):
    torch, x, out = buffers
    plan = chx.Plan(N, F, T, 256, 1, request.param)
    say(f"  method {chx.METHOD_NAMES[request.param]}")
    yield torch, x, out, plan
    plan.close()


def _run(torch, plan, x, out):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    plan.correlate_device(x.data_ptr(), out.data_ptr())
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    say(f"  N={N} F={F} T={T}: {ms:.1f} ms, "
        f"{F * T * (N * (N + 1) // 2) / ms / 1e9:.1f} TCMAC/s")


def test_saturation_closed_form(big):
    torch, x, out, plan = big
    x.zero_()
    _run(torch, plan, x, out)
    assert bool((out[..., 0] == 128 * T).all())
    assert bool((out[..., 1] == 0).all())


def test_repeat_mode_closed_form(big):
    torch, x, out, plan = big
    row = np.random.default_rng(42).integers(0, 256, size=(1, F, N), dtype=np.uint8)
    x.copy_(torch.from_numpy(row).cuda().expand(T, F, N))
    _run(torch, plan, x, out)
    rng = np.random.default_rng(1)
    blocks = np.sort(rng.choice(chx.num_blocks(N), size=64, replace=False))
    v1 = O.correlate_blocks(row, 1)[:, blocks].astype(np.int64)  # one sample
    got = out[:, torch.as_tensor(blocks, device="cuda")].cpu().numpy()
    want = (v1 * T).astype(np.uint32).view(np.int32)  # int32 wrap
    np.testing.assert_array_equal(got, want)


def test_uniform_subset_oracle_and_time_linearity(big):
    torch, x, out, plan = big
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    torch.randint(0, 256, x.shape, dtype=torch.uint8, device="cuda", generator=g, out=x)
    _run(torch, plan, x, out)
    rng = np.random.default_rng(7)
    E = np.sort(rng.choice(N, size=16, replace=False))
    chans = [0, 13, 31]
    ci = torch.as_tensor(chans, device="cuda")[:, None]
    ei = torch.as_tensor(E, device="cuda")[None, :]
    sub = x[:, ci, ei].contiguous().cpu().numpy()  # [T][3][16]
    want = O.correlate_triangle(sub, 1)  # [3][136][2]
    ys, xs = np.triu_indices(16)
    Y, X = E[ys], E[xs]
    nb = N // 32
    by, bx = Y // 32, X // 32
    b = by * nb - by * (by - 1) // 2 + (bx - by)
    got = out[torch.as_tensor(chans, device="cuda")][
        :, torch.as_tensor(b, device="cuda"), torch.as_tensor(Y % 32, device="cuda"),
        torch.as_tensor(X % 32, device="cuda")].cpu().numpy()
    np.testing.assert_array_equal(got, want)
    # linearity in time: two half-length plans add up to the whole (mod 2^32)
    half = T // 2
    a = torch.empty_like(out)
    with chx.Plan(N, F, half, 256, 1, plan.cfg.kernel_batch) as p:
        p.correlate_device(x.data_ptr(), a.data_ptr())
        p.corr(x[half:].data_ptr(), a.data_ptr())  # adds the second half
        torch.cuda.synchronize()
    assert torch.equal(a, out)
