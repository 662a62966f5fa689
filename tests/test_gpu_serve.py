"""Streaming engine (chx_serve_*) on the GPU, bit-exact.

A stream's accumulator must equal the oracle / a single Plan over the
concatenation of its frames, whether the stream runs alone or mixed with
other streams whose frames are submitted concurrently (continuous batching:
frames of several streams share launches, copies overlap computation, slots
run out and are waited for).  Admission by HBM budget and by integration
length is checked, and the streaming workload (N=2048, F=8, 1024-sample
frames, 5% of the elements held at (7,-8), 2^20 samples) integrates to the
same visibilities as one Plan over the whole input.
"""
import ctypes
import sys
import threading
import time

import numpy as np
import pytest

import chx
import oracle as O
import serve_driver as D
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


def _frames(N, F, T, k, seed):
    return [np.random.default_rng(seed + i).integers(0, 256, size=(T, F, N), dtype=np.uint8)
            for i in range(k)]


def _feed(sv, sid, frames, device_tensors=None):
    """Submits frames in order, waiting for the oldest in flight when busy."""
    inflight = []
    for i, fr in enumerate(frames):
        while True:
            if device_tensors is not None:
                fid = sv.submit_device(sid, device_tensors[i].data_ptr())
            else:
                fid = sv.submit(sid, fr)
            if fid is not None:
                inflight.append(fid)
                break
            if inflight:
                sv.wait_frame(inflight.pop(0))
            else:  # other streams hold every slot
                time.sleep(1e-4)
    return inflight


@pytest.mark.parametrize("U", [1, 0])
@pytest.mark.parametrize("method", [chx.INT8_MFMA, chx.FP6_MFMA])
def test_stream_alone_equals_oracle(U, method):
    N, F, T, k = 96, 2, 256, 6
    frames = _frames(N, F, T, k, 300)
    with chx.Serve(frame_slots=3, slot_bytes=N * F * T, method=method) as sv:
        sid = sv.open(N, F, T, U)
        _feed(sv, sid, frames)
        got = sv.read(sid)
        say(f"  alone U={U}: {sv.stats()}")
    np.testing.assert_array_equal(got, O.correlate_blocks(np.concatenate(frames), U))


MIX = [  # (N, F, T per frame, frames, U, device frames?)
    (2048, 2, 1024, 12, 1, False),
    (256, 1, 300, 16, 1, True),
    (160, 3, 512, 10, 0, False),
    (64, 4, 96, 20, 1, True),
]


@pytest.mark.parametrize("max_item_stages", [0, 8])
@pytest.mark.parametrize("method", [chx.INT8_MFMA, chx.FP6_MFMA])
def test_streams_in_a_mix_equal_alone(max_item_stages, method):
    torch = _torch()
    data = [_frames(N, F, T, k, 1000 * j) for j, (N, F, T, k, U, dev) in enumerate(MIX)]
    dev = [[torch.from_numpy(fr).cuda() for fr in frs] if d else None
           for frs, (N, F, T, k, U, d) in zip(data, MIX)]
    torch.cuda.synchronize()
    slot = max(N * F * T for N, F, T, *_ in MIX)

    def engine():
        return chx.Serve(frame_slots=6, slot_bytes=slot, max_batch_frames=8,
                         max_item_stages=max_item_stages, method=method)

    # alone: one stream per engine run
    alone = []
    for j, (N, F, T, k, U, d) in enumerate(MIX):
        with engine() as sv:
            sid = sv.open(N, F, T, U)
            _feed(sv, sid, data[j], dev[j])
            alone.append(sv.read(sid))
    # mixed: all streams at once, one submitting thread each
    with engine() as sv:
        sids = [sv.open(N, F, T, U) for N, F, T, k, U, d in MIX]
        errs = []

        def run(j):
            try:
                _feed(sv, sids[j], data[j], dev[j])
            except Exception as e:  # reported below
                errs.append(e)

        th = [threading.Thread(target=run, args=(j,)) for j in range(len(MIX))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, errs
        mixed = [sv.read(s) for s in sids]
        st = sv.stats()
        say(f"  mix max_item_stages={max_item_stages}: {st}")
        assert st["frames_done"] == sum(m[3] for m in MIX)
    for j, (N, F, T, k, U, d) in enumerate(MIX):
        np.testing.assert_array_equal(mixed[j], alone[j], err_msg=f"stream {j}")
        np.testing.assert_array_equal(alone[j], O.correlate_blocks(np.concatenate(data[j]), U),
                                      err_msg=f"stream {j}")


@pytest.mark.parametrize("method", [chx.INT8_MFMA, chx.FP6_MFMA])
def test_admission_by_budget_and_integration_length(method):
    N, F, T = 256, 1, 512
    acc = F * chx.num_blocks(N) * 2048 * 4
    slot = N * F * T
    fixed = 2 * slot + 2 * 2 * slot  # slots + default workspace
    with chx.Serve(frame_slots=2, slot_bytes=slot, hbm_budget_bytes=fixed + acc,
                   method=method) as sv:
        a = sv.open(N, F, T, integration_frames=2)
        assert a is not None
        assert sv.open(N, F, T) is None  # no room for a second accumulator
        frames = _frames(N, F, T, 3, 77)
        _feed(sv, a, frames[:2])
        with pytest.raises(chx.ChxError) as e:
            sv.submit(a, frames[2])  # past its integration length
        assert e.value.status == chx.ERR_BUDGET
        got = sv.read(a)
        sv.close_stream(a)
        b = sv.open(N, F, T)  # admitted once memory is released
        assert b is not None
        _feed(sv, b, frames[2:])
        np.testing.assert_array_equal(sv.read(b), O.correlate_blocks(frames[2], 1))
    np.testing.assert_array_equal(got, O.correlate_blocks(np.concatenate(frames[:2]), 1))


def test_streaming_workload_2pow20_samples():
    """N=2048, F=8, 1024 frames of 1024 samples (5% of elements at 0xF0),
    device-resident frames at saturation and at Poisson arrivals; host
    frames over PCIe from a pinned pool.  The accumulators equal a Plan over
    the same samples."""
    torch = _torch()
    N, F, T, n = 2048, 8, 1024, 1024
    fb = T * F * N
    data = W.streaming_frame(torch, n * T, F, N, seed=4)
    held = W.held_elements(N)
    assert (data[:, :, torch.as_tensor(held, device="cuda")] == 0xF0).all()
    torch.cuda.synchronize()
    frames = [data.data_ptr() + i * fb for i in range(n)]
    with chx.Serve(frame_slots=64, slot_bytes=fb) as sv:
        sid = sv.open(N, F, T)
        sat = D.saturated(sv, sid, frames, device=True)
        say(f"  device frames, saturated: {sat}")
        got = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device="cuda")
        sv.read_device(sid, got.data_ptr())
        sv.close_stream(sid)
        # Poisson arrivals at half the sustainable rate, on a fresh stream
        sid = sv.open(N, F, T)
        half = D.run_arrivals(sv, sid, frames[:256],
                              D.poisson_arrivals(256, 0.5 * sat["achieved_rate"], 1), True)
        say(f"  device frames, Poisson 0.5x: {half}")
        assert half["p99_ms"] < 50 * sat["seconds"] * 1e3 / n + 50
        sv.close_stream(sid)
        # host frames: a pinned pool of 32 frames cycled 4 times
        pool = torch.empty((32, T, F, N), dtype=torch.uint8, pin_memory=True)
        pool.copy_(data[:32 * T].view(32, T, F, N))
        sid = sv.open(N, F, T)
        host = D.saturated(sv, sid, [pool[i % 32].data_ptr() for i in range(128)], False)
        say(f"  host frames, saturated: {host}")
        host_vis = sv.read(sid)
    with chx.Plan(N, F, n * T, 256, 1, chx.INT8_MFMA) as p:
        want = torch.empty_like(got)
        p.correlate_device(data.data_ptr(), want.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(got, want)
    with chx.Plan(N, F, 32 * T, 256, 1, chx.INT8_MFMA) as p:
        want32 = torch.empty_like(got)
        p.correlate_device(data.data_ptr(), want32.data_ptr())
        torch.cuda.synchronize()
    # linearity: 4 passes over the pool = 4 x the pool's visibilities (mod 2^32)
    np.testing.assert_array_equal(host_vis, (want32.cpu().numpy().astype(np.int64) * 4)
                                  .astype(np.uint32).view(np.int32))
    # saturation of the held elements' pairs: (7,-8) x (7,-8) -> re = 113 per sample
    hb, hl = held[0] // 32, held[0] % 32
    NBn, nb = chx.num_blocks(N), N // 32
    b = hb * nb - hb * (hb - 1) // 2
    assert int(got[0, b, hl, hl, 0]) == 113 * n * T


def test_close_and_read_after_a_failed_launch():
    """After a launch fails (forced on the host: nothing reaches the GPU),
    the engine's error is sticky; read and close of streams with frames
    still queued wait until the worker has retired them, then return the
    error -- and the accumulators they free are named by no queued frame."""
    torch = _torch()
    N, F, T = 256, 2, 1024
    frames = [torch.randint(0, 256, (T, F, N), dtype=torch.uint8, device="cuda")
              for _ in range(12)]
    torch.cuda.synchronize()
    with chx.Serve(frame_slots=12, slot_bytes=N * F * T, max_batch_frames=1,
                   method=chx.FP6_MFMA) as sv:
        a, b = sv.open(N, F, T, 1), sv.open(N, F, T, 0)
        assert chx.lib().chx_serve_debug_fail_next_launch(sv._h) == 0
        submitted = 0
        for i, fr in enumerate(frames):
            try:
                sv.submit_device(a if i % 2 == 0 else b, fr.data_ptr())
                submitted += 1
            except chx.ChxError as e:  # the error is sticky once the worker hit it
                assert e.status == chx.ERR_HIP
                break
        say(f"  submitted {submitted} frames before the error showed")
        with pytest.raises(chx.ChxError) as e:
            sv.close_stream(a)
        assert e.value.status == chx.ERR_HIP
        with pytest.raises(chx.ChxError) as e:
            sv.read(b)
        assert e.value.status == chx.ERR_HIP
        with pytest.raises(chx.ChxError) as e:
            sv.close_stream(b)
        assert e.value.status == chx.ERR_HIP
        st = chx.ServeStats()
        assert chx.lib().chx_serve_get_stats(sv._h, ctypes.byref(st)) == chx.ERR_HIP
        assert st.frames_submitted == submitted and st.frames_done == 0
        assert st.streams_open == 0
