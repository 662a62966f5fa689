"""chx_correlate_noalloc: the plan-free, allocation-free correlate entry
point of the packed-integer methods, through the C ABI.

Its output must equal chx_correlate's (a Plan) bit for bit, on consecutive
calls that share the one per-GPU scratch array (the second call smaller, so
stale maps or sums of the first would show), and the GPU's free memory must
not change across calls.  (Its host-side refusals are CPU tests in
test_abi.py.)
"""
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
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


CASES = [  # (T, F, N, ta): the second call is smaller than the first
    (2048, 4, 512, 256),
    (288, 2, 96, 96),
]


def _plan_output(torch, d_in, T, F, N, ta, U, method):
    with chx.Plan(N, F, T, ta, U, method) as p:
        out = torch.full(p.output_shape, 11, dtype=torch.int32, device="cuda")
        p.correlate_device(d_in.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("method", [chx.CONFERENCE, chx.HIGHLY_PACKED])
@pytest.mark.parametrize("U", [1, 0])
def test_noalloc_equals_plan_on_consecutive_calls(method, U):
    torch = _torch()
    for k, (T, F, N, ta) in enumerate(CASES):
        say(f"  noalloc method={chx.METHOD_NAMES[method]} U={U} T={T} F={F} N={N} ta={ta}")
        data = np.random.default_rng(300 + k).integers(0, 256, size=(T, F, N), dtype=np.uint8)
        d_in = torch.from_numpy(data).cuda()
        want = _plan_output(torch, d_in, T, F, N, ta, U, method)
        got = torch.full_like(want, -3)
        chx.correlate_noalloc(N, F, T, d_in.data_ptr(), got.data_ptr(), ta, U, method,
                              stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"call {k}"
        if k == len(CASES) - 1:
            np.testing.assert_array_equal(got.cpu().numpy(), O.correlate_blocks(data, U))


def test_noalloc_leaves_free_memory_unchanged():
    torch = _torch()
    T, F, N, ta = 1024, 2, 256, 256
    data = np.random.default_rng(9).integers(0, 256, size=(T, F, N), dtype=np.uint8)
    d_in = torch.from_numpy(data).cuda()
    out = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device="cuda")
    chx.correlate_noalloc(N, F, T, d_in.data_ptr(), out.data_ptr(), ta)  # loads the code
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        chx.correlate_noalloc(N, F, T, d_in.data_ptr(), out.data_ptr(), ta)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    np.testing.assert_array_equal(out.cpu().numpy(), O.correlate_blocks(data, 1))


def test_noalloc_in_a_captured_graph_replays():
    """Captured into a graph (its scratch is rewritten by kernels, no memset
    nodes), replayed several times with eager work and host synchronisation
    between the replays, each replay on new input equals chx_correlate."""
    torch = _torch()
    T, F, N, ta = 1024, 2, 192, 256
    rng = np.random.default_rng(77)
    d_in = torch.empty((T, F, N), dtype=torch.uint8, device="cuda")
    out = torch.empty((F, chx.num_blocks(N), 32, 32, 2), dtype=torch.int32, device="cuda")
    chx.correlate_noalloc(N, F, T, d_in.data_ptr(), out.data_ptr(), ta, 1, chx.HIGHLY_PACKED)
    torch.cuda.synchronize()  # resolves the scratch address before capture
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            chx.correlate_noalloc(N, F, T, d_in.data_ptr(), out.data_ptr(), ta, 1,
                                  chx.HIGHLY_PACKED, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(4):
        data = rng.integers(0, 256, size=(T, F, N), dtype=np.uint8)
        d_in.copy_(torch.from_numpy(data))
        # eager work of the program between replays (another shape on the
        # same scratch, then torch's own kernels)
        other = torch.empty((1, chx.num_blocks(64), 32, 32, 2), dtype=torch.int32,
                            device="cuda")
        x64 = torch.randint(0, 256, (256, 1, 64), dtype=torch.uint8, device="cuda")
        chx.correlate_noalloc(64, 1, 256, x64.data_ptr(), other.data_ptr(), 64)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        want = _plan_output(torch, d_in, T, F, N, ta, 1, chx.HIGHLY_PACKED)
        assert torch.equal(out, want), f"replay {rep}"
