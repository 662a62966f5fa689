"""Bit-exact parity of the HIP paths against the oracle, through the C ABI.

Every method (conference, highly packed, int8 MFMA) x both conventions is
checked on: the reference's own CPU-correlator fixtures (config 1: N=256,
F=1, T=256), seeded random inputs of ragged shapes against the oracle's full
block buffer (diagonal lower halves included), the stage operations one by
one, the worst-case lane data at each method's time_accum cap, and
closed forms (saturation, conjugation, linearity in T).
"""
import sys

import numpy as np
import pytest

import chx
import oracle as O

pytestmark = pytest.mark.gpu

METHODS = [chx.CONFERENCE, chx.HIGHLY_PACKED, chx.INT8_MFMA, chx.FP6_MFMA]
MATRIX = [chx.INT8_MFMA, chx.FP6_MFMA]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(data, method, U, ta):
    T, F, N = data.shape
    say(f"  case method={chx.METHOD_NAMES[method]} U={U} T={T} F={F} N={N} ta={ta}")
    with chx.Plan(N, F, T, ta, U, method) as p:
        return p.correlate(data)


def rand(T, F, N, seed):
    return np.random.default_rng(seed).integers(0, 256, size=(T, F, N), dtype=np.uint8)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("U", [1, 0])
@pytest.mark.parametrize("case", ["c1_repeat", "c1_continuous"])
def test_config1_matches_reference_cpu_correlator(golden, method, U, case):
    data = golden[f"{case}__input"]
    blocks = run(data, method, U, 256)
    tri = chx.blocks_to_triangle(blocks, 256)
    np.testing.assert_array_equal(tri, golden[f"{case}__tri_U{U}"])


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("U", [1, 0])
def test_reference_fixture_shapes(golden, method, U):
    for case in ("rand_cont_s7", "single_freq_1"):
        data = golden[f"{case}__input"]
        T = data.shape[0]
        blocks = run(data, method, U, 32 if T % 64 else 64)
        np.testing.assert_array_equal(chx.blocks_to_triangle(blocks, data.shape[2]),
                                      golden[f"{case}__tri_U{U}"], err_msg=case)


SHAPES = [  # (T, F, N, ta)
    (32, 1, 32, 8),
    (64, 2, 64, 32),
    (256, 3, 96, 128),
    (288, 1, 160, 288),
    (1024, 2, 128, 256),
    (2304, 1, 256, 1152),
    (512, 1, 384, 64),
]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("U", [1, 0])
def test_random_shapes_against_oracle(method, U):
    for i, (T, F, N, ta) in enumerate(SHAPES):
        if method == chx.CONFERENCE and ta > 288:
            ta = 288 if T % 288 == 0 else 256
            if T % ta:
                ta = 64
        data = rand(T, F, N, 1000 + i)
        np.testing.assert_array_equal(run(data, method, U, ta), O.correlate_blocks(data, U),
                                      err_msg=f"shape {T, F, N, ta}")


@pytest.mark.parametrize("method,ta", [(chx.CONFERENCE, 288), (chx.HIGHLY_PACKED, 2272),
                                       (chx.INT8_MFMA, 2272), (chx.FP6_MFMA, 2272)])
def test_worst_case_lanes_at_cap(method, ta):
    T = ta * 2
    data = np.full((T, 1, 64), 0xFF, np.uint8)
    data[:, :, 1::2] = 0xF0
    data[T // 2:, :, ::3] = 0x0F
    for U in (1, 0):
        np.testing.assert_array_equal(run(data, method, U, ta), O.correlate_blocks(data, U))


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


@pytest.mark.parametrize("method", [chx.CONFERENCE, chx.HIGHLY_PACKED])
@pytest.mark.parametrize("U", [1, 0])
def test_stage_operations(method, U):
    torch = _torch()
    T, F, N, ta = 512, 2, 96, 128
    data = rand(T, F, N, 77)
    say(f"  stages method={method} U={U}")
    with chx.Plan(N, F, T, ta, U, method) as p:
        d_in = torch.from_numpy(data).cuda()
        d_sums = torch.zeros((F, N, 2), dtype=torch.int32, device="cuda")
        d_vis = torch.zeros(p.output_shape, dtype=torch.int32, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        p.offset_accumulate(d_in.data_ptr(), d_sums.data_ptr(), s)
        torch.cuda.synchronize()
        sums = d_sums.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(sums, O.offset_sums(data))
        p.preseed(d_sums.data_ptr(), d_vis.data_ptr(), s)
        torch.cuda.synchronize()
        pre = O.preseed(sums, T, method, U)
        np.testing.assert_array_equal(d_vis.cpu().numpy(), pre)
        p.corr(d_in.data_ptr(), d_vis.data_ptr(), s)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_vis.cpu().numpy(), O.correlate_blocks(data, U))
        # raw products alone == the packed-lane emulation
        d_vis.zero_()
        p.corr(d_in.data_ptr(), d_vis.data_ptr(), s)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_vis.cpu().numpy(), O.packed_raw(data, method, U, ta))


@pytest.mark.parametrize("method", MATRIX)
def test_mfma_corr_adds_and_device_triangle(method):
    torch = _torch()
    T, F, N = 256, 2, 128
    data = rand(T, F, N, 5)
    with chx.Plan(N, F, T, 256, 1, method) as p:
        d_in = torch.from_numpy(data).cuda()
        d_vis = torch.full(p.output_shape, 7, dtype=torch.int32, device="cuda")
        p.corr(d_in.data_ptr(), d_vis.data_ptr())
        torch.cuda.synchronize()
        want = O.correlate_blocks(data, 1)
        np.testing.assert_array_equal(d_vis.cpu().numpy(), want + 7)
        p.correlate_device(d_in.data_ptr(), d_vis.data_ptr())
        d_tri = torch.empty((F, N * (N + 1) // 2, 2), dtype=torch.int32, device="cuda")
        chx.blocks_to_triangle_device(F, N, d_vis.data_ptr(), d_tri.data_ptr())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_tri.cpu().numpy(), O.blocks_to_triangle(want, N))


@pytest.mark.parametrize("method", MATRIX)
def test_mfma_long_integration_saturation_closed_form(method):
    # Every sample (-8,-8): re = 128*T, im = 0, over several accumulator
    # flush chunks (the largest accumulator values either path holds).
    T, F, N = 131072, 2, 128
    data = np.zeros((T, F, N), np.uint8)
    blocks = run(data, method, 1, 256)
    assert (blocks[..., 0] == 128 * T).all() and (blocks[..., 1] == 0).all()


@pytest.mark.parametrize("method", MATRIX)
@pytest.mark.parametrize("U", [1, 0])
def test_mfma_every_code_pair_over_a_full_item(method, U):
    # Element e holds byte e at every sample, so the 256 elements cover
    # every (R, I) code pair, including (7,7) and (-8,-8), and every pair
    # of elements meets in one block; over 131072 samples (one work item of
    # the most stages) the int8 path's offset-binary code products reach
    # their largest sums.  Constant data: the result is T/32 times the
    # oracle's over 32 samples.
    T, F, N = 131072, 1, 256
    row = np.arange(N, dtype=np.uint8)
    data = np.broadcast_to(row, (T, F, N)).copy()
    want = O.correlate_blocks(np.ascontiguousarray(data[:32]), U).astype(np.int64) * (T // 32)
    got = run(data, method, U, 256)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("method", MATRIX)
def test_mfma_time_split_and_linearity(method):
    # Small N with long T takes the split-time (atomic) path.
    T, F, N = 16384, 1, 64
    data = rand(T, F, N, 9)
    full = run(data, method, 1, 256)
    a = run(np.ascontiguousarray(data[:8192]), method, 1, 256)
    b = run(np.ascontiguousarray(data[8192:]), method, 1, 256)
    np.testing.assert_array_equal(full, a + b)
    np.testing.assert_array_equal(full, O.correlate_blocks(data, 1))
    conj = run(data, method, 0, 256)
    np.testing.assert_array_equal(conj[..., 0], full[..., 0])
    np.testing.assert_array_equal(conj[..., 1], -full[..., 1])


def test_fp6_uneven_time_pieces():
    # 2049 stages of 64 samples plus a ragged tail: FP6 work items hold at
    # most 2048 stages, so the time is cut into pieces with a shorter last
    # one, added atomically.
    T, F, N = 2049 * 64 + 96, 1, 64  # plans take T % 32 == 0
    data = rand(T, F, N, 2049)
    got = run(data, chx.FP6_MFMA, 1, 32)
    np.testing.assert_array_equal(got, run(data, chx.INT8_MFMA, 1, 32))
    np.testing.assert_array_equal(got, O.correlate_blocks(data, 1))


@pytest.mark.parametrize("method", MATRIX)
def test_half_tile_pairs_against_oracle(method):
    # Arrays whose row tiles end in half tiles: N=512 (4 tiles: one pair),
    # 640 (5 tiles: a pair and a lone half tile), 1152 (9 tiles: two pairs
    # and a lone one), 1376 (11 tiles, N not a multiple of 128: three
    # pairs), 2944 (23 tiles: full tiles in three bands of row tiles, the
    # last one of 7, and six pairs), checked against the oracle for both
    # conventions.
    for N, seed in ((512, 51), (640, 64), (1152, 11), (1376, 13), (2944, 29)):
        data = rand(256, 2, N, seed)
        for U in (1, 0):
            np.testing.assert_array_equal(run(data, method, U, 32), O.correlate_blocks(data, U),
                                          err_msg=f"N={N} U={U}")


def test_methods_agree_on_array_scale_channel_slice():
    # N=2048 (config 2's array) on a T slice: all three methods bit-exact
    # against each other and against the oracle.
    T, F, N = 512, 2, 2048
    data = rand(T, F, N, 2048)
    want = O.correlate_blocks(data, 1)
    for method, ta in ((chx.INT8_MFMA, 512), (chx.FP6_MFMA, 512), (chx.CONFERENCE, 256),
                       (chx.HIGHLY_PACKED, 512)):
        np.testing.assert_array_equal(run(data, method, 1, ta), want,
                                      err_msg=chx.METHOD_NAMES[method])


@pytest.mark.parametrize("method", MATRIX)
def test_matrix_long_small_array_multi_chunk(method):
    # A small array over 2^21 + 96 samples: several plane chunks (each capped
    # so the unpack grid stays within its limits) and a ragged last chunk.
    T, F, N = (1 << 21) + 96, 1, 32
    data = rand(T, F, N, 21)
    np.testing.assert_array_equal(run(data, method, 1, 32), O.correlate_blocks(data, 1))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("U", [1, 0])
def test_paired_16_feed_run_against_oracle(method, U):
    # A 16-feed array at F frequencies runs as 32 elements at F/2 (the bytes
    # are read as they lie); the split and the 16-element triangles give
    # each frequency's own 16-feed correlation.
    T, F = 256, 8
    data = rand(T, F, 16, 160 + U)
    blocks = run(data.reshape(T, F // 2, 32), method, U, 64)
    tri = chx.feed16_to_triangle(chx.split_paired_feeds(chx.blocks_to_full(blocks, 32), F, 16))
    np.testing.assert_array_equal(tri, O.correlate_triangle(data, U))


@pytest.mark.parametrize("T,F,N", [(4192, 3, 416), (65632, 1, 32), (2048, 16, 2048)])
def test_offset_sums_ragged_columns_and_slices(T, F, N):
    # column groups of 64 x 16 bytes with a partial last one (F*N/16 = 78),
    # many slices with a ragged last one (T = 65536 + 96), and a full row of
    # the array-scale shape
    torch = _torch()
    data = rand(T, F, N, 5)
    say(f"  offset sums T={T} F={F} N={N}")
    with chx.Plan(N, F, T, 32, 1, chx.CONFERENCE) as p:
        d_in = torch.from_numpy(data).cuda()
        d_sums = torch.zeros((F, N, 2), dtype=torch.int32, device="cuda")
        p.offset_accumulate(d_in.data_ptr(), d_sums.data_ptr(),
                            torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_sums.cpu().numpy().view(np.uint32), O.offset_sums(data))


def test_full_matrix_expansion_on_device(golden):
    torch = _torch()
    for N, F in ((96, 2), (64, 1)):
        nb = chx.num_blocks(N)
        blocks = (np.arange(F * nb * 2048, dtype=np.int64) * 37 - 100000).astype(np.int32)
        d_b = torch.from_numpy(blocks).cuda()
        d_full = torch.empty((F, N, N, 2), dtype=torch.int32, device="cuda")
        chx.blocks_to_full_device(F, N, d_b.data_ptr(), d_full.data_ptr())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_full.cpu().numpy(), golden[f"reorgfull_N{N}_F{F}__full"])
    # a real correlation at array scale: device expansion == oracle expansion
    T, F, N = 256, 2, 2048
    data = rand(T, F, N, 4)
    with chx.Plan(N, F, T, 256, 1, chx.FP6_MFMA) as p:
        d_in = torch.from_numpy(data).cuda()
        d_vis = torch.empty(p.output_shape, dtype=torch.int32, device="cuda")
        p.correlate_device(d_in.data_ptr(), d_vis.data_ptr())
        d_full = torch.empty((F, N, N, 2), dtype=torch.int32, device="cuda")
        chx.blocks_to_full_device(F, N, d_vis.data_ptr(), d_full.data_ptr())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_full.cpu().numpy(),
                                      O.blocks_to_full(d_vis.cpu().numpy(), N))
        # Hermitian: full[x][y] == conj(full[y][x])
        full = d_full.cpu().numpy()
        np.testing.assert_array_equal(full[..., 0], full[..., 0].transpose(0, 2, 1))
        np.testing.assert_array_equal(full[..., 1], -full[..., 1].transpose(0, 2, 1))


# ---- Config 2 (N=2048, F=16) at the launch shapes the headline runs ----

CFG2_T, CFG2_F, CFG2_N = 65536, 16, 2048


def _launch_shape(F, N, T, cus):
    """The matrix-core plan launch for one chunk (both methods): whole rounds
    of one workgroup per CU take whole work items; the items of the last
    partial round are cut in time into `split` pieces that add atomically
    into output zeroed beforehand (the channels from head // items-per-channel
    on)."""
    nt = (N + 127) // 128
    m = nt // 2
    wide = (m + 1) * (m + 1) if nt % 2 else m * (m + 1)
    per_f = wide - (nt + 1) // 2 // 2
    items = F * per_f
    nst = -(-T // 256) * 4
    head = items // cus * cus if nst <= 2048 else 0
    tail = items - head
    split = 1
    if tail:
        best = -(-tail // cus)
        for sp in range(2, 65):
            if nst // sp < 16:
                break
            r = -(-(tail * sp) // cus) / sp
            if r < best - 1e-9:
                best, split = r, sp
    sps = min(-(-(-(-nst // split)) // 4) * 4, 2048)
    split = -(-nst // sps)
    return per_f, head, tail, split


def _oracle_blocks_by_channel(data, U):
    """The oracle's block buffer, one channel per thread (channels are
    independent in the reference loop, cpu_corr_test.c:176-198)."""
    from concurrent.futures import ThreadPoolExecutor

    T, F, N = data.shape
    out = np.empty((F, chx.num_blocks(N), 32, 32, 2), np.int32)

    def one(f):
        out[f] = O.correlate_blocks(np.ascontiguousarray(data[:, f:f + 1, :]), U)[0]

    with ThreadPoolExecutor(max_workers=16) as ex:
        list(ex.map(one, range(F)))
    return out


@pytest.fixture(scope="module")
def cfg2_slice_oracle():
    T = 2048
    data = rand(T, CFG2_F, CFG2_N, 22)
    say(f"  oracle T={T} F={CFG2_F} N={CFG2_N} (threads over channels)")
    return data, _oracle_blocks_by_channel(data, 1)


@pytest.mark.parametrize("method", MATRIX)
def test_config2_head_and_tail_pieces_against_oracle(method, cfg2_slice_oracle):
    # N=2048, F=16, T=2048 takes the headline launch: whole items with plain
    # stores for the whole rounds, then the last partial round's items cut
    # into time pieces that add atomically after a zero fill of the last
    # channel; checked against the oracle's full block buffer.
    torch = _torch()
    data, want = cfg2_slice_oracle
    T = data.shape[0]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    per_f, head, tail, split = _launch_shape(CFG2_F, CFG2_N, T, cus)
    say(f"  launch: {cus} CUs, {per_f} items per channel, head {head}, tail {tail}, "
        f"split {split}, zero fill from channel {head // per_f}")
    assert head > 0 and tail > 0 and split > 1, "shape must take whole items and tail pieces"
    assert head // per_f < CFG2_F
    # output starts non-zero, so the fill and the plain stores both matter
    with chx.Plan(CFG2_N, CFG2_F, T, 256, 1, method) as p:
        d_in = torch.from_numpy(data).cuda()
        d_vis = torch.full(p.output_shape, -5, dtype=torch.int32, device="cuda")
        p.correlate_device(d_in.data_ptr(), d_vis.data_ptr())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_vis.cpu().numpy(), want)


def test_config2_full_size_all_methods():
    # Config 2 at full size (2 GiB of input, 16 x 2,098,176 visibilities):
    # all four methods bit-equal, the conjugate convention the conjugate, and
    # every channel equal to the oracle on a sorted 16-element subset over
    # all 65,536 samples (elements of the first and the last tile included).
    torch = _torch()
    T, F, N = CFG2_T, CFG2_F, CFG2_N
    x = torch.empty((T, F, N), dtype=torch.uint8, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    torch.randint(0, 256, x.shape, dtype=torch.uint8, device="cuda", generator=g, out=x)
    outs = {}
    for method, ta in ((chx.FP6_MFMA, 256), (chx.INT8_MFMA, 256), (chx.CONFERENCE, 256),
                       (chx.HIGHLY_PACKED, 2048)):
        say(f"  config 2 full size: {chx.METHOD_NAMES[method]}")
        with chx.Plan(N, F, T, ta, 1, method) as p:
            out = torch.empty(p.output_shape, dtype=torch.int32, device="cuda")
            p.correlate_device(x.data_ptr(), out.data_ptr())
            torch.cuda.synchronize()
        outs[method] = out
    ref = outs[chx.FP6_MFMA]
    for method, out in outs.items():
        assert torch.equal(out, ref), chx.METHOD_NAMES[method]
    with chx.Plan(N, F, T, 256, 0, chx.INT8_MFMA) as p:
        conj = torch.empty_like(ref)
        p.correlate_device(x.data_ptr(), conj.data_ptr())
        torch.cuda.synchronize()
    assert torch.equal(conj[..., 0], ref[..., 0])
    assert torch.equal(conj[..., 1], -ref[..., 1])
    del conj, outs
    rng = np.random.default_rng(65536)
    fixed = [0, 37, 1930, 2047]
    rest = rng.choice(np.setdiff1d(np.arange(N), fixed), size=12, replace=False)
    E = np.sort(np.concatenate([fixed, rest]))
    sub = x[:, :, torch.as_tensor(E, device="cuda")].contiguous().cpu().numpy()  # [T][F][16]
    want = O.correlate_triangle(sub, 1)  # [F][136][2]
    ys, xs = np.triu_indices(16)
    Y, X = E[ys], E[xs]
    nb = N // 32
    by, bx = Y // 32, X // 32
    b = by * nb - by * (by - 1) // 2 + (bx - by)
    got = ref[:, torch.as_tensor(b, device="cuda"), torch.as_tensor(Y % 32, device="cuda"),
              torch.as_tensor(X % 32, device="cuda")].cpu().numpy()
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("method", METHODS)
def test_plan_correlate_in_a_captured_graph(method):
    # chx_correlate captured into a graph (fills are kernels, not memset
    # nodes), replayed several times on new input with eager work and host
    # synchronisation between the replays: every replay equals the eager
    # result and the oracle.
    torch = _torch()
    T, F, N = 2048, 2, 256
    rng = np.random.default_rng(31)
    with chx.Plan(N, F, T, 256, 1, method) as p:
        d_in = torch.empty((T, F, N), dtype=torch.uint8, device="cuda")
        out = torch.empty(p.output_shape, dtype=torch.int32, device="cuda")
        eager = torch.empty_like(out)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                p.correlate_device(d_in.data_ptr(), out.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        for rep in range(3):
            data = rng.integers(0, 256, size=(T, F, N), dtype=np.uint8)
            d_in.copy_(torch.from_numpy(data))
            p.correlate_device(d_in.data_ptr(), eager.data_ptr())
            torch.cuda.synchronize()
            out.fill_(-1)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {rep}"
            np.testing.assert_array_equal(out.cpu().numpy(), O.correlate_blocks(data, 1))
