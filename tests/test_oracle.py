"""The CPU oracle, pinned before it is trusted.

Pins, per oracle function:
  chxo_generate              reference generator fixtures (every mode), glibc rand()
  chxo_correlate_triangle    reference CPU correlator fixtures, both conventions
  chxo_generate_and_correlate_triangle  same fixtures, reference loop order
  chxo_correlate_blocks      fixtures via the reorg, plus the lower halves of
                             diagonal blocks against the reference full matrix
  chxo_blocks_to_triangle    reference reorg fixtures
  chxo_compare_triangle      reference comparator on planted differences
  chxo_offset_sums           closed form (numpy bincount-free sum) on random data
  chxo_preseed+packed_raw    sum equals the correlator exactly (App. A identities),
                             and lane overflow past the cap is visible
  all                        closed forms: saturation, constant, conjugation,
                             linearity in T, and a pure-Python brute force
"""
import ctypes

import numpy as np
import pytest

import oracle as O


def params(golden, name):
    T, F, N, g, seed, dre, dim, ire, iim, sf, nr = (int(v) for v in golden[f"{name}__params"])
    return dict(T=T, F=F, N=N, gen_type=g, seed=seed, default_real=dre, default_imag=dim,
                initial_real=ire, initial_imag=iim, single_freq=sf, no_repeat=bool(nr))


def test_generator_matches_reference(golden, golden_cases):
    for name in golden_cases:
        kw = params(golden, name)
        T, F, N = kw.pop("T"), kw.pop("F"), kw.pop("N")
        data = O.generate(T, F, N, **kw)
        np.testing.assert_array_equal(data, golden[f"{name}__input"], err_msg=name)


def test_glibc_rand_stream_is_the_generator_source(golden):
    # c1 continuous: byte k of the stream is (rand()%16 << 4) | rand()%16
    libc = ctypes.CDLL(None)
    libc.srand(42)
    vals = [libc.rand() % 16 for _ in range(512)]
    expect = np.array([(vals[2 * k] << 4) | vals[2 * k + 1] for k in range(256)], np.uint8)
    np.testing.assert_array_equal(golden["c1_continuous__input"][0, 0], expect)


@pytest.mark.parametrize("U", [0, 1])
def test_triangle_matches_reference(golden, golden_cases, U):
    for name in golden_cases:
        tri = O.correlate_triangle(golden[f"{name}__input"], U)
        np.testing.assert_array_equal(tri, golden[f"{name}__tri_U{U}"], err_msg=name)


@pytest.mark.parametrize("U", [0, 1])
def test_reference_loop_order_restatement(golden, golden_cases, U):
    for name in golden_cases:
        kw = params(golden, name)
        if kw["T"] * kw["N"] * kw["N"] > 2e7:
            continue
        tri = O.generate_and_correlate_triangle(U=U, **kw)
        np.testing.assert_array_equal(tri, golden[f"{name}__tri_U{U}"], err_msg=name)


@pytest.mark.parametrize("U", [0, 1])
def test_blocks_match_reference(golden, golden_cases, U):
    for name in golden_cases:
        data = golden[f"{name}__input"]
        N = data.shape[2]
        blocks = O.correlate_blocks(data, U)
        np.testing.assert_array_equal(O.blocks_to_triangle(blocks, N),
                                      golden[f"{name}__tri_U{U}"], err_msg=name)
        key = f"{name}__full_U{U}"
        if key in golden:
            full = golden[key]
            nb = N // 32
            b = 0
            for by in range(nb):
                for bx in range(by, nb):
                    ref = full[:, by * 32:by * 32 + 32, bx * 32:bx * 32 + 32, :]
                    np.testing.assert_array_equal(blocks[:, b], ref, err_msg=f"{name} b{b}")
                    b += 1


def test_reorg_matches_reference(golden):
    for N, F in ((96, 2), (32, 1), (256, 1)):
        nb = O.num_blocks(N)
        blocks = np.arange(F * nb * 2048, dtype=np.int32)
        np.testing.assert_array_equal(O.blocks_to_triangle(blocks, N),
                                      golden[f"reorg_N{N}_F{F}__tri"])


def test_block_maps():
    x, y = O.block_maps(256)
    assert len(x) == 36
    pairs = list(zip(y.tolist(), x.tolist()))
    assert pairs == [(by, bx) for by in range(8) for bx in range(by, 8)]


def test_compare_matches_reference(golden):
    n, e2, _ = O.compare_triangle(golden["compare__gpu"], golden["compare__cpu"], 64)
    assert [n, e2] == golden["compare__result"].tolist()


def test_offset_sums_closed_form():
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, size=(96, 3, 64), dtype=np.uint8)
    s = O.offset_sums(data)
    np.testing.assert_array_equal(s[..., 0], (data >> 4).astype(np.int64).sum(0))
    np.testing.assert_array_equal(s[..., 1], (data & 15).astype(np.int64).sum(0))


@pytest.mark.parametrize("method,ta", [(0, 8), (0, 64), (0, 288), (1, 8), (1, 120),
                                       (1, 256), (1, 1024), (1, 2272)])
@pytest.mark.parametrize("U", [0, 1])
def test_packed_method_is_exact(method, ta, U):
    rng = np.random.default_rng(100 + ta + method)
    T = ta * 2 if ta < 1024 else ta
    data = rng.integers(0, 256, size=(T, 2, 64), dtype=np.uint8)
    got = O.packed_pipeline(data, method, U, ta)
    np.testing.assert_array_equal(got, O.correlate_blocks(data, U))


@pytest.mark.parametrize("method,ta", [(0, 288), (1, 2272)])
def test_packed_method_worst_case_lanes(method, ta):
    # All-0xFF on X with 0xF0 on Y maximises the conference lanes (225/sample)
    # and the HP a0-hi lane (450/sample).
    data = np.full((ta, 1, 64), 0xFF, np.uint8)
    data[:, :, 1::2] = 0xF0
    for U in (0, 1):
        np.testing.assert_array_equal(O.packed_pipeline(data, method, U, ta),
                                      O.correlate_blocks(data, U))


def test_packed_method_overflows_past_cap():
    # Conference lanes hold 16 bits: 225 * 296 > 65535, so 0xFF data over a
    # 296-sample chunk wraps and the result is wrong.  This is the cap the
    # product enforces.
    data = np.full((296, 1, 32), 0xFF, np.uint8)
    got = O.packed_pipeline(data, 0, 1, 296)
    assert not np.array_equal(got, O.correlate_blocks(data, 1))


def test_saturation_closed_form():
    T = 72
    data = np.zeros((T, 2, 64), np.uint8)  # every sample (-8, -8)
    tri = O.correlate_triangle(data, 1)
    assert (tri[..., 0] == 128 * T).all() and (tri[..., 1] == 0).all()


def test_constant_closed_form():
    T, r, i = 40, -3, 5
    data = O.generate(T, 2, 32, gen_type=O.GEN_CONSTANT, initial_real=r, initial_imag=i)
    tri = O.correlate_triangle(data, 1)
    assert (tri[..., 0] == T * (r * r + i * i)).all() and (tri[..., 1] == 0).all()


def test_conjugation_and_linearity():
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, size=(48, 2, 64), dtype=np.uint8)
    u1 = O.correlate_blocks(data, 1)
    u0 = O.correlate_blocks(data, 0)
    np.testing.assert_array_equal(u0[..., 0], u1[..., 0])
    np.testing.assert_array_equal(u0[..., 1], -u1[..., 1])
    parts = O.correlate_blocks(data[:20], 1) + O.correlate_blocks(data[20:], 1)
    np.testing.assert_array_equal(parts, u1)
    tri = O.correlate_triangle(data, 1)
    N = 64
    diag = [y * N - (y - 1) * y // 2 for y in range(N)]
    assert (tri[:, diag, 1] == 0).all()


def test_brute_force_python():
    rng = np.random.default_rng(9)
    T, F, N = 5, 2, 32
    data = rng.integers(0, 256, size=(T, F, N), dtype=np.uint8)
    want = np.zeros((F, N * (N + 1) // 2, 2), np.int64)
    for f in range(F):
        k = 0
        for y in range(N):
            for x in range(y, N):
                for t in range(T):
                    yb, xb = int(data[t, f, y]), int(data[t, f, x])
                    yc = complex((yb >> 4) - 8, (yb & 15) - 8)
                    xc = complex((xb >> 4) - 8, (xb & 15) - 8)
                    v = yc * xc.conjugate()
                    want[f, k, 0] += int(v.real)
                    want[f, k, 1] += int(v.imag)
                k += 1
    np.testing.assert_array_equal(O.correlate_triangle(data, 1), want)


def test_full_reorg_matches_reference(golden):
    """chxo_blocks_to_full restates reorganize_GPU_to_full_Matrix_for_comparison
    (gpu_data_reorg.c:36-74): equal to the reference's output on signed
    index-valued block buffers."""
    for N, F in ((96, 2), (64, 1)):
        nb = O.num_blocks(N)
        blocks = (np.arange(F * nb * 2048, dtype=np.int64) * 37 - 100000).astype(np.int32)
        np.testing.assert_array_equal(O.blocks_to_full(blocks, N),
                                      golden[f"reorgfull_N{N}_F{F}__full"])


def test_full_expansion_of_blocks_equals_reference_full_correlator(golden, golden_cases):
    """Blocks -> Hermitian expansion equals the reference's full N x N CPU
    correlator (cpu_corr_test.c:108-154 and :11-57) in both conventions."""
    done = 0
    for case in golden_cases:
        if f"{case}__full_U1" not in golden:
            continue
        data = golden[f"{case}__input"]
        N = data.shape[2]
        for U in (1, 0):
            np.testing.assert_array_equal(O.blocks_to_full(O.correlate_blocks(data, U), N),
                                          golden[f"{case}__full_U{U}"], err_msg=f"{case} U={U}")
        done += 1
    assert done >= 5


SPLIT_CASES = ((4, 16), (5, 16), (6, 7), (1, 16))   # (F, E)
TRI16_CASES = ((4, 4), (6, 3), (2, 5), (3, 0))      # (F_out, F_in)


def test_split_paired_feeds_matches_reference(golden):
    # whole buffer, in place: the tail past the output keeps its input values
    for F, E in SPLIT_CASES:
        buf = golden[f"split_F{F}_E{E}__in"].copy()
        O.lib().chxo_split_paired_feeds(F, E, O._p(buf))
        np.testing.assert_array_equal(buf, golden[f"split_F{F}_E{E}__out"], err_msg=f"F{F} E{E}")


def test_feed16_to_triangle_matches_reference(golden):
    for F_out, F_in in TRI16_CASES:
        src = golden[f"tri16_out{F_out}_in{F_in}__in"]
        out = np.full(F_out * 272, 0x5A5A5A5A, np.int32)
        O.lib().chxo_feed16_to_triangle(F_out, F_in, O._p(src), O._p(out))
        np.testing.assert_array_equal(out, golden[f"tri16_out{F_out}_in{F_in}__out"])


@pytest.mark.parametrize("U", [1, 0])
def test_paired_feed_run_is_the_16_feed_correlation(U):
    # Bytes [T][F][16] read as [T][F/2][32] hold frequencies 2g and 2g+1 side
    # by side; the diagonal quadrants of the 32-element correlation are the
    # 16-element correlations of those two frequencies.  Taking an
    # off-diagonal quadrant, the wrong half, or the wrong triangle fails.
    T, F = 40, 6
    data = np.random.default_rng(16).integers(0, 256, size=(T, F, 16), dtype=np.uint8)
    full = O.blocks_to_full(O.correlate_blocks(data.reshape(T, F // 2, 32), U), 32)
    tri = O.feed16_to_triangle(O.split_paired_feeds(full, F, 16), F)
    np.testing.assert_array_equal(tri, O.correlate_triangle(data, U))
