"""Arrays well beyond the reference's sizes on both matrix-core methods:
N = 8192 (F = 2) and N = 12320 (F = 1, not a multiple of the 128-element
tile), both conventions.  Sampled 32x32 blocks -- the first, the last,
diagonal ones, the last tile's partial columns and random ones -- equal the
oracle run on the sub-array of the block's 64 (or, on the diagonal, 32)
elements, whose block (0, 1) is the full array's block (by, bx)."""
import sys

import numpy as np
import pytest

import chx
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def block_index(by, bx, nb):
    return by * nb - by * (by - 1) // 2 + (bx - by)


@pytest.mark.parametrize("method", [chx.FP6_MFMA, chx.INT8_MFMA])
@pytest.mark.parametrize("N,F,T,U", [(8192, 2, 512, 1), (12320, 1, 256, 0)])
def test_very_large_arrays_sampled_blocks(method, N, F, T, U):
    rng = np.random.default_rng(N + F + T + method)
    data = rng.integers(0, 256, size=(T, F, N), dtype=np.uint8)
    say(f"  case method={chx.METHOD_NAMES[method]} N={N} F={F} T={T} U={U}")
    with chx.Plan(N, F, T, 256, U, method) as p:
        got = p.correlate(data)
    nb = N // 32
    assert got.shape == (F, chx.num_blocks(N), 32, 32, 2)
    picks = {(0, 0), (0, nb - 1), (nb - 1, nb - 1), (nb - 2, nb - 1), (nb // 2, nb // 2),
             (nb // 2, nb // 2 + 1), (3, 4), (4, 7)}
    while len(picks) < 40:
        by = int(rng.integers(0, nb))
        picks.add((by, int(rng.integers(by, nb))))
    for by, bx in sorted(picks):
        if by == bx:
            sub = data[:, :, 32 * by:32 * by + 32]
            want = O.correlate_blocks(np.ascontiguousarray(sub), U)[:, 0]
        else:
            idx = np.r_[32 * by:32 * by + 32, 32 * bx:32 * bx + 32]
            want = O.correlate_blocks(np.ascontiguousarray(data[:, :, idx]), U)[:, 1]
        np.testing.assert_array_equal(got[:, block_index(by, bx, nb)], want,
                                      err_msg=f"block ({by}, {bx})")
