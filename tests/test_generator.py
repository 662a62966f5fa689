"""The reference generator (input_generator.c:17-103) rebuilt for the device.

CPU: the host restatement of glibc's rand() (TYPE_3 additive feedback,
jump-ahead by matrix powers) equals libc rand() itself for several seeds
(0 and negative seeds included) and offsets.
GPU: chx_generate_device writes the reference generator's exact bytes for
every case the reference fixtures hold (all generation types, repeat and
continuous random modes, clipping, single-frequency incl. the q == F quirk)
and for larger continuous runs that span many generator segments.
"""
import ctypes

import numpy as np
import pytest

import chx
import oracle as O

LIBC = ctypes.CDLL("libc.so.6")


def libc_rand(seed, n):
    LIBC.srand(ctypes.c_uint(seed & 0xFFFFFFFF))
    return np.array([LIBC.rand() for _ in range(n)], np.int32)


@pytest.mark.parametrize("seed", [42, 0, 1, 7, 11, -5, 2**31 - 1, 3_000_000_000 - 2**32])
def test_rand_restatement_matches_libc(seed):
    ref = libc_rand(seed, 2500)
    np.testing.assert_array_equal(chx.rand_host(seed, 0, 2500), ref)
    for k0 in (1, 31, 344, 1000, 2400):
        np.testing.assert_array_equal(chx.rand_host(seed, k0, 100), ref[k0:k0 + 100])


def test_rand_jump_far_ahead():
    ref = libc_rand(42, 300_000)
    np.testing.assert_array_equal(chx.rand_host(42, 299_900, 100), ref[-100:])


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


@pytest.mark.gpu
def test_device_generator_matches_reference_fixtures(golden, golden_cases):
    torch = _torch()
    for case in golden_cases:
        if f"{case}__input" not in golden:
            continue
        T, F, N, g, seed, dre, dim, ire, iim, sf, nr = (int(v) for v in golden[f"{case}__params"])
        d = torch.empty((T, F, N), dtype=torch.uint8, device="cuda")
        chx.generate_device(d.data_ptr(), T, F, N, g, seed, dre, dim, ire, iim, sf, bool(nr))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d.cpu().numpy(), golden[f"{case}__input"], err_msg=case)


@pytest.mark.gpu
@pytest.mark.parametrize("T,F,N,seed,no_repeat", [
    (4096, 4, 512, 42, True),    # 16 MiB: ~1000 generator segments
    (777, 3, 96, 99, True),      # ragged tail inside a segment
    (300, 5, 64, 5, False),      # repeat mode, several rows
    (1, 1, 32, 1, True),
])
def test_device_generator_matches_oracle(T, F, N, seed, no_repeat):
    torch = _torch()
    d = torch.empty((T, F, N), dtype=torch.uint8, device="cuda")
    chx.generate_device(d.data_ptr(), T, F, N, chx.GEN_RANDOM, seed, no_repeat=no_repeat)
    torch.cuda.synchronize()
    want = O.generate(T, F, N, gen_type=O.GEN_RANDOM, seed=seed, no_repeat=no_repeat)
    np.testing.assert_array_equal(d.cpu().numpy(), want)
