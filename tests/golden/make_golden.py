"""Writes the golden fixtures under tests/golden/ from the reference's own C.

Run in a container that has the reference tree (it is compiled by
oracle/Makefile into oracle/_ref/libchxref.so from cpu_corr_test.c,
input_generator.c, gpu_data_reorg.c and gpu_cpu_helpers.c where they lie).
The fixtures are data only: generator inputs, the reference CPU correlator's
outputs, and the reference reorg's outputs, saved with numpy (no pickles).

    python tests/golden/make_golden.py

Reference entry points used (file:line):
  generate_char_data_set                               input_generator.c:17
  cpu_data_generate_and_correlate_upper_triangle_only  cpu_corr_test.c:156
  ..._upper_triangle_only_nonstandard_convention        cpu_corr_test.c:59
  cpu_data_generate_and_correlate                      cpu_corr_test.c:108
  cpu_data_generate_and_correlate_nonstandard_convention cpu_corr_test.c:11
  reorganize_GPU_to_upper_triangle                     gpu_data_reorg.c:76
  reorganize_GPU_to_full_Matrix_for_comparison         gpu_data_reorg.c:36
  reorganize_32_to_16_feed_GPU_Correlated_Data         gpu_data_reorg.c:6
  reorganize_data_16_element_with_triangle_conversion  gpu_data_reorg.c:126
  compare_NSquared_correlator_results_data_has_upper_triangle_only
                                                        cpu_corr_test.c:269
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libchxref.so")


def load_ref():
    if not os.path.exists(REF_LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    L = ctypes.CDLL(REF_LIB)
    i, vp = ctypes.c_int, ctypes.c_void_p
    L.generate_char_data_set.argtypes = [i] * 11 + [vp]
    for name in ("cpu_data_generate_and_correlate_upper_triangle_only",
                 "cpu_data_generate_and_correlate_upper_triangle_only_nonstandard_convention",
                 "cpu_data_generate_and_correlate",
                 "cpu_data_generate_and_correlate_nonstandard_convention"):
        getattr(L, name).argtypes = [i, i, i, vp] + [i] * 9
    L.reorganize_GPU_to_upper_triangle.argtypes = [i, i, i, i, vp, vp]
    L.reorganize_GPU_to_full_Matrix_for_comparison.argtypes = [i, i, i, i, vp, vp]
    L.compare_NSquared_correlator_results_data_has_upper_triangle_only.argtypes = [
        vp, vp, i, i, vp, vp, vp, vp, i]
    L.reorganize_32_to_16_feed_GPU_Correlated_Data.argtypes = [i, i, vp]
    L.reorganize_data_16_element_with_triangle_conversion.argtypes = [i, i, vp, vp]
    return L


def p(a):
    return a.ctypes.data


# (name, T, F, N, gen_type, seed, default_re, default_im, init_re, init_im,
#  single_freq, no_repeat)
CASES = [
    ("c1_repeat", 256, 1, 256, 4, 42, 0, 0, 0, 0, -1, 0),
    ("c1_continuous", 256, 1, 256, 4, 42, 0, 0, 0, 0, -1, 1),
    ("rand_cont_s7", 64, 2, 64, 4, 7, 0, 0, 0, 0, -1, 1),
    ("rand_repeat_s7", 16, 2, 32, 4, 7, 0, 0, 0, 0, -1, 0),
    ("const_m3_5", 16, 2, 32, 1, 42, 0, 0, -3, 5, -1, 0),
    ("const_clip", 8, 1, 32, 1, 42, 0, 0, 20, -20, -1, 0),
    ("ramp_up", 24, 3, 32, 2, 42, 0, 0, 2, -5, -1, 0),
    ("ramp_down", 24, 3, 64, 3, 42, 0, 0, -1, 6, -1, 0),
    ("single_freq_1", 32, 3, 32, 4, 11, 3, -2, 0, 0, 1, 1),
    ("single_freq_eqF", 8, 2, 32, 4, 11, -4, 7, 0, 0, 2, 1),
    ("gen_default", 8, 1, 32, 9, 42, 5, -6, 0, 0, -1, 0),
    ("saturate_00", 40, 1, 32, 1, 0, 0, 0, -8, -8, -1, 0),
]


def main() -> int:
    L = load_ref()
    out = {}
    for (name, T, F, N, g, seed, dre, dim, ire, iim, sf, nr) in CASES:
        data = np.zeros((T, F, N), np.uint8)
        L.generate_char_data_set(g, seed, dre, dim, ire, iim, sf, T, F, N, nr, p(data))
        out[f"{name}__input"] = data
        out[f"{name}__params"] = np.array([T, F, N, g, seed, dre, dim, ire, iim, sf, nr],
                                          np.int64)
        for U, fn in ((1, L.cpu_data_generate_and_correlate_upper_triangle_only),
                      (0, L.cpu_data_generate_and_correlate_upper_triangle_only_nonstandard_convention)):
            tri = np.zeros((F, N * (N + 1) // 2, 2), np.int32)
            rc = fn(T, F, N, p(tri), g, seed, dre, dim, ire, iim, sf, nr, 0)
            assert rc == 0
            out[f"{name}__tri_U{U}"] = tri
        if N <= 64:
            for U, fn in ((1, L.cpu_data_generate_and_correlate),
                          (0, L.cpu_data_generate_and_correlate_nonstandard_convention)):
                full = np.zeros((F, N, N, 2), np.int32)
                rc = fn(T, F, N, p(full), g, seed, dre, dim, ire, iim, sf, nr, 0)
                assert rc == 0
                out[f"{name}__full_U{U}"] = full

    # Reorg: a block buffer whose entries are their own index.
    for N, F in ((96, 2), (32, 1), (256, 1)):
        nb = (N // 32) * (N // 32 + 1) // 2
        blocks = np.arange(F * nb * 2048, dtype=np.int32)
        tri = np.zeros((F, N * (N + 1) // 2, 2), np.int32)
        L.reorganize_GPU_to_upper_triangle(32, nb, F, N, p(blocks), p(tri))
        out[f"reorg_N{N}_F{F}__tri"] = tri

    # Full Hermitian reorg: signed values (their own index, shifted).
    for N, F in ((96, 2), (64, 1)):
        nb = (N // 32) * (N // 32 + 1) // 2
        blocks = (np.arange(F * nb * 2048, dtype=np.int64) * 37 - 100000).astype(np.int32)
        full = np.zeros((F, N, N, 2), np.int32)
        L.reorganize_GPU_to_full_Matrix_for_comparison(32, nb, F, N, p(blocks), p(full))
        out[f"reorgfull_N{N}_F{F}__full"] = full

    # 16-feed helpers: signed values that name their own position.
    for F, E in ((4, 16), (5, 16), (6, 7), (1, 16)):
        fin = F // 2
        full = (np.arange(max(fin, 1) * (2 * E) ** 2 * 2, dtype=np.int64) * 13
                - 50000).astype(np.int32)
        out[f"split_F{F}_E{E}__in"] = full.copy()
        L.reorganize_32_to_16_feed_GPU_Correlated_Data(F, E, p(full))
        out[f"split_F{F}_E{E}__out"] = full
    for F_out, F_in in ((4, 4), (6, 3), (2, 5), (3, 0)):
        src = (np.arange(max(F_in, 1) * 512, dtype=np.int64) * -7 + 3000).astype(np.int32)
        tri = np.full(F_out * 272, 0x5A5A5A5A, np.int32)
        L.reorganize_data_16_element_with_triangle_conversion(F_out, F_in, p(src), p(tri))
        out[f"tri16_out{F_out}_in{F_in}__in"] = src
        out[f"tri16_out{F_out}_in{F_in}__out"] = tri

    # Comparator: planted differences.
    N, F = 64, 2
    rng = np.random.default_rng(5)
    cpu = rng.integers(-1000, 1000, size=(F, N * (N + 1) // 2, 2), dtype=np.int32)
    gpu = cpu.copy()
    idx = rng.choice(cpu.shape[1], size=17, replace=False)
    gpu[0, idx, 0] += rng.integers(1, 50, size=17).astype(np.int32)
    gpu[1, idx[:5], 1] -= 3
    nerr = ctypes.c_int()
    err2 = ctypes.c_int64()
    ratio = np.zeros(F * N * N, np.float64)
    phase = np.zeros(F * N * N, np.float64)
    sys.stdout.flush()
    L.compare_NSquared_correlator_results_data_has_upper_triangle_only(
        ctypes.addressof(nerr), ctypes.addressof(err2), F, N, p(gpu), p(cpu),
        p(ratio), p(phase), 0)
    out["compare__cpu"] = cpu
    out["compare__gpu"] = gpu
    out["compare__result"] = np.array([nerr.value, err2.value], np.int64)

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
