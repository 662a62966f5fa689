"""ctypes wrapper over the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker, never as the thing
measured or shipped.  The product path (chime-x_amd) never imports it.

Each function mirrors one C entry point of chx_oracle.c; that file cites the
reference file:line each one restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libchxref.so")

GEN_CONSTANT, GEN_RAMP_UP, GEN_RAMP_DOWN, GEN_RANDOM = 1, 2, 3, 4
ALL_FREQUENCIES = -1

_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i32, i64, vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
        L.chxo_generate.argtypes = [i32] * 7 + [i64, i32, i32, i32, vp]
        L.chxo_correlate_triangle.argtypes = [i64, i32, i32, vp, i32, vp]
        L.chxo_correlate_blocks.argtypes = [i64, i32, i32, vp, i32, vp]
        L.chxo_generate_and_correlate_triangle.argtypes = [i32, i32, i32, vp] + [i32] * 9
        L.chxo_num_blocks.argtypes = [i32]
        L.chxo_num_blocks.restype = i64
        L.chxo_block_maps.argtypes = [i32, vp, vp]
        L.chxo_blocks_to_triangle.argtypes = [i32, i32, vp, vp]
        L.chxo_blocks_to_full.argtypes = [i32, i32, vp, vp]
        L.chxo_offset_sums.argtypes = [i64, i32, i32, vp, vp]
        L.chxo_split_paired_feeds.argtypes = [i32, i32, vp]
        L.chxo_feed16_to_triangle.argtypes = [i32, i32, vp, vp]
        L.chxo_preseed.argtypes = [i32, i32, i64, i32, i32, vp, vp]
        L.chxo_packed_raw.argtypes = [i32, i32, i64, i32, i32, i32, vp, vp]
        L.chxo_compare_triangle.argtypes = [i32, i32, vp, vp, vp, vp]
        L.chxo_offset_and_clip.argtypes = [i32] * 4
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def num_blocks(N: int) -> int:
    return (N // 32) * (N // 32 + 1) // 2


def generate(T, F, N, gen_type=GEN_RANDOM, seed=42, default_real=0, default_imag=0,
             initial_real=0, initial_imag=0, single_freq=ALL_FREQUENCIES,
             no_repeat=False) -> np.ndarray:
    out = np.empty((T, F, N), np.uint8)
    lib().chxo_generate(gen_type, seed, default_real, default_imag, initial_real,
                        initial_imag, single_freq, T, F, N, int(no_repeat), _p(out))
    return out


def correlate_triangle(data: np.ndarray, U: int = 1) -> np.ndarray:
    T, F, N = data.shape
    data = np.ascontiguousarray(data, np.uint8)
    out = np.empty((F, N * (N + 1) // 2, 2), np.int32)
    lib().chxo_correlate_triangle(T, F, N, _p(data), U, _p(out))
    return out


def correlate_blocks(data: np.ndarray, U: int = 1) -> np.ndarray:
    T, F, N = data.shape
    data = np.ascontiguousarray(data, np.uint8)
    out = np.empty((F, num_blocks(N), 32, 32, 2), np.int32)
    lib().chxo_correlate_blocks(T, F, N, _p(data), U, _p(out))
    return out


def generate_and_correlate_triangle(T, F, N, gen_type=GEN_RANDOM, seed=42,
                                    default_real=0, default_imag=0, initial_real=0,
                                    initial_imag=0, single_freq=ALL_FREQUENCIES,
                                    no_repeat=False, U=1) -> np.ndarray:
    out = np.empty((F, N * (N + 1) // 2, 2), np.int32)
    rc = lib().chxo_generate_and_correlate_triangle(
        T, F, N, _p(out), gen_type, seed, default_real, default_imag, initial_real,
        initial_imag, single_freq, int(no_repeat), U)
    assert rc == 0
    return out


def block_maps(N: int):
    nb = num_blocks(N)
    x = np.empty(nb, np.uint32)
    y = np.empty(nb, np.uint32)
    lib().chxo_block_maps(N, _p(x), _p(y))
    return x, y


def blocks_to_triangle(blocks: np.ndarray, N: int) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, np.int32)
    F = blocks.size // (num_blocks(N) * 2048)
    tri = np.empty((F, N * (N + 1) // 2, 2), np.int32)
    lib().chxo_blocks_to_triangle(F, N, _p(blocks), _p(tri))
    return tri


def blocks_to_full(blocks: np.ndarray, N: int) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, np.int32)
    F = blocks.size // (num_blocks(N) * 2048)
    full = np.zeros((F, N, N, 2), np.int32)
    lib().chxo_blocks_to_full(F, N, _p(blocks), _p(full))
    return full


def split_paired_feeds(full: np.ndarray, F: int, E: int) -> np.ndarray:
    """[F//2][2E][2E][2] full matrices -> [F//2*2][E][E][2] (a copy)."""
    work = np.ascontiguousarray(full, np.int32).copy()
    assert work.size >= (F // 2) * (2 * E) ** 2 * 2
    lib().chxo_split_paired_feeds(F, E, _p(work))
    return work.reshape(-1)[: (F // 2) * 2 * E * E * 2].reshape(F // 2 * 2, E, E, 2).copy()


def feed16_to_triangle(full16: np.ndarray, F_out: int) -> np.ndarray:
    """[F_in][16][16][2] -> [F_out][136][2], zero past F_in."""
    full16 = np.ascontiguousarray(full16, np.int32)
    F_in = full16.size // 512
    out = np.empty((F_out, 136, 2), np.int32)
    lib().chxo_feed16_to_triangle(F_out, F_in, _p(full16), _p(out))
    return out


def offset_sums(data: np.ndarray) -> np.ndarray:
    T, F, N = data.shape
    data = np.ascontiguousarray(data, np.uint8)
    sums = np.zeros((F, N, 2), np.uint32)
    lib().chxo_offset_sums(T, F, N, _p(data), _p(sums))
    return sums


def preseed(sums: np.ndarray, T: int, method: int, U: int) -> np.ndarray:
    F, N, _ = sums.shape
    sums = np.ascontiguousarray(sums, np.uint32)
    out = np.empty((F, num_blocks(N), 32, 32, 2), np.int32)
    lib().chxo_preseed(method, U, T, F, N, _p(sums), _p(out))
    return out


def packed_raw(data: np.ndarray, method: int, U: int, time_accum: int,
               into: np.ndarray | None = None) -> np.ndarray:
    T, F, N = data.shape
    data = np.ascontiguousarray(data, np.uint8)
    out = np.zeros((F, num_blocks(N), 32, 32, 2), np.int32) if into is None else into
    lib().chxo_packed_raw(method, U, T, F, N, time_accum, _p(data), _p(out))
    return out


def packed_pipeline(data: np.ndarray, method: int, U: int, time_accum: int) -> np.ndarray:
    """offset sums -> preseed (assign) -> packed raw products (+=): the
    reference's per-iteration kernel order (main_wrapper.c:691-778)."""
    T = data.shape[0]
    blocks = preseed(offset_sums(data), T, method, U)
    return packed_raw(data, method, U, time_accum, into=blocks)


def compare_triangle(gpu: np.ndarray, cpu: np.ndarray, N: int):
    gpu = np.ascontiguousarray(gpu, np.int32)
    cpu = np.ascontiguousarray(cpu, np.int32)
    F = gpu.size // (N * (N + 1))
    e2 = ctypes.c_int64()
    mx = ctypes.c_int64()
    n = lib().chxo_compare_triangle(F, N, _p(gpu), _p(cpu), ctypes.addressof(e2),
                                    ctypes.addressof(mx))
    return n, e2.value, mx.value
