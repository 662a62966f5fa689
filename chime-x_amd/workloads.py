"""Synthetic workloads of the correlator (job lists and input makers).

Inputs are [time][freq][element] bytes, real in the high nibble and
imaginary in the low nibble, offset binary (input_generator.c:84-96); a
uniform random byte is a uniform 4-bit complex sample in [-8, 7].
"""
from __future__ import annotations

import numpy as np

UNEVEN_N = (256, 512, 1024, 2048)


def uneven_batch_jobs(njobs: int = 64, seed: int = 3, U: int = 1,
                      sizes: tuple = UNEVEN_N) -> list[dict]:
    """The uneven batch: N from {256, 512, 1024, 2048} (or `sizes`), 1-8
    channels, an integration length in [4096, 131072] (a multiple of 256),
    and a data seed per job."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(njobs):
        jobs.append(dict(num_elements=int(rng.choice(sizes)),
                         num_freq=int(rng.integers(1, 9)),
                         time_steps=int(rng.integers(4096 // 256, 131072 // 256 + 1)) * 256,
                         upper_triangle_convention=U,
                         seed=1000 + k))
    return jobs


def cmacs(N: int, F: int, T: int) -> int:
    """Algorithmic complex MACs of one correlation: F * T * N(N+1)/2."""
    return F * T * (N * (N + 1) // 2)


def device_uniform(torch, shape, seed: int, device: str = "cuda"):
    """Uniform random bytes generated on the device from a seed."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return torch.randint(0, 256, shape, dtype=torch.uint8, device=device, generator=g)


def streaming_frame(torch, T: int, F: int, N: int, seed: int, const_frac: float = 0.05,
                    const_byte: int = 0xF0, device: str = "cuda"):
    """One frame of streaming input: uniform bytes, with a fixed subset
    (const_frac of the elements, picked from the seed) held at const_byte
    ((7, -8) for 0xF0)."""
    x = device_uniform(torch, (T, F, N), seed, device)
    held = held_elements(N, const_frac)
    if len(held):
        x[:, :, torch.as_tensor(held, device=device)] = const_byte
    return x


def held_elements(N: int, const_frac: float = 0.05, seed: int = 11) -> np.ndarray:
    """The elements held constant in streaming frames (same for every frame)."""
    k = int(round(N * const_frac))
    return np.sort(np.random.default_rng(seed).choice(N, size=k, replace=False))
