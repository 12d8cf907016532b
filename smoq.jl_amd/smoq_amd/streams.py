"""Seeded uniform streams shared by host code, tests and the CPU baseline.

The reference draws from Julia's global `rand()` (unseeded, Xoshiro).  Here
every random input — angles, fields, sampling and jump uniforms — comes from
a SplitMix64 stream so that the device path and the CPU oracle see the same
bytes.  u_i = (splitmix64(seed + (i+1)·γ) >> 11) · 2^-53  ∈ [0, 1).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Raw 64-bit outputs n..n+offset of the SplitMix64 sequence."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniforms(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1) with 53 random bits each."""
    z = splitmix64(seed, n, offset)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
