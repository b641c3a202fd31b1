"""Deterministic test inputs: splitmix64 in numpy uint64 arithmetic, mapped to
uniform [-1, 1). Stable across numpy versions (no Generator stream involved)."""
import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, count):
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_real(seed, n):
    z = splitmix64(seed, n)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / (1 << 53)) - 1.0


def uniform_complex(seed, n):
    v = uniform_real(seed, 2 * n)
    return v[0::2] + 1j * v[1::2]
