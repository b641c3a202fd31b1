"""Shared helpers for the GPU parity tests (inputs seeded like the reference's
tests: a fixed seed, several random vectors per length)."""
import numpy as np

SEED = 0xDAE14B0B


def rng(tag=0):
    return np.random.default_rng([SEED, tag])


def crand(r, shape, dtype=np.complex128):
    return (r.random(shape) + 1j * r.random(shape)).astype(dtype)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


# FP64 plain path: rel-L2 <= 1e-12 against the FP64 oracle.
TOL_F64 = 1e-12


def tol_f32(n):
    """FP32 form budget against the FP64 oracle: 1e-6 * log2(N), at least 2e-6."""
    return max(2e-6, 1e-6 * np.log2(max(n, 2)))
