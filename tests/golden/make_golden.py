"""Writes tests/golden/golden.npz: independent known answers that pin the CPU
oracle (oracle/chfft_oracle.cpp) without running the reference (a Rust crate
that cannot be built here).

Sources of truth, none of them the oracle:
  * mpmath at 40 digits: exact-definition DFTs (cfft1d.rs:481-494) for every
    N in 1..40 and the primes up to 97, one vector each;
  * numpy/pocketfft (FP64) at the benchmark lengths, stored as a fixed subset
    of bins (the full vectors would bloat the repository);
  * numpy float64 direct sums with integer-reduced angles for DCT-II
    (dct1d.rs:423-434) and MDCT/IMDCT (mdct1d.rs:305-339);
  * the README example [2, 1+i, 3i, 2+4i] -> [5+8i, -1-2i, -1-2i, 5-4i].
Inputs come from splitmix64 (tests/golden/gen.py), so they can be regenerated
bit-for-bit at test time from their seed.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import mpmath
import numpy as np
import scipy.fft as sfft

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from gen import uniform_complex, uniform_real  # noqa: E402

mpmath.mp.dps = 40
PRIMES = [41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]
BIG_C2C = [4096, 43200, 4093, 10007, 65537, 2048, 512]
NBINS = 257


def mp_dft(x, sign):
    n = len(x)
    xs = [mpmath.mpc(complex(v)) for v in x]
    out = []
    for k in range(n):
        s = mpmath.mpc(0)
        for j in range(n):
            s += xs[j] * mpmath.expjpi(sign * mpmath.mpf(2 * ((j * k) % n)) / n)
        out.append(complex(s))
    return np.array(out)


def bins(n, seed):
    r = np.random.default_rng(seed)
    b = np.unique(np.concatenate([[0, 1, n // 2, n - 1], r.integers(0, n, NBINS)]))
    return b[b < n]


def dct2_direct(x):
    n = len(x)
    j = np.arange(n)
    out = np.empty(n)
    for k in range(n):
        m = ((2 * j + 1) * k) % (4 * n)
        out[k] = np.sum(x * np.cos(np.pi * m / (2 * n)))
    return out


def mdct_direct(x, w):
    n = len(x)
    k = np.arange(n)
    out = np.empty(n // 2)
    for m in range(n // 2):
        a = ((1 + n // 2 + 2 * k) * (1 + 2 * m)) % (4 * n)
        out[m] = np.sum(w * x * np.cos(np.pi / (2 * n) * a))
    return out


def imdct_direct(c, w):
    n = 2 * len(c)
    m = np.arange(n // 2)
    out = np.empty(n)
    for k in range(n):
        a = ((1 + n // 2 + 2 * k) * (1 + 2 * m)) % (4 * n)
        out[k] = 4.0 * w[k] / n * np.sum(c * np.cos(np.pi / (2 * n) * a))
    return out


def main():
    g = {}
    g["kat_in"] = np.array([2, 1 + 1j, 3j, 2 + 4j])
    g["kat_out"] = np.array([5 + 8j, -1 - 2j, -1 - 2j, 5 - 4j])
    for n in list(range(1, 41)) + PRIMES:
        x = uniform_complex(1000 + n, n)
        g[f"dft_fwd_{n}"] = mp_dft(x, -1)
        g[f"dft_bwd_{n}"] = mp_dft(x, +1)
    for n in BIG_C2C:
        x = uniform_complex(2000 + n, n)
        b = bins(n, n)
        g[f"np_bins_{n}"] = b
        g[f"np_fwd_{n}"] = np.fft.fft(x)[b]
        g[f"np_bwd_{n}"] = (np.fft.ifft(x) * n)[b]
    for n in [8192, 4096, 10, 6, 2]:
        x = uniform_real(3000 + n, n)
        g[f"rfft_{n}"] = np.fft.rfft(x)
    for n in [1024, 2, 6, 12, 30, 126]:
        x = uniform_real(4000 + n, n)
        g[f"dct2_{n}"] = dct2_direct(x)
        # scipy's type-2 is 2x chfft's unnormalised DCT-II; type-3 is 2x chfft's
        # non-ortho DCT-III (x0/2 + sum ...)
        g[f"scipy_dct2_{n}"] = sfft.dct(x, 2) / 2
        g[f"scipy_dct3_{n}"] = sfft.dct(x, 3) / 2
    for n in [2048, 4, 8, 12, 64, 396]:
        x = uniform_real(5000 + n, n)
        w = np.sin(np.pi * (np.arange(n) + 0.5) / n)
        c = mdct_direct(x, w)
        g[f"mdct_{n}"] = c
        g[f"imdct_{n}"] = imdct_direct(c, w)
    x = uniform_complex(6000, 7 * 5).reshape(7, 5)
    g["fft2_7x5"] = np.fft.fft2(x)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **g)
    print("wrote", len(g), "arrays")


if __name__ == "__main__":
    main()
