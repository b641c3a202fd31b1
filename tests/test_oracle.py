"""CPU tests pinning the oracle (oracle/chfft_oracle.cpp) — no GPU needed.

Each oracle function is checked against something that is not the oracle:
mpmath exact DFTs, numpy/pocketfft, scipy DCTs, float64 direct sums (the
golden fixtures, tests/golden/make_golden.py), the README known answer, the
long-double brute-force definitions that the reference's own tests use, the
plan-structure table of the reference (factor orders), and identities the
mathematics fixes (round trips, Parseval, linearity, shift, Hermitian symmetry,
DCT-II/III inverse pairs, MDCT TDAC).
"""
import os

import numpy as np
import pytest

import oracle
from tests.golden.gen import uniform_complex, uniform_real

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
METHODS = ["forward", "forward0", "forwardu", "forwardn", "backward", "backward0", "backwardu", "backwardn"]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb else 1.0)


def scale_of(m, n):
    return {"forward": 1, "forward0": 1, "forwardu": n ** -0.5, "forwardn": 1 / n, "backward": 1 / n,
            "backwardn": 1 / n, "backward0": 1, "backwardu": n ** -0.5}[m]


def test_readme_known_answer():
    assert rel(oracle.c2c(G["kat_in"], "forward"), G["kat_out"]) < 1e-15


@pytest.mark.parametrize("n", list(range(1, 41)) + [41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97])
def test_c2c_vs_mpmath(n):
    # mixed radix for 7-smooth n, Bluestein otherwise (cfft1d.rs:104-111)
    x = uniform_complex(1000 + n, n)
    for m in METHODS:
        y = oracle.c2c(x, m)
        if n == 1:  # cfft1d.rs:213-214: returned unchanged, no scaling
            assert np.array_equal(y, x)
            continue
        ref = (G[f"dft_fwd_{n}"] if m.startswith("f") else G[f"dft_bwd_{n}"]) * scale_of(m, n)
        assert rel(y, ref) < 2e-15 * max(1, np.log2(n)), (m, rel(y, ref))
        yi = oracle.c2c(x, m, inplace=True)
        assert rel(yi, ref) < 2e-15 * max(1, np.log2(n))
    y32 = oracle.c2c(x.astype(np.complex64), "forward")
    assert rel(y32, G[f"dft_fwd_{n}"]) < 1e-6 * max(1, np.log2(n)) * 4


@pytest.mark.parametrize("n", [4096, 43200, 4093, 10007, 65537, 2048, 512])
def test_c2c_vs_numpy_config_lengths(n):
    x = uniform_complex(2000 + n, n)
    b = G[f"np_bins_{n}"]
    assert rel(oracle.c2c(x, "forward")[b], G[f"np_fwd_{n}"]) < 1e-14
    assert rel(oracle.c2c(x, "backward0")[b], G[f"np_bwd_{n}"]) < 1e-14


def test_brute_force_long_double_agrees_with_mpmath():
    for n in (7, 16, 31, 40):
        x = uniform_complex(1000 + n, n)
        assert rel(oracle.dft_brute(x, -1), G[f"dft_fwd_{n}"]) < 1e-15
        assert rel(oracle.dft_brute(x, +1), G[f"dft_bwd_{n}"]) < 1e-15


@pytest.mark.parametrize("n", [2 * 3 * 5 * 7, 256, 1000, 2 * 11, 101, 127 * 2, 7 ** 3, 1 << 13])
def test_c2c_vs_brute_force(n):
    x = uniform_complex(7000 + n, n)
    assert rel(oracle.c2c(x, "forward"), oracle.dft_brute(x, -1)) < 1e-14
    assert rel(oracle.c2c(x, "backward0"), oracle.dft_brute(x, +1)) < 1e-14


def test_factor_orders_match_reference_table():
    # prime_factorization.rs:66-113 (4-group deferred behind odd primes)
    assert oracle.factors(4096) == [(4, 6)]
    assert oracle.factors(43200) == [(3, 3), (5, 2), (4, 3)]
    assert oracle.factors(512) == [(2, 1), (4, 4)]
    assert oracle.factors(2048) == [(2, 1), (4, 5)]
    assert oracle.factors(2 * 7 * 7) == [(2, 1), (7, 2)]
    # the 4-group is pushed in front of a group with prime > 4 only when that
    # group is closed inside the loop (prime_factorization.rs:79-87)
    assert oracle.factors(4 * 9 * 7) == [(3, 2), (7, 1), (4, 1)]
    assert oracle.factors(4 * 3 * 5 * 7) == [(3, 1), (4, 1), (5, 1), (7, 1)]
    for p in (4093, 10007, 65537, 11, 2 * 13):
        assert oracle.factors(p) == []
    assert oracle.factors(1) == []


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 10, 12, 15, 16, 30, 97, 1000])
def test_omega_table(n):
    # precompute_utils.rs:24-68: W_k = exp(-2 pi i k / n), k = 0..n, last = 1
    import mpmath
    mpmath.mp.dps = 30
    w = oracle.omega(n)
    ref = np.array([complex(mpmath.expjpi(-2 * mpmath.mpf(k % n) / n)) for k in range(n + 1)])
    # odd lengths evaluate angles up to 2 pi directly: ulp(2 pi) ~ 9e-16
    assert np.max(np.abs(w - ref)) < 2e-15
    assert w[n] == 1.0 and w[0] == 1.0


def test_bitreverse_is_digit_reversal():
    # radix-2 only: classic bit reversal
    n = 64
    ids = oracle.bitreverse(n, [(2, 6)])
    ref = [int(format(i, "06b")[::-1], 2) for i in range(n)]
    assert list(ids) == ref
    # mixed radix: a permutation, and the swap sequence realises the same gather
    for facs, n in [([(3, 3), (5, 2), (4, 1)], 27 * 25 * 4), ([(2, 1), (4, 3)], 128)]:
        ids = oracle.bitreverse(n, facs)
        assert sorted(ids) == list(range(n))
        sw = oracle.bitreverse(n, facs, inplace=True)
        a = np.arange(n)
        for i, s in enumerate(int(v) for v in sw):
            if i != s:
                a[[i, s]] = a[[s, i]]
        assert np.array_equal(a, ids)


def test_identities():
    r = np.random.default_rng(3)
    for n in (30, 64, 97, 4093 // 3):
        x = r.standard_normal(n) + 1j * r.standard_normal(n)
        y = r.standard_normal(n) + 1j * r.standard_normal(n)
        X = oracle.c2c(x, "forward")
        # round trips at every scaling pair (cfft1d.rs:496-543)
        assert rel(oracle.c2c(X, "backward"), x) < 1e-14
        assert rel(oracle.c2c(oracle.c2c(x, "forwardu"), "backwardu"), x) < 1e-14
        assert rel(oracle.c2c(oracle.c2c(x, "forwardn"), "backward0"), x) < 1e-14
        # Parseval, linearity, shift theorem
        assert abs(np.vdot(X, X).real / n - np.vdot(x, x).real) < 1e-12 * np.vdot(x, x).real
        assert rel(oracle.c2c(2 * x - 3j * y, "forward"), 2 * X - 3j * oracle.c2c(y, "forward")) < 1e-14
        k = np.arange(n)
        assert rel(oracle.c2c(np.roll(x, 1), "forward"), X * np.exp(-2j * np.pi * k / n)) < 1e-13
        # delta -> ones, constant -> N * delta
        d = np.zeros(n, complex)
        d[0] = 1
        assert rel(oracle.c2c(d, "forward"), np.ones(n)) < 1e-15
        e = np.zeros(n, complex)
        e[0] = n
        assert rel(oracle.c2c(np.ones(n, complex), "forward"), e) < 1e-14


def test_sequence_replanning_matches_fresh_plans():
    # cfft1d.rs:216-218 replans on a length change; chirp-z level reuse at
    # cfft1d.rs:142-160 (11 and 13 share M = 32)
    lens = [11, 13, 16, 17, 100, 11]
    xs = [uniform_complex(10 + n, n) for n in lens]
    got = oracle.c2c_sequence(xs)
    for x, y in zip(xs, got):
        assert rel(y, oracle.c2c(x, "forward")) < 1e-15


@pytest.mark.parametrize("n", [8192, 4096, 10, 6, 2])
def test_rfft_vs_numpy(n):
    x = uniform_real(3000 + n, n)
    assert rel(oracle.rfft(x, "forward"), G[f"rfft_{n}"]) < 1e-14
    X = oracle.rfft(x, "forward")
    assert rel(oracle.rfft(X, "backward", n=n), x) < 1e-14
    for m in ("forwardu", "forwardn"):
        assert rel(oracle.rfft(x, m), G[f"rfft_{n}"] * scale_of(m, n)) < 1e-14


def test_rfft_ignores_dc_and_nyquist_imaginary_parts():
    # rfft1d.rs:164-167 reads only the real parts of X[0] and X[N/2]
    n = 16
    X = oracle.rfft(uniform_real(1, n), "forward")
    Y = X.copy()
    Y[0] += 5j
    Y[n // 2] -= 2j
    assert np.array_equal(oracle.rfft(X, "backward", n=n), oracle.rfft(Y, "backward", n=n))


@pytest.mark.parametrize("n", [1024, 2, 6, 12, 30, 126])
def test_dct_vs_direct_and_scipy(n):
    x = uniform_real(4000 + n, n)
    d2 = oracle.dct(x, 2)
    assert rel(d2, G[f"dct2_{n}"]) < 1e-13
    assert rel(d2, G[f"scipy_dct2_{n}"]) < 1e-13
    assert rel(oracle.dct(x, 3), G[f"scipy_dct3_{n}"]) < 1e-13
    # DCT-III_{2/N}(DCT-II(x)) = x and the unitary pair (dct1d.rs:436-462)
    assert rel(oracle.dct(d2, 3, method="forwardn"), x) < 1e-13
    assert rel(oracle.dct(oracle.dct(x, 2, method="forwardu"), 3, method="forwardu"), x) < 1e-13


def test_dct_ortho_is_orthonormal():
    # is_ortho scales element 0 by sqrt(1/2) on input (dct1d.rs:220-224,249,251 and
    # 343-347,369); with forwardu the DCT-II then equals scipy's orthonormal one
    # up to the element-0 convention of the reference, and DCT-III(ortho) inverts it.
    import scipy.fft as sfft
    for n in (4, 10, 64):
        x = uniform_real(n, n)
        y = oracle.dct(x, 3, True, "forwardu")
        assert rel(y, sfft.dct(x, 3, norm="ortho")) < 1e-13
        assert rel(oracle.dct(sfft.dct(x, 2, norm="ortho"), 3, True, "forwardu"), x) < 1e-13


@pytest.mark.parametrize("n", [2048, 4, 8, 12, 64, 396])
def test_mdct_vs_direct(n):
    x = uniform_real(5000 + n, n)
    w = oracle.window(n, "sine")
    assert np.max(np.abs(w - np.sin(np.pi * (np.arange(n) + 0.5) / n))) < 1e-15
    c = oracle.mdct(x, w, "forward")
    assert rel(c, G[f"mdct_{n}"]) < 1e-12
    assert rel(oracle.mdct(c, w, "backward"), G[f"imdct_{n}"]) < 1e-12
    assert rel(oracle.mdct(c, w, "backward"), oracle.imdct_brute(c, w)) < 1e-12


def test_mdct_tdac_reconstruction():
    # overlap-adding backward(forward(frame)) at hop N/2 reconstructs the
    # interior samples (Princen-Bradley with the sine window)
    n, hop = 64, 32
    sig = uniform_real(77, hop * 20 + n)
    w = oracle.window(n, "sine")
    frames = (len(sig) - n) // hop + 1
    c = oracle.mdct(sig, w, "forward", n=n, hop=hop, frames=frames)
    y = oracle.mdct(c, w, "backward")
    ola = np.zeros_like(sig)
    for f in range(frames):
        ola[f * hop:f * hop + n] += y[f]
    inner = slice(hop, (frames - 1) * hop + hop)
    assert rel(ola[inner], sig[inner]) < 1e-13


def test_cfft2d_vs_numpy():
    x = uniform_complex(6000, 35).reshape(7, 5)
    assert rel(oracle.c2c2d(x, "forward"), G["fft2_7x5"]) < 1e-14
    assert rel(oracle.c2c2d(oracle.c2c2d(x, "forward"), "backward"), x) < 1e-14


def test_f32_within_budget():
    # chfft f32 computes its own tables in f32; its error against FP64 stays
    # far inside the FP32 budget of 1e-6 * log2(N)
    for n in (4096, 43200, 4093):
        x = uniform_complex(9, n)
        y = oracle.c2c(x.astype(np.complex64), "forward")
        ref = oracle.c2c(x.astype(np.complex64).astype(np.complex128), "forward")
        assert rel(y, ref) < 1e-6 * np.log2(n)


def test_planted_mistakes_are_caught():
    # the checks above distinguish forward from backward, scaled from unscaled
    # and a transposed 2-D result
    n = 12
    x = uniform_complex(1000 + n, n)
    assert rel(oracle.c2c(x, "backward0"), G[f"dft_fwd_{n}"]) > 0.1
    assert rel(oracle.c2c(x, "forwardn"), G[f"dft_fwd_{n}"]) > 0.1
    x2 = uniform_complex(6000, 35).reshape(7, 5)
    assert rel(oracle.c2c2d(x2.T.copy(), "forward").T, G["fft2_7x5"]) < 1e-14
    assert rel(oracle.c2c2d(x2, "backward0"), G["fft2_7x5"]) > 0.1


# ---------------------------------------------------------------- branches beyond the default methods
# Each check below compares an oracle branch with something the oracle does
# not compute: scipy/numpy transforms, mpmath closed forms, the golden direct
# sums, or an identity, and shows that a plausible mistake in that branch
# (a dropped factor, the other convention) would fail it.

@pytest.mark.parametrize("n", [1024, 2, 6, 12, 30, 126])
def test_dct2_ortho_scales_input_element_zero(n):
    # dct1d.rs:220-224 and 249/251: with is_ortho the DCT-II multiplies the
    # *input* x[0] by sqrt(1/2) before the unnormalised transform; every
    # method then applies its own scale (1, sqrt(2/N), 2/N: dct1d.rs:98-144)
    import scipy.fft as sfft
    x = uniform_real(4000 + n, n)
    xo = x.copy()
    xo[0] *= np.sqrt(0.5)
    ref = sfft.dct(xo, 2) / 2  # scipy's type 2 is twice chfft's unnormalised DCT-II
    for m, s in (("forward", 1.0), ("forward0", 1.0), ("forwardu", np.sqrt(2.0 / n)), ("forwardn", 2.0 / n)):
        assert rel(oracle.dct(x, 2, True, m), ref * s) < 1e-13, m
    # the golden direct sum is linear in x: DCT-II(x') = DCT-II(x) - (1 - sqrt(1/2)) x0 * DCT-II(e0)
    e0 = sfft.dct(np.eye(n)[0], 2) / 2  # = 1 for every bin
    assert rel(oracle.dct(x, 2, True), G[f"dct2_{n}"] - (1 - np.sqrt(0.5)) * x[0] * e0) < 1e-13
    # planted mistakes: ignoring the flag, or scaling the output bin 0 (the
    # orthonormal convention of scipy) instead of the input element 0
    if n > 2:
        assert rel(oracle.dct(x, 2, False), ref) > 1e-3
        out0 = sfft.dct(x, 2) / 2
        out0[0] *= np.sqrt(0.5)
        assert rel(out0, ref) > 1e-3


@pytest.mark.parametrize("n", [4, 8, 64, 396, 2048])
def test_vorbis_window_closed_form(n):
    # mdct1d.rs:46-53: w(l, p) = sin(pi/2 * sin^2(pi (p + 1/2) / l)), pinned
    # against mpmath at 30 digits, in both precisions, and the API mirror's copy
    import mpmath
    from chfft_amd.api import vorbis_window as api_vorbis
    mpmath.mp.dps = 30
    ref = np.array([float(mpmath.sin(mpmath.pi / 2 * mpmath.sin(mpmath.pi * (p + mpmath.mpf(1) / 2) / n) ** 2))
                    for p in range(n)])
    w = oracle.window(n, "vorbis")
    # a few ulps: the reference evaluates the formula in T (f64 / f32)
    assert np.max(np.abs(w - ref)) < 2e-15
    assert np.max(np.abs(oracle.window(n, "vorbis", np.float32) - ref)) < 4e-7
    assert np.max(np.abs(np.array([api_vorbis(n, p) for p in range(n)]) - ref)) < 2e-15
    # Princen-Bradley: w[p]^2 + w[p + n/2]^2 = 1, the condition for TDAC
    h = n // 2
    assert np.max(np.abs(w[:h] ** 2 + w[h:] ** 2 - 1)) < 1e-15
    # planted mistake: the window without the square inside (sin(pi/2 sin(.)))
    bad = np.sin(np.pi / 2 * np.sin(np.pi * (np.arange(n) + 0.5) / n))
    assert np.max(np.abs(bad - ref)) > 1e-2
    assert np.max(np.abs(bad[:h] ** 2 + bad[h:] ** 2 - 1)) > 1e-2


@pytest.mark.parametrize("kind", ["sine", "vorbis"])
def test_mdct_tdac_reconstruction_both_windows(kind):
    # overlap-add of backward(forward(frame)) at hop N/2 gives back the
    # interior samples for any Princen-Bradley window (mdct1d.rs:197-291)
    n, hop = 128, 64
    sig = uniform_real(78, hop * 24 + n)
    w = oracle.window(n, kind)
    frames = (len(sig) - n) // hop + 1
    c = oracle.mdct(sig, w, "forward", n=n, hop=hop, frames=frames)
    y = oracle.mdct(c, w, "backward")
    ola = np.zeros_like(sig)
    for f in range(frames):
        ola[f * hop:f * hop + n] += y[f]
    inner = slice(hop, frames * hop)
    assert rel(ola[inner], sig[inner]) < 1e-13
    # and the windowed brute-force definitions with this window (mdct1d.rs:305-339)
    x = uniform_real(79, n)
    assert rel(oracle.mdct(x, w, "forward"), oracle.mdct_brute(x, w)) < 1e-12
    cb = oracle.mdct_brute(x, w)
    assert rel(oracle.mdct(cb, w, "backward"), oracle.imdct_brute(cb, w)) < 1e-12


@pytest.mark.parametrize("n", [2048, 4, 8, 12, 64, 396])
def test_mdct_unitary_scalings(n):
    # mdct1d.rs:149-195: forward = scaler_ui = sqrt(N/2) times the folded core,
    # which is the unscaled definition (the golden direct sum); forwardu = 1, so
    # forwardu = forward * sqrt(2/N) (the doc's "sqrt(2/n) scaling factor");
    # backward = scaler_u = 1/sqrt(N/2) gives the (4/N) w sum definition;
    # backwardu = 1, so backwardu = backward * sqrt(N/2)
    x = uniform_real(5000 + n, n)
    w = np.sin(np.pi * (np.arange(n) + 0.5) / n)
    h = n // 2
    assert rel(oracle.mdct(x, w, "forwardu"), G[f"mdct_{n}"] * np.sqrt(2.0 / n)) < 1e-12
    c = G[f"mdct_{n}"]
    assert rel(oracle.mdct(c, w, "backwardu"), G[f"imdct_{n}"] * np.sqrt(h)) < 1e-12
    # the unitary pair reconstructs by TDAC too: backwardu(forwardu(.)) = backward(forward(.))
    assert rel(oracle.mdct(oracle.mdct(x, w, "forwardu"), w, "backwardu"),
               oracle.mdct(oracle.mdct(x, w, "forward"), w, "backward")) < 1e-12
    # planted mistake: the reciprocal factor
    if n > 2:
        assert rel(oracle.mdct(x, w, "forward") / np.sqrt(2.0 / n), oracle.mdct(x, w, "forwardu")) > 0.1


@pytest.mark.parametrize("n", [8192, 4096, 10, 6, 2])
def test_rfft_backward_scalings(n):
    # rfft1d.rs:153-187 and 259-312: convert_back doubles the scaler and runs
    # the unscaled inner inverse, so backward (1/N) is numpy's irfft (which,
    # like rfft1d.rs:164-167, ignores Im X[0] and Im X[N/2]); backward0 = N
    # times it, backwardu = sqrt(N) times it, backwardn = backward
    r = np.random.default_rng(n)
    X = r.standard_normal(n // 2 + 1) + 1j * r.standard_normal(n // 2 + 1)
    inv = np.fft.irfft(X, n)
    for m, s in (("backward", 1.0), ("backwardn", 1.0), ("backward0", float(n)), ("backwardu", np.sqrt(n))):
        assert rel(oracle.rfft(X, m, n=n), inv * s) < 1e-13, m
    # planted mistake: the inner inverse without the doubling of convert_back
    assert rel(oracle.rfft(X, "backward0", n=n) / 2, inv * n) > 0.1
    # and the forward family against the golden numpy rfft (forward0 = forward)
    x = uniform_real(3000 + n, n)
    assert rel(oracle.rfft(x, "forward0"), G[f"rfft_{n}"]) < 1e-14


@pytest.mark.parametrize("shape", [(7, 5), (1, 9), (8, 1), (4, 6), (16, 12)])
def test_cfft2d_all_scalings_vs_numpy(shape):
    # cfft2d.rs:150-324: forward = forward0 = 1, forwardu = 1/sqrt(MN),
    # forwardn = 1/(MN), backward = 1/(MN), backward0 = 1, backwardu =
    # 1/sqrt(MN); the 2-D transform is row FFTs then column FFTs (cfft2d.rs:326-363)
    m, n = shape
    x = uniform_complex(6100 + m * 16 + n, m * n).reshape(m, n)
    N = m * n
    F = np.fft.fft2(x)
    B = np.fft.ifft2(x) * N  # unnormalised inverse
    want = {"forward": F, "forward0": F, "forwardu": F / np.sqrt(N), "forwardn": F / N,
            "backward": B / N, "backward0": B, "backwardu": B / np.sqrt(N)}
    for meth, ref in want.items():
        assert rel(oracle.c2c2d(x, meth), ref) < 1e-14, meth
    with pytest.raises(ValueError):
        oracle.c2c2d(x, "backwardn")  # CFft2D has no backwardn
    # the transposed problem gives the transposed spectrum
    assert rel(oracle.c2c2d(x.T.copy(), "forward"), F.T) < 1e-14
    # planted mistakes: the conjugate direction, one axis left untransformed
    if N > 2:
        assert rel(oracle.c2c2d(x, "backward0"), F) > 0.1
    if m > 1:
        assert rel(np.fft.fft(x, axis=1), F) > 0.1
    if n > 1:
        assert rel(np.fft.fft(x, axis=0), F) > 0.1


@pytest.mark.parametrize("n", [4, 12, 64, 396])
def test_brute_force_definitions_vs_golden(n):
    # the oracle's long-double brute-force definitions (the reference's test
    # formulas, mdct1d.rs:305-339 and dct1d.rs:423-434) against the golden
    # float64 direct sums with integer-reduced angles
    x = uniform_real(5000 + n, n)
    w = np.sin(np.pi * (np.arange(n) + 0.5) / n)
    assert rel(oracle.mdct_brute(x, w), G[f"mdct_{n}"]) < 1e-13
    assert rel(oracle.imdct_brute(G[f"mdct_{n}"], w), G[f"imdct_{n}"]) < 1e-13
    if n in (12,):
        xd = uniform_real(4000 + n, n)
        assert rel(oracle.dct2_brute(xd), G[f"dct2_{n}"]) < 1e-13
