"""ctypes front-end for the CPU oracle (liboracle.so, built from chfft_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as the product. The product library (chfft_amd) does not import this.

Every function takes numpy arrays (complex128/complex64 for complex data,
float64/float32 for real data) laid out batch-major, and mirrors the scalings of
the reference methods (cfft1d.rs:270-455, rfft1d.rs:199-312, cfft2d.rs:150-324,
dct1d.rs:98-144, mdct1d.rs:149-195).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FORWARD, FORWARD0, FORWARDU, FORWARDN = 0, 1, 2, 3
BACKWARD, BACKWARD0, BACKWARDU, BACKWARDN = 4, 5, 6, 7
METHODS = {
    "forward": FORWARD, "forward0": FORWARD0, "forwardu": FORWARDU, "forwardn": FORWARDN,
    "backward": BACKWARD, "backward0": BACKWARD0, "backwardu": BACKWARDU, "backwardn": BACKWARDN,
}


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_NATIVE_NOTE = None


def use_native(tmpdir):
    """Build the oracle for this host (-march=native) into tmpdir and use it
    from now on (the CPU baseline). Returns a note saying which build runs."""
    global _LIB, _NATIVE_NOTE
    out = os.path.join(tmpdir, "liboracle_native.so")
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native", f"OUT={out}"], check=True, timeout=300,
                       capture_output=True)
        _LIB = None
        _load(out)
        _NATIVE_NOTE = "built with g++ -O3 -march=native on this host"
    except Exception as e:  # noqa: BLE001 - keep the shipped x86-64-v3 build
        _NATIVE_NOTE = f"shipped -march=x86-64-v3 build (native build failed: {type(e).__name__})"
    return _NATIVE_NOTE


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _load(path)
    return _LIB


def _load(path):
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(path)
        vp, sz, ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        for suf in ("f64", "f32"):
            getattr(L, f"oracle_c2c_{suf}").argtypes = [vp, vp, sz, sz, ci, ci, ci]
            getattr(L, f"oracle_rfft_{suf}").argtypes = [vp, vp, sz, sz, ci, ci]
            getattr(L, f"oracle_c2c2d_{suf}").argtypes = [vp, vp, sz, sz, sz, ci, ci]
            getattr(L, f"oracle_dct_{suf}").argtypes = [vp, vp, sz, sz, ci, ci, ci, ci]
            getattr(L, f"oracle_mdct_{suf}").argtypes = [vp, vp, sz, sz, vp, sz, ci, ci]
            getattr(L, f"oracle_window_{suf}").argtypes = [sz, ci, vp]
        L.oracle_c2c_sequence_f64.argtypes = [vp, vp, vp, sz, ci]
        L.oracle_factors.argtypes = [sz, sz, vp, vp, sz]
        L.oracle_omega_f64.argtypes = [sz, vp]
        L.oracle_bitreverse.argtypes = [sz, vp, vp, sz, vp, ci]
        L.oracle_dft_brute.argtypes = [vp, vp, sz, ci]
        L.oracle_dct2_brute.argtypes = [vp, vp, sz]
        L.oracle_mdct_brute.argtypes = [vp, vp, sz, vp]
        L.oracle_imdct_brute.argtypes = [vp, vp, sz, vp]
        L.oracle_time_c2c.argtypes = [ci, vp, vp, sz, sz, ci, ci, ci]
        L.oracle_time_c2c.restype = ctypes.c_double
        L.oracle_time_op.argtypes = [ci, ci, vp, vp, sz, sz, sz, sz, vp, ci, ci]
        L.oracle_time_op.restype = ctypes.c_double
        L.oracle_time_lens.argtypes = [ci, vp, vp, vp, sz, ci, ci]
        L.oracle_time_lens.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _threads(threads):
    return threads if threads else min(os.cpu_count() or 1, 16)


def _method(m):
    return METHODS[m] if isinstance(m, str) else int(m)


def _suffix(dtype):
    return "f64" if np.dtype(dtype) in (np.complex128, np.float64) else "f32"


def c2c(x, method="forward", inplace=False, threads=0):
    """CFft1D over the last axis of x (complex64 -> f32 path, complex128 -> f64)."""
    x = np.ascontiguousarray(x)
    assert x.dtype in (np.complex64, np.complex128)
    n = x.shape[-1]
    out = np.empty_like(x)
    rc = getattr(lib(), "oracle_c2c_" + _suffix(x.dtype))(
        _p(x), _p(out), n, x.size // max(n, 1), _method(method), int(inplace), _threads(threads))
    if rc:
        raise ValueError("invalid length")
    return out


def c2c_sequence(xs, method="forward"):
    """One CFft1D::new() plan fed signals of changing lengths (replans each change)."""
    lens = np.array([len(x) for x in xs], dtype=np.uint64)
    flat = np.ascontiguousarray(np.concatenate(xs).astype(np.complex128))
    out = np.empty_like(flat)
    rc = lib().oracle_c2c_sequence_f64(_p(flat), _p(out), _p(lens), len(xs), _method(method))
    if rc:
        raise ValueError("invalid length")
    res, off = [], 0
    for n in lens:
        res.append(out[off:off + int(n)])
        off += int(n)
    return res


def rfft(x, method="forward", n=None, threads=0):
    """RFft1D: forward methods take real [..., n], backward take complex [..., n/2+1]."""
    x = np.ascontiguousarray(x)
    m = _method(method)
    back = m >= BACKWARD
    if back:
        assert x.dtype in (np.complex64, np.complex128)
        n = n if n is not None else 2 * (x.shape[-1] - 1)
        if x.shape[-1] != n // 2 + 1:
            raise ValueError("invalid length")
        rdt = np.float64 if x.dtype == np.complex128 else np.float32
        out = np.empty(x.shape[:-1] + (n,), dtype=rdt)
        batch = x.size // x.shape[-1]
    else:
        assert x.dtype in (np.float32, np.float64)
        n = x.shape[-1]
        cdt = np.complex128 if x.dtype == np.float64 else np.complex64
        out = np.empty(x.shape[:-1] + (n // 2 + 1,), dtype=cdt)
        batch = x.size // n
    rc = getattr(lib(), "oracle_rfft_" + _suffix(x.dtype))(_p(x), _p(out), n, batch, m, _threads(threads))
    if rc:
        raise ValueError("invalid length")
    return out


def c2c2d(x, method="forward", threads=0):
    x = np.ascontiguousarray(x)
    m, n = x.shape[-2], x.shape[-1]
    out = np.empty_like(x)
    rc = getattr(lib(), "oracle_c2c2d_" + _suffix(x.dtype))(
        _p(x), _p(out), m, n, x.size // max(m * n, 1), _method(method), _threads(threads))
    if rc:
        raise ValueError("unsupported method")
    return out


def dct(x, dct_type=2, is_ortho=False, method="forward", threads=0):
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    rc = getattr(lib(), "oracle_dct_" + _suffix(x.dtype))(
        _p(x), _p(out), n, x.size // max(n, 1), dct_type, int(is_ortho), _method(method), _threads(threads))
    if rc:
        raise ValueError("invalid length")
    return out


def window(n, kind="sine", dtype=np.float64):
    w = np.empty(n, dtype=dtype)
    getattr(lib(), "oracle_window_" + _suffix(dtype))(n, 0 if kind == "sine" else 1, _p(w))
    return w


def mdct(x, win, method="forward", n=None, hop=None, frames=None, threads=0):
    """Mdct1D. Forward: frames of n samples (independent rows of x, or `frames`
    frames `hop` apart in a 1-D signal) -> n/2 coefficients each. Backward:
    [..., n/2] -> [..., n] windowed samples."""
    x = np.ascontiguousarray(x)
    win = np.ascontiguousarray(win, dtype=x.dtype)
    m = _method(method)
    back = m >= BACKWARD
    if back:
        n = 2 * x.shape[-1]
        batch = x.size // x.shape[-1]
        out = np.empty(x.shape[:-1] + (n,), dtype=x.dtype)
        stride = n // 2
    elif hop is not None:
        batch = frames
        stride = hop
        out = np.empty((frames, n // 2), dtype=x.dtype)
    else:
        n = x.shape[-1]
        batch = x.size // n
        stride = n
        out = np.empty(x.shape[:-1] + (n // 2,), dtype=x.dtype)
    rc = getattr(lib(), "oracle_mdct_" + _suffix(x.dtype))(
        _p(x), _p(out), n, batch, _p(win), stride, m, _threads(threads))
    if rc:
        raise ValueError("invalid length")
    return out


def factors(n, max_prime=7):
    v = np.zeros(64, dtype=np.uint64)
    c = np.zeros(64, dtype=np.uint64)
    k = lib().oracle_factors(n, max_prime, _p(v), _p(c), 64)
    return [(int(v[i]), int(c[i])) for i in range(k)]


def omega(n):
    out = np.empty(n + 1, dtype=np.complex128)
    lib().oracle_omega_f64(n, _p(out))
    return out


def bitreverse(n, facs, inplace=False):
    v = np.array([f[0] for f in facs], dtype=np.uint64)
    c = np.array([f[1] for f in facs], dtype=np.uint64)
    out = np.empty(n, dtype=np.uint64)
    rc = lib().oracle_bitreverse(n, _p(v), _p(c), len(facs), _p(out), int(inplace))
    if rc:
        raise ValueError("factor product differs from n")
    return out


def dft_brute(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().oracle_dft_brute(_p(x), _p(out), len(x), sign)
    return out


def dct2_brute(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().oracle_dct2_brute(_p(x), _p(out), len(x))
    return out


def mdct_brute(x, win):
    x = np.ascontiguousarray(x, dtype=np.float64)
    win = np.ascontiguousarray(win, dtype=np.float64)
    out = np.empty(len(x) // 2)
    lib().oracle_mdct_brute(_p(x), _p(out), len(x), _p(win))
    return out


def imdct_brute(c, win):
    c = np.ascontiguousarray(c, dtype=np.float64)
    win = np.ascontiguousarray(win, dtype=np.float64)
    out = np.empty(2 * len(c))
    lib().oracle_imdct_brute(_p(c), _p(out), 2 * len(c), _p(win))
    return out


def time_c2c(x, method="forward", threads=0, reps=3):
    """Seconds per pass of the CPU baseline over the whole batch in x."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    n = x.shape[-1]
    return lib().oracle_time_c2c(int(x.dtype == np.complex128), _p(x), _p(out), n, x.size // n,
                                 _method(method), _threads(threads), reps)


def time_op(kind, x, n, threads=0, reps=3, m=1, batch=None, hop=None, window=None):
    """Seconds per pass of the CPU baseline for the other transform types
    (plans built before the clock): kind "rfft" (x real [batch, n]), "dct2",
    "dct3" (x real [batch, n]), "mdct" (x a 1-D signal, `batch` frames `hop`
    apart, `window` of n entries), "c2c2d" (x complex [batch, m, n])."""
    code = {"rfft": 1, "dct2": 2, "dct3": 3, "mdct": 4, "c2c2d": 5}[kind]
    x = np.ascontiguousarray(x)
    f64 = x.dtype in (np.float64, np.complex128)
    rdt = np.float64 if f64 else np.float32
    if kind == "rfft":
        batch = x.size // n
        out = np.empty((batch, n // 2 + 1), dtype=np.complex128 if f64 else np.complex64)
    elif kind == "mdct":
        out = np.empty((batch, n // 2), dtype=rdt)
    else:
        batch = x.size // (n * m)
        out = np.empty_like(x)
    w = None if window is None else np.ascontiguousarray(window, dtype=rdt)
    return lib().oracle_time_op(code, int(f64), _p(x), _p(out), n, m, batch, hop or n,
                                None if w is None else _p(w), _threads(threads), reps)


def time_lens(x, lens, threads=0, reps=3):
    """Seconds per pass of the CPU baseline over transforms of uneven lengths
    packed back to back in x (CFft1D::forward of each, plans built first)."""
    x = np.ascontiguousarray(x)
    lv = np.ascontiguousarray(np.asarray(lens, dtype=np.uint64))
    out = np.empty_like(x)
    return lib().oracle_time_lens(int(x.dtype == np.complex128), _p(x), _p(out), _p(lv), len(lv),
                                  _threads(threads), reps)
