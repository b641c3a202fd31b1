"""Python mirror of chfft's public API over the C ABI (include/chfft_hip.h).

The five reference types keep their names, constructors, methods, scalings and
"invalid length" errors (lib.rs:21-25):

  CFft1D  cfft1d.rs:53-468   new / with_len / setup, forward{,0,u,n},
                             backward{,0,u,n}, in place forward0i / backward0i /
                             forwardui / backwardui
  RFft1D  rfft1d.rs:32-313   new / setup, forward{,0,u,n}, backward{,0,u,n}
  CFft2D  cfft2d.rs:57-376   new / with_len / setup, forward{,0,u,n},
                             backward{,0,u}
  Dct1D   dct1d.rs:30-161    new / setup, forward{,0,u,n}; DctType.Two/Three
  Mdct1D  mdct1d.rs:30-195   new / with_sine / with_vorbis / setup,
                             forward{,u}, backward{,u}

`dtype` plays the role of the reference's type parameter T (np.float64 = the
plain FP64 method, np.float32 = the FP32 form). Inputs are numpy arrays (host:
copied through the plan's staging buffers) or torch CUDA tensors (device:
asynchronous on torch's current stream). A leading batch dimension is allowed
everywhere: [..., n] transforms each row. Every call runs the HIP kernels;
there is no CPU path.
"""
import ctypes
import enum
import math

import numpy as np

from . import _lib

__all__ = ["CFft1D", "RFft1D", "CFft2D", "Dct1D", "DctType", "Mdct1D", "MultiFFT", "Plan", "sine_window",
           "vorbis_window"]


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _torch():
    import torch
    return torch


def _real_dtype(dtype):
    dt = np.dtype(dtype)
    if dt not in (np.float32, np.float64):
        raise TypeError("dtype must be np.float32 or np.float64")
    return dt


class Plan:
    """Owning handle of one chfft_plan (a length or shape plus a batch count)."""

    def __init__(self, kind, f64, batch, *args, tuned=True):
        """`tuned`: apply the tuning table's spec for this shape, if it has
        one (chfft_amd.tuning)."""
        self.kind = kind
        self.f64 = f64
        self.batch = batch
        self.args = args
        h = ctypes.c_void_p()
        L = _lib.lib()
        prec = _lib.F64 if f64 else _lib.F32
        if kind == "c2c":
            st = L.chfft_plan_c2c_1d(ctypes.byref(h), args[0], batch, prec)
        elif kind == "r2c":
            st = L.chfft_plan_r2c_1d(ctypes.byref(h), args[0], batch, prec)
        elif kind == "c2c2d":
            st = L.chfft_plan_c2c_2d(ctypes.byref(h), args[0], args[1], batch, prec)
        elif kind == "dct":
            st = L.chfft_plan_dct_1d(ctypes.byref(h), args[0], args[1], int(args[2]), batch, prec)
        elif kind == "mdct":
            win = np.ascontiguousarray(args[1], dtype=np.float64)
            st = L.chfft_plan_mdct_1d(ctypes.byref(h), args[0], win.ctypes.data, batch, args[2], prec)
        elif kind == "c2c_multi":
            lens = np.ascontiguousarray(args[0], dtype=np.uint64)
            st = L.chfft_plan_c2c_multi(ctypes.byref(h), lens.ctypes.data, len(lens), prec)
        else:
            raise ValueError(kind)
        _lib.check(st, f"plan {kind}{args}")
        self._h = h
        self.tuned_spec = None
        if tuned and kind in ("c2c", "c2c_multi"):
            from . import tuning
            self.tuned_spec = tuning.apply(self)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._LIB is not None:
            _lib.lib().chfft_plan_destroy(h)
            self._h = None

    def io_bytes(self, method):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(_lib.lib().chfft_plan_io_bytes(self._h, _lib.METHODS[method], ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def scratch_bytes(self):
        return _lib.lib().chfft_plan_scratch_bytes(self._h)

    def describe(self):
        buf = ctypes.create_string_buffer(65536)
        _lib.lib().chfft_plan_describe(self._h, buf, 65536)
        return buf.value.decode()

    def autotune(self, src, dst, reps=3, stream=0):
        """Time this plan's alternative schedules on device buffers (ints or
        torch tensors; contents are overwritten) and keep the fastest."""
        if _is_torch(src):
            stream = stream or _torch().cuda.current_stream(src.device).cuda_stream
            src, dst = src.data_ptr(), dst.data_ptr()
        r = _lib.lib().chfft_plan_autotune(self._h, ctypes.c_void_p(src), ctypes.c_void_p(dst), reps,
                                           ctypes.c_void_p(stream))
        if r < 0:
            _lib.check(-r, "autotune")
        return r

    def tune(self, spec):
        """Rebuild this plan under planner overrides (chfft_plan_tune), e.g.
        "generic=1 order=2 tpb=4 tpt=64" or "n1=180 chunk=-1"; "default" is
        the planner's own choice. Not while an exec of the plan is running."""
        _lib.check(_lib.lib().chfft_plan_tune(self._h, spec.encode()), "tune")

    def tuning(self):
        """(spec, report): the overrides the plan runs with and the last
        autotune's candidate times (chfft_plan_tuning)."""
        buf = ctypes.create_string_buffer(16384)
        _lib.check(_lib.lib().chfft_plan_tuning(self._h, buf, 16384), "tuning")
        spec, _, report = buf.value.decode().partition("\n")
        return spec, report

    def check(self, stream=0):
        """Wait for `stream` and raise ChfftError if a multi-length launch of
        this plan since the last check did not complete (chfft_plan_check)."""
        _lib.check(_lib.lib().chfft_plan_check(self._h, ctypes.c_void_p(stream)), "multi-length launch")

    def exec_ptr(self, src, dst, method, stream=0):
        """Device pointers (ints); asynchronous on `stream` (hipStream_t as int)."""
        _lib.check(_lib.lib().chfft_exec(self._h, ctypes.c_void_p(src), ctypes.c_void_p(dst),
                                         _lib.METHODS[method], ctypes.c_void_p(stream)), "exec")

    def _dtypes(self, method):
        """(input, output) element types of `method`: complex or real, in the
        plan's precision."""
        back = _lib.METHODS[method] >= _lib.METHODS["backward"]
        cin = {"c2c": True, "c2c2d": True, "c2c_multi": True, "r2c": back, "dct": False, "mdct": False}[self.kind]
        cout = {"c2c": True, "c2c2d": True, "c2c_multi": True, "r2c": not back, "dct": False,
                "mdct": False}[self.kind]
        return cin, cout

    def _check_tensor(self, t, complex_want, what):
        torch = _torch()
        want = (torch.complex128 if self.f64 else torch.complex64) if complex_want else \
            (torch.float64 if self.f64 else torch.float32)
        # complex data may also come as its interleaved real view
        ok = t.dtype == want or (complex_want and t.dtype == (torch.float64 if self.f64 else torch.float32))
        if not ok:
            raise ValueError(f"{what}: dtype {t.dtype} does not match the plan ({want})")
        if not t.is_contiguous():
            raise ValueError(f"{what}: tensor must be contiguous")

    def exec(self, src, dst, method):
        """numpy arrays (host, blocking) or torch CUDA tensors (device, async).
        Buffers must be contiguous, of the plan's precision, and hold at least
        the plan's batch."""
        cin, cout = self._dtypes(method)
        if _is_torch(src):
            torch = _torch()
            if not (src.is_cuda and dst.is_cuda):
                raise ValueError("torch tensors must live on the GPU")
            self._check_tensor(src, cin, "input")
            self._check_tensor(dst, cout, "output")
            ib, ob = self.io_bytes(method)
            if src.numel() * src.element_size() < ib or dst.numel() * dst.element_size() < ob:
                raise ValueError("invalid length (buffer smaller than the plan's batch)")
            stream = torch.cuda.current_stream(src.device).cuda_stream
            self.exec_ptr(src.data_ptr(), dst.data_ptr(), method, stream)
        else:
            fdt = np.float64 if self.f64 else np.float32
            for a, c, what in ((src, cin, "input"), (dst, cout, "output")):
                want = (np.complex128 if self.f64 else np.complex64) if c else fdt
                if a.dtype != want and not (c and a.dtype == fdt):
                    raise ValueError(f"{what}: dtype {a.dtype} does not match the plan ({np.dtype(want)})")
                if not a.flags.c_contiguous:
                    raise ValueError(f"{what}: array must be C-contiguous")
            ib, ob = self.io_bytes(method)
            if src.nbytes < ib or dst.nbytes < ob:
                raise ValueError("invalid length (buffer smaller than the plan's batch)")
            _lib.check(_lib.lib().chfft_exec_host(self._h, src.ctypes.data, dst.ctypes.data,
                                                  _lib.METHODS[method]), "exec_host")


class _Base:
    def __init__(self, dtype):
        self.dtype = _real_dtype(dtype)
        self._plan = None
        self._plan_key = None

    @property
    def f64(self):
        return self.dtype == np.float64

    @property
    def cdtype(self):
        return np.complex128 if self.f64 else np.complex64

    def _get_plan(self, key, make):
        if self._plan_key != key:
            self._plan = None
            self._plan = make()
            self._plan_key = key
        return self._plan

    def _prep(self, x, complex_in):
        """Contiguous input of the plan's precision; returns (array, is_torch)."""
        if _is_torch(x):
            torch = _torch()
            want = (torch.complex128 if self.f64 else torch.complex64) if complex_in else \
                (torch.float64 if self.f64 else torch.float32)
            return x.to(want).contiguous(), True
        want = self.cdtype if complex_in else self.dtype
        return np.ascontiguousarray(np.asarray(x, dtype=want)), False

    def _empty(self, like, shape, complex_out, is_torch):
        if is_torch:
            torch = _torch()
            dt = (torch.complex128 if self.f64 else torch.complex64) if complex_out else \
                (torch.float64 if self.f64 else torch.float32)
            return torch.empty(shape, dtype=dt, device=like.device)
        return np.empty(shape, dtype=self.cdtype if complex_out else self.dtype)


def _batch(shape):
    b = 1
    for s in shape[:-1]:
        b *= s
    return b


class CFft1D(_Base):
    """Complex-to-complex 1-D transform (cfft1d.rs:53-468)."""

    def __init__(self, dtype=np.float64):
        super().__init__(dtype)
        self.len = 0

    @classmethod
    def new(cls, dtype=np.float64):
        return cls(dtype)

    @classmethod
    def with_len(cls, len, dtype=np.float64):
        f = cls(dtype)
        f.setup(len)
        return f

    def setup(self, len):
        self.len = len
        self._plan = None
        self._plan_key = None

    # cfft1d.rs:208-230 / 232-256: length <= 1 returns the input unchanged;
    # a different length replans.
    def _convert(self, source, method, inplace=False):
        x, tch = self._prep(source, True)
        n = x.shape[-1]
        if n <= 1:
            if inplace:
                return source
            return x.clone() if tch else x.copy()
        if n != self.len:
            self.setup(n)
        batch = _batch(x.shape)
        plan = self._get_plan((n, batch), lambda: Plan("c2c", self.f64, batch, n))
        if inplace:
            plan.exec(x, x, method)
            if x is not source:
                if tch:
                    source.copy_(x)
                else:
                    source[...] = x
            return source
        out = self._empty(x, x.shape, True, tch)
        plan.exec(x, out, method)
        return out

    def forward(self, source):
        return self._convert(source, "forward")

    def forward0(self, source):
        return self._convert(source, "forward0")

    def forwardu(self, source):
        return self._convert(source, "forwardu")

    def forwardn(self, source):
        return self._convert(source, "forwardn")

    def backward(self, source):
        return self._convert(source, "backward")

    def backward0(self, source):
        return self._convert(source, "backward0")

    def backwardu(self, source):
        return self._convert(source, "backwardu")

    def backwardn(self, source):
        return self._convert(source, "backwardn")

    def forward0i(self, source):
        self._convert(source, "forward0", inplace=True)

    def backward0i(self, source):
        self._convert(source, "backward0", inplace=True)

    def forwardui(self, source):
        self._convert(source, "forwardu", inplace=True)

    def backwardui(self, source):
        self._convert(source, "backwardu", inplace=True)


class RFft1D(_Base):
    """Real-to-complex 1-D transform of even length (rfft1d.rs:32-313)."""

    def __init__(self, len, dtype=np.float64):
        super().__init__(dtype)
        self.setup(len)

    @classmethod
    def new(cls, len, dtype=np.float64):
        return cls(len, dtype)

    def setup(self, len):
        if len & 1 != 0:
            raise ValueError("invalid length")
        self.len = len
        self._plan = None
        self._plan_key = None

    def _plan_for(self, batch):
        return self._get_plan(batch, lambda: Plan("r2c", self.f64, batch, self.len))

    def _convert(self, source, method):
        x, tch = self._prep(source, False)
        if x.shape[-1] != self.len:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, rdft.len: {self.len})")
        batch = _batch(x.shape)
        out = self._empty(x, tuple(x.shape[:-1]) + (self.len // 2 + 1,), True, tch)
        self._plan_for(batch).exec(x, out, method)
        return out

    def _convert_back(self, source, method):
        x, tch = self._prep(source, True)
        if x.shape[-1] != (self.len >> 1) + 1:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, rdft.len: {self.len})")
        batch = _batch(x.shape)
        out = self._empty(x, tuple(x.shape[:-1]) + (self.len,), False, tch)
        self._plan_for(batch).exec(x, out, method)
        return out

    def forward(self, source):
        return self._convert(source, "forward")

    def forward0(self, source):
        return self._convert(source, "forward0")

    def forwardu(self, source):
        return self._convert(source, "forwardu")

    def forwardn(self, source):
        return self._convert(source, "forwardn")

    def backward(self, source):
        return self._convert_back(source, "backward")

    def backward0(self, source):
        return self._convert_back(source, "backward0")

    def backwardu(self, source):
        return self._convert_back(source, "backwardu")

    def backwardn(self, source):
        return self._convert_back(source, "backwardn")


class CFft2D(_Base):
    """Complex 2-D transform of len_m rows x len_n columns (cfft2d.rs:57-376).

    Input is a 2-D array [len_m, len_n] (or a batch [..., len_m, len_n]); a list
    of equal-length rows, the reference's &[Vec<Complex<T>>], is accepted too.
    """

    def __init__(self, dtype=np.float64):
        super().__init__(dtype)
        self.len_m = 0
        self.len_n = 0

    @classmethod
    def new(cls, dtype=np.float64):
        return cls(dtype)

    @classmethod
    def with_len(cls, len_m, len_n, dtype=np.float64):
        f = cls(dtype)
        f.setup(len_m, len_n)
        return f

    def setup(self, len_m, len_n):
        self.len_m = len_m
        self.len_n = len_n
        self._plan = None
        self._plan_key = None

    def _convert(self, source, method):
        if not _is_torch(source) and isinstance(source, (list, tuple)):
            if len(source) == 0:
                return []  # cfft2d.rs:333-335
            source = np.array(source)
        x, tch = self._prep(source, True)
        if x.ndim < 2:
            raise ValueError("invalid length (expected rows x columns)")
        m, n = x.shape[-2], x.shape[-1]
        if m == 0 or n == 0:
            return x.clone() if tch else x.copy()
        if m != self.len_m or n != self.len_n:
            self.setup(m, n)  # cfft2d.rs:336-338
        batch = 1
        for s in x.shape[:-2]:
            batch *= s
        plan = self._get_plan((m, n, batch), lambda: Plan("c2c2d", self.f64, batch, m, n))
        out = self._empty(x, x.shape, True, tch)
        plan.exec(x, out, method)
        return out

    def forward(self, source):
        return self._convert(source, "forward")

    def forward0(self, source):
        return self._convert(source, "forward0")

    def forwardu(self, source):
        return self._convert(source, "forwardu")

    def forwardn(self, source):
        return self._convert(source, "forwardn")

    def backward(self, source):
        return self._convert(source, "backward")

    def backward0(self, source):
        return self._convert(source, "backward0")

    def backwardu(self, source):
        return self._convert(source, "backwardu")


class DctType(enum.Enum):
    """dct1d.rs:35-44"""
    Two = 2
    Three = 3


class Dct1D(_Base):
    """DCT-II / DCT-III of even length (dct1d.rs:30-161)."""

    def __init__(self, dct_type, len, is_ortho, dtype=np.float64):
        super().__init__(dtype)
        if len & 1 != 0:
            raise ValueError("invalid length")
        self.dct_type = DctType(dct_type)
        self.len = len
        self.is_ortho = bool(is_ortho)

    @classmethod
    def new(cls, dct_type, len, is_ortho, dtype=np.float64):
        return cls(dct_type, len, is_ortho, dtype)

    def setup(self, dct_type, len, is_ortho):
        if len & 1 != 0:
            raise ValueError("invalid length")
        self.dct_type = DctType(dct_type)
        self.len = len
        self.is_ortho = bool(is_ortho)
        self._plan = None
        self._plan_key = None

    def _convert(self, source, method):
        x, tch = self._prep(source, False)
        if x.shape[-1] != self.len:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, dct.len: {self.len})")
        batch = _batch(x.shape)
        plan = self._get_plan(batch, lambda: Plan("dct", self.f64, batch, self.dct_type.value, self.len,
                                                   self.is_ortho))
        out = self._empty(x, x.shape, False, tch)
        plan.exec(x, out, method)
        return out

    def forward(self, source):
        return self._convert(source, "forward")

    def forward0(self, source):
        return self._convert(source, "forward0")

    def forwardu(self, source):
        return self._convert(source, "forwardu")

    def forwardn(self, source):
        return self._convert(source, "forwardn")


def sine_window(l, p):
    """mdct1d.rs:42-44"""
    return math.sin(math.pi * (p + 0.5) / l)


def vorbis_window(l, p):
    """mdct1d.rs:46-53"""
    s = math.sin(math.pi * (p + 0.5) / l)
    return math.sin(s * s * math.pi * 0.5)


class Mdct1D(_Base):
    """MDCT / IMDCT with a window (mdct1d.rs:30-195). len % 4 == 0."""

    def __init__(self, window_func, len, dtype=np.float64):
        super().__init__(dtype)
        self.window_func = window_func
        self.setup(len)

    @classmethod
    def new(cls, window_func, len, dtype=np.float64):
        return cls(window_func, len, dtype)

    @classmethod
    def with_sine(cls, len, dtype=np.float64):
        return cls(sine_window, len, dtype)

    @classmethod
    def with_vorbis(cls, len, dtype=np.float64):
        return cls(vorbis_window, len, dtype)

    def setup(self, len):
        if len & 3 != 0:
            raise ValueError("invalid length")
        self.len = len
        # mdct1d.rs:91-93: the window closure is evaluated once into a table
        self.window = np.array([self.window_func(len, i) for i in range(len)], dtype=np.float64)
        self._plan = None
        self._plan_key = None

    def _run(self, x, tch, batch, stride, out_shape, method):
        plan = self._get_plan((batch, stride), lambda: Plan("mdct", self.f64, batch, self.len, self.window, stride))
        out = self._empty(x, out_shape, False, tch)
        plan.exec(x, out, method)
        return out

    def _convert(self, source, method):
        x, tch = self._prep(source, False)
        if x.shape[-1] != self.len:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, mdct.len: {self.len})")
        return self._run(x, tch, _batch(x.shape), self.len, tuple(x.shape[:-1]) + (self.len // 2,), method)

    def _convert_back(self, source, method):
        x, tch = self._prep(source, False)
        if x.shape[-1] << 1 != self.len:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, mdct.len: {self.len})")
        return self._run(x, tch, _batch(x.shape), self.len, tuple(x.shape[:-1]) + (self.len,), method)

    def forward(self, source):
        return self._convert(source, "forward")

    def backward(self, source):
        return self._convert_back(source, "backward")

    def forwardu(self, source):
        return self._convert(source, "forwardu")

    def backwardu(self, source):
        return self._convert_back(source, "backwardu")

    def forward_signal(self, signal, hop, method="forward"):
        """All full frames of a 1-D signal, frame b starting at sample b*hop
        (the caller-side framing the reference leaves to its user)."""
        x, tch = self._prep(signal, False)
        frames = (x.shape[-1] - self.len) // hop + 1
        if x.ndim != 1 or frames < 1:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, mdct.len: {self.len})")
        return self._run(x, tch, frames, hop, (frames, self.len // 2), method)

    def backward_signal(self, coefs, hop, method="backward"):
        """Inverse of forward_signal's framing: backward (or backwardu) of each
        row of coefs [frames, len/2], overlap-added on the device at `hop`
        (chfft_exec_mdct_ola); returns (frames - 1) * hop + len samples. With
        hop = len/2 the samples len/2 .. (frames - 1) * hop + len/2 reproduce
        the analysed signal (TDAC)."""
        x, tch = self._prep(coefs, False)
        if x.ndim != 2 or x.shape[-1] << 1 != self.len:
            raise ValueError(f"invalid length (soure: {x.shape[-1]}, mdct.len: {self.len})")
        frames = x.shape[0]
        plan = self._get_plan((frames, hop), lambda: Plan("mdct", self.f64, frames, self.len, self.window, hop))
        out = self._empty(x, ((frames - 1) * hop + self.len,), False, tch)
        if tch:
            stream = _torch().cuda.current_stream(x.device).cuda_stream
            _lib.check(_lib.lib().chfft_exec_mdct_ola(plan.handle, ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(out.data_ptr()), _lib.METHODS[method],
                                                      ctypes.c_void_p(stream)), "exec_mdct_ola")
            return out
        torch = _torch()
        xd = torch.from_numpy(x).cuda()
        od = torch.empty(out.shape, dtype=xd.dtype, device=xd.device)
        _lib.check(_lib.lib().chfft_exec_mdct_ola(plan.handle, ctypes.c_void_p(xd.data_ptr()),
                                                  ctypes.c_void_p(od.data_ptr()), _lib.METHODS[method],
                                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "exec_mdct_ola")
        return od.cpu().numpy()


class MultiFFT(_Base):
    """Many complex transforms of uneven lengths in one persistent launch
    (chfft_plan_c2c_multi). Each transform behaves as CFft1D of its own length
    (cfft1d.rs:208-230, including the scalings per method); the signals are
    packed back to back in one 1-D array of sum(lens) elements."""

    def __init__(self, lens, dtype=np.float32):
        super().__init__(dtype)
        self.lens = [int(n) for n in lens]
        self.total = sum(self.lens)
        self.offsets = np.concatenate([[0], np.cumsum(self.lens)]).astype(np.int64)
        self.plan = Plan("c2c_multi", self.f64, 1, self.lens)

    def split(self, packed):
        return [packed[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.lens))]

    def _convert(self, packed, method):
        x, tch = self._prep(packed, True)
        if x.ndim != 1 or x.shape[0] != self.total:
            raise ValueError(f"invalid length (soure: {x.shape[0]}, total: {self.total})")
        out = self._empty(x, x.shape, True, tch)
        self.plan.exec(x, out, method)
        if tch:
            # the launch's grid-wide barriers are bounded waits: make a
            # timed-out one an error rather than a silently incomplete output
            self.plan.check(_torch().cuda.current_stream(x.device).cuda_stream)
        return out

    def forward(self, packed):
        return self._convert(packed, "forward")

    def forwardu(self, packed):
        return self._convert(packed, "forwardu")

    def forwardn(self, packed):
        return self._convert(packed, "forwardn")

    def backward(self, packed):
        return self._convert(packed, "backward")

    def backward0(self, packed):
        return self._convert(packed, "backward0")

    def backwardu(self, packed):
        return self._convert(packed, "backwardu")
