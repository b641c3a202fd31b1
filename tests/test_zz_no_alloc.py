"""GPU test of chfft_exec_noalloc (the entry point that allocates nothing:
its four-step scratch and work buffer live in a library-owned device array).
Called through the C ABI and compared with chfft_exec on the same inputs:
the same kernels run, so the outputs must agree bit for bit. Two calls in a
row, the second with a smaller plan, then plans of the other kinds that use
scratch or work memory, and the limit."""
import ctypes

import numpy as np
import pytest

from chfft_amd import MultiFFT, Plan, _lib
from tests.gpu_util import rng

pytestmark = pytest.mark.gpu


def run_both(plan, x, out_numel, out_dtype, method="forward"):
    import torch
    L = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    y_ref = torch.full((out_numel,), float("nan"), dtype=out_dtype, device="cuda")
    y_na = torch.full((out_numel,), float("nan"), dtype=out_dtype, device="cuda")
    m = _lib.METHODS[method]
    _lib.check(L.chfft_exec(plan.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y_ref.data_ptr()), m,
                            ctypes.c_void_p(st)), "exec")
    rc = L.chfft_exec_noalloc(plan.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y_na.data_ptr()), m,
                              ctypes.c_void_p(st))
    _lib.check(rc, "exec_noalloc")
    torch.cuda.synchronize()
    return y_ref, y_na


def cuda_normal(numel, dtype, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.empty(numel, dtype=dtype, device="cuda").normal_(generator=g)


@pytest.mark.parametrize("f64", [False, True])
def test_noalloc_matches_exec_two_calls(f64):
    import torch
    cdt = torch.complex128 if f64 else torch.complex64
    # first call: the headline four-step (43200 x 512, scratch of a 2- or
    # 4-chunk split: 88.5 MB, within the 96 MiB array)
    n1, b1 = 43200, 512
    p1 = Plan("c2c", f64, b1, n1)
    assert 0 < p1.scratch_bytes() <= _lib.NOALLOC_SCRATCH_BYTES
    x1 = cuda_normal(n1 * b1, cdt, 1)
    for method in ("forward", "backward"):
        a, b = run_both(p1, x1, n1 * b1, cdt, method)
        assert torch.equal(a, b), method
    # second call, smaller: a four-step Bluestein (20011 x 16), whose scratch
    # overwrites the start of what the first call left
    n2, b2 = 20011, 16
    p2 = Plan("c2c", f64, b2, n2)
    assert 0 < p2.scratch_bytes() < p1.scratch_bytes()
    x2 = cuda_normal(n2 * b2, cdt, 2)
    a, b = run_both(p2, x2, n2 * b2, cdt, "forward")
    assert torch.equal(a, b)
    xs = x2.cpu().numpy().reshape(b2, n2)[:2].astype(np.complex128)
    ref = np.fft.fft(xs)
    got = b.cpu().numpy().reshape(b2, n2)[:2]
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert e < (1e-12 if f64 else 2e-5), e


def test_noalloc_work_buffer_and_multi():
    import torch
    # a long-length RFft1D (four-step scratch + work buffer in the array)
    n, batch = 65536, 8
    p = Plan("r2c", False, batch, n)
    x = cuda_normal(n * batch, torch.float32, 3)
    a, b = run_both(p, x, (n // 2 + 1) * batch, torch.complex64, "forward")
    assert torch.equal(a, b)
    X = a.clone()
    a, b = run_both(p, X, n * batch, torch.float32, "backward")
    assert torch.equal(a, b)
    # a multi-length plan with four-step items (per-item scratch slots)
    lens = [20011, 4096, 97, 100000, 1, 3000]
    f = MultiFFT(lens, np.float32)
    xm = cuda_normal(f.total, torch.complex64, 4)
    a, b = run_both(f.plan, xm, f.total, torch.complex64, "forward")
    f.plan.check()
    assert torch.equal(a, b)


def test_noalloc_rejects_plans_beyond_the_limit():
    import torch
    p = Plan("c2c", True, 512, 10007)  # FP64 Bluestein over M = 20480: 168 MB of scratch
    assert p.scratch_bytes() > _lib.NOALLOC_SCRATCH_BYTES
    x = torch.zeros(512 * 10007, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    rc = _lib.lib().chfft_exec_noalloc(p.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                       _lib.METHODS["forward"], ctypes.c_void_p(0))
    assert rc == _lib.ERR_SCRATCH_LIMIT
    assert _lib.lib().chfft_status_string(rc).startswith(b"the plan's scratch exceeds")
