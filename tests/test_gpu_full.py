"""GPU tests at the benchmark workloads' full sizes, the long-length paths
(inner transforms above one workgroup's LDS), the MDCT stream synthesis,
large non-square 2-D transforms and the multi-length launch's failure
report.

Full-size batches are checked two ways: spot transforms (first, last and
chunk-boundary indices) against the oracle (numpy's FP64 FFT above 70000
points), and size-independent properties over the whole batch computed on
the device (round trips, DCT-III(DCT-II) = x, MDCT TDAC reconstruction,
Parseval per transform). Bars as in the other parity tests: FP64 rel-L2 <=
1e-12 against the FP64 oracle, FP32 <= 1e-6 log2 N.
"""
import numpy as np
import pytest

import oracle
from chfft_amd import CFft1D, CFft2D, Dct1D, DctType, Mdct1D, MultiFFT, Plan, RFft1D
from tests.gpu_util import TOL_F64, rel_l2, rng, tol_f32

pytestmark = pytest.mark.gpu


def check(gpu, ref, f64, n, what=""):
    e = rel_l2(gpu, ref)
    tol = TOL_F64 if f64 else tol_f32(n)
    assert e <= tol, f"{what} rel-L2 {e:.3e} > {tol:.1e} (n={n})"
    return e


def dev_rel(a, b):
    """rel-L2 of two device tensors, in float64 on the device."""
    import torch
    a = a.to(torch.complex128) if a.is_complex() else a.to(torch.float64)
    b = b.to(torch.complex128) if b.is_complex() else b.to(torch.float64)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def cuda_uniform(shape, complex_, f64, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    dt = torch.float64 if f64 else torch.float32
    n = int(np.prod(shape)) * (2 if complex_ else 1)
    x = torch.empty(n, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)
    if complex_:
        x = x.view(torch.complex128 if f64 else torch.complex64)
    return x.view(*shape)


def c2c_ref(x, method="forward"):
    n = x.shape[-1]
    if n <= 70000:
        return oracle.c2c(x.astype(np.complex128), method)
    y = np.fft.fft(x.astype(np.complex128)) if method == "forward" else np.fft.ifft(x.astype(np.complex128))
    return y


# ---------------------------------------------------------------- c3: Bluestein primes at full batch
@pytest.mark.parametrize("n,batch", [(4093, 2048), (10007, 512), (65537, 64)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bluestein_full_batch(n, batch, dtype):
    import torch
    f64 = dtype == np.float64
    x = cuda_uniform((batch, n), True, f64, 0x5EED0003 + n)
    p = Plan("c2c", f64, batch, n)
    y = torch.empty_like(x)
    p.exec(x, y, "forward")
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    ys = y.cpu().numpy()
    for b in sorted({0, 1, batch // 2 - 1, batch // 2, batch - 2, batch - 1}):
        check(ys[b], c2c_ref(xs[b]), f64, n, f"transform {b}")
    # whole batch: backward(forward(x)) = x on the device
    z = torch.empty_like(x)
    p.exec(y, z, "backward")
    torch.cuda.synchronize()
    e = dev_rel(z, x)
    assert e <= (TOL_F64 if f64 else 2 * tol_f32(n)), e


# ---------------------------------------------------------------- c4: real and cosine transforms at full batch
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rfft_full_batch(dtype):
    import torch
    f64 = dtype == np.float64
    n, batch = 8192, 4096
    x = cuda_uniform((batch, n), False, f64, 0x5EED0004)
    f = RFft1D.new(n, dtype)
    X = f.forward(x)
    xb = f.backward(X)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    Xs = X.cpu().numpy()
    for b in (0, 1, 2047, 2048, 4095):
        check(Xs[b], oracle.rfft(xs[b].astype(np.float64), "forward"), f64, n, f"row {b}")
    e = dev_rel(xb, x)
    assert e <= (TOL_F64 if f64 else 2 * tol_f32(n)), e


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dct_full_batch(dtype):
    import torch
    f64 = dtype == np.float64
    n, batch = 1024, 16384
    x = cuda_uniform((batch, n), False, f64, 0x5EED0014)
    d2 = Dct1D.new(DctType.Two, n, False, dtype)
    d3 = Dct1D.new(DctType.Three, n, False, dtype)
    y2 = d2.forward(x)
    y3 = d3.forward(x)
    back = d3.forwardn(y2)  # DCT-III_{2/N}(DCT-II(x)) = x (dct1d.rs:436-462)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    y2s, y3s = y2.cpu().numpy(), y3.cpu().numpy()
    for b in (0, 1, 8191, 8192, 16383):
        xr = xs[b].astype(np.float64)
        check(y2s[b], oracle.dct(xr, 2), f64, n, f"DCT-II row {b}")
        check(y3s[b], oracle.dct(xr, 3), f64, n, f"DCT-III row {b}")
    e = dev_rel(back, x)
    assert e <= (TOL_F64 if f64 else 2 * tol_f32(n)), e


@pytest.mark.parametrize("kind", ["sine", "vorbis"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mdct_stream_full_size(dtype, kind):
    # the MDCT workload: N = 2048, hop 1024 over 2^24 samples (16383 frames);
    # spot frames against the oracle, and the device overlap-add of the
    # inverse reconstructs the interior of the signal (TDAC)
    import torch
    f64 = dtype == np.float64
    n, hop, S = 2048, 1024, 1 << 24
    r = np.random.default_rng(0x5EED0004)
    t = np.arange(S)
    sig = sum(r.uniform(0.1, 1) * np.sin(2 * np.pi * r.uniform(0, 0.5) * t + r.uniform(0, 6.3))
              for _ in range(8)) + r.normal(0, 0.01, S)
    sig = sig.astype(dtype)
    md = Mdct1D.with_sine(n, dtype) if kind == "sine" else Mdct1D.with_vorbis(n, dtype)
    xs = torch.from_numpy(sig).cuda()
    c = md.forward_signal(xs, hop)
    frames = c.shape[0]
    assert frames == (S - n) // hop + 1
    y = md.backward_signal(c, hop)
    torch.cuda.synchronize()
    assert y.shape[0] == (frames - 1) * hop + n == S
    cs = c.cpu().numpy()
    w = oracle.window(n, kind)
    for f in (0, 1, frames // 2, frames - 1):
        ref = oracle.mdct(sig[f * hop:f * hop + n].astype(np.float64), w, "forward")
        check(cs[f], ref, f64, n, f"frame {f}")
    inner = slice(hop, (frames - 1) * hop + hop)
    e = dev_rel(y[inner], xs[inner])
    assert e <= (TOL_F64 if f64 else 2 * tol_f32(n)), e
    # the unitary pair reconstructs as well
    yu = md.backward_signal(md.forward_signal(xs, hop, "forwardu"), hop, "backwardu")
    torch.cuda.synchronize()
    assert dev_rel(yu[inner], xs[inner]) <= (TOL_F64 if f64 else 2 * tol_f32(n))


def test_mdct_ola_matches_host_overlap_add():
    # the device overlap-add equals overlap-adding the backward frames on the
    # host, also for a hop that is not len/2 (three frames overlap)
    n, hop, frames = 64, 16, 40
    r = rng(70)
    c = r.standard_normal((frames, n // 2))
    md = Mdct1D.with_vorbis(n)
    y = md.backward_signal(c, hop)
    fr = oracle.mdct(c, oracle.window(n, "vorbis"), "backward")
    ref = np.zeros((frames - 1) * hop + n)
    for f in range(frames):
        ref[f * hop:f * hop + n] += fr[f]
    check(y, ref, True, n)


# ---------------------------------------------------------------- c5: the uneven batch at full size
def test_uneven_full_workload():
    # the full 10,000-length uneven batch (<= 2^28 points, FP32): sampled
    # transforms against numpy's FP64 FFT, and for every transform Parseval
    # (sum |X|^2 = n sum |x|^2) and the round trip backward(forward(x)) = x,
    # both reduced per transform on the device
    import torch
    from chfft_amd.workloads import uneven_lengths
    lens = uneven_lengths()
    fft = MultiFFT(lens, np.float32)
    x = cuda_uniform((fft.total,), True, False, 0x5EED0005)
    y = fft.forward(x)
    z = fft.backward(y)
    torch.cuda.synchronize()
    nn = torch.tensor(lens, dtype=torch.float64, device="cuda")
    # per-transform sums in float64 (a running-sum difference would lose the
    # small transforms' digits behind the large ones)
    seg_len = torch.tensor(lens, dtype=torch.int64, device="cuda")

    def seg_sum(v):
        return torch.segment_reduce(v, "sum", lengths=seg_len)
    ex = seg_sum(x.abs().to(torch.float64) ** 2)
    ey = seg_sum(y.abs().to(torch.float64) ** 2)
    ed = seg_sum((z - x).abs().to(torch.float64) ** 2)
    tol = torch.tensor([tol_f32(n) for n in lens], dtype=torch.float64, device="cuda")
    pars = (ey / (nn * ex) - 1).abs()
    rt = (ed / ex).sqrt()
    del z
    xs = x.cpu().numpy()
    ys = y.cpu().numpy()
    bad_p = torch.nonzero(pars > 2 * tol).flatten().cpu().tolist()
    bad_r = torch.nonzero(rt > 2 * tol).flatten().cpu().tolist()
    worst = [int(pars.argmax()), int(rt.argmax())]
    report = []
    for i in sorted(set(bad_p[:8] + bad_r[:8] + worst)):
        o0, o1 = fft.offsets[i], fft.offsets[i + 1]
        e = rel_l2(ys[o0:o1], np.fft.fft(xs[o0:o1].astype(np.complex128)))
        report.append(f"transform {i} n={lens[i]} parseval {float(pars[i]):.2e} round trip {float(rt[i]):.2e} "
                      f"rel-L2 vs numpy {e:.2e}")
    print("\n".join(report))
    assert not bad_p and not bad_r, f"{len(bad_p)} Parseval, {len(bad_r)} round-trip failures:\n" + "\n".join(report)
    big = np.argsort(lens)[-10:]
    for i in sorted(set(range(0, len(lens), 97)) | set(big.tolist())):
        o0, o1 = fft.offsets[i], fft.offsets[i + 1]
        check(ys[o0:o1], np.fft.fft(xs[o0:o1].astype(np.complex128)), False, lens[i], f"transform {i}")


# ---------------------------------------------------------------- long-length operator paths
@pytest.mark.parametrize("n,dtype", [(65536, np.float32), (32768, np.float64), (2 * 8209, np.float32),
                                     (2 * 10007, np.float64)])
def test_rfft_long_lengths(n, dtype):
    # half lengths above one workgroup's LDS (four-step or four-step
    # Bluestein inner transform, split and un-split as passes of their own)
    f64 = dtype == np.float64
    r = rng(n)
    x = r.standard_normal((3, n)).astype(dtype)
    f = RFft1D.new(n, dtype)
    assert "operator" in f._plan_for(3).describe()
    xr = x.astype(np.float64)
    for m in ("forward", "forward0", "forwardu", "forwardn"):
        check(getattr(f, m)(x), oracle.rfft(xr, m), f64, n, m)
    X = oracle.rfft(xr, "forward")
    Xc = X.astype(np.complex128 if f64 else np.complex64)
    for m in ("backward", "backward0", "backwardu", "backwardn"):
        check(getattr(f, m)(Xc), oracle.rfft(X, m, n=n), f64, n, m)


@pytest.mark.parametrize("n,dtype", [(65536, np.float32), (32768, np.float64), (2 * 10007, np.float32)])
def test_dct_long_lengths(n, dtype):
    f64 = dtype == np.float64
    r = rng(n + 1)
    x = (2 * r.random((3, n)) - 1).astype(dtype)
    xr = x.astype(np.float64)
    for t in (2, 3):
        for ortho in (False, True):
            d = Dct1D.new(DctType(t), n, ortho, dtype)
            for m in ("forward", "forwardu", "forwardn"):
                check(getattr(d, m)(x), oracle.dct(xr, t, ortho, m), f64, n, f"DCT-{t} ortho={ortho} {m}")


@pytest.mark.parametrize("n,dtype", [(131072, np.float32), (65536, np.float64), (4 * 10007, np.float32)])
def test_mdct_long_lengths(n, dtype):
    f64 = dtype == np.float64
    r = rng(n + 2)
    x = r.standard_normal((2, n)).astype(dtype)
    md = Mdct1D.with_vorbis(n, dtype)
    w = oracle.window(n, "vorbis")
    xr = x.astype(np.float64)
    for m in ("forward", "forwardu"):
        check(getattr(md, m)(x), oracle.mdct(xr, w, m), f64, n, m)
    c = oracle.mdct(xr, w, "forward")
    for m in ("backward", "backwardu"):
        check(getattr(md, m)(c.astype(dtype)), oracle.mdct(c, w, m), f64, n, m)


# ---------------------------------------------------------------- CFft2D: large, non-square, long columns
@pytest.mark.parametrize("shape", [(1024, 8192), (8192, 1024)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cfft2d_large_non_square(shape, dtype):
    # cfft2d.rs:326-363 at 2^23 points, both orientations
    import torch
    f64 = dtype == np.float64
    m, n = shape
    x = cuda_uniform((m, n), True, f64, 0x2D00 + m)
    f = CFft2D.with_len(m, n, dtype)
    y = f.forward(x)
    z = f.backward(y)
    torch.cuda.synchronize()
    xs = x.cpu().numpy().astype(np.complex128)
    check(y.cpu().numpy(), np.fft.fft2(xs), f64, m * n, "forward")
    e = dev_rel(z, x)
    assert e <= (TOL_F64 if f64 else 2 * tol_f32(m * n)), e


@pytest.mark.parametrize("shape,dtype", [((32768, 3), np.float32), ((16384, 2), np.float64),
                                         ((20011, 2), np.float32), ((8209, 1), np.float64)])
def test_cfft2d_long_columns(shape, dtype):
    # columns above one workgroup's LDS: transposed, run as rows, transposed back
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    m, n = shape
    r = rng(m + n)
    x = (r.standard_normal((2, m, n)) + 1j * r.standard_normal((2, m, n))).astype(cdt)
    f = CFft2D.with_len(m, n, dtype)
    for meth in ("forward", "forward0", "forwardu", "forwardn", "backward", "backward0", "backwardu"):
        got = getattr(f, meth)(x)
        check(got, oracle.c2c2d(x.astype(np.complex128), meth), f64, m * n, meth)


def test_cfft2d_single_row_chunked(monkeypatch):
    # one row of a four-step length whose row pass runs in chunks: every
    # chunk's last launch carries the method's scale (forwardn, backward)
    monkeypatch.setenv("CHFFT_FS_CHUNK", "2")
    n, batch = 43200, 5
    r = rng(71)
    x = (r.standard_normal((batch, 1, n)) + 1j * r.standard_normal((batch, 1, n))).astype(np.complex64)
    f = CFft2D.with_len(1, n, np.float32)
    xr = x.astype(np.complex128)
    check(f.forwardn(x), np.fft.fft2(xr) / n, False, n, "forwardn")
    check(f.backward(x), np.fft.ifft2(xr), False, n, "backward")
    check(f.forwardu(x), np.fft.fft2(xr) / np.sqrt(n), False, n, "forwardu")


# ---------------------------------------------------------------- error paths
def test_multi_barrier_failure_is_reported(monkeypatch):
    # test mode of the multi-length launch: every grid-wide barrier waits for
    # one arrival more than the grid makes, for 10 us, then each workgroup
    # records the failure and exits; the torch path and the host path both
    # report it instead of returning an incomplete result
    import torch
    from chfft_amd import ChfftError
    from chfft_amd import _lib
    lens = [20011, 4096, 97, 100000]  # four-step items: the launch has phases and barriers
    r = rng(72)
    tot = sum(lens)
    xh = (r.standard_normal(tot) + 1j * r.standard_normal(tot)).astype(np.complex64)
    monkeypatch.setenv("CHFFT_MULTI_BARRIER_TEST", "1")
    bad = MultiFFT(lens, np.float32)
    monkeypatch.delenv("CHFFT_MULTI_BARRIER_TEST")
    with pytest.raises(ChfftError) as ei:
        bad.forward(torch.from_numpy(xh).cuda())
    assert ei.value.status == _lib.ERR_INCOMPLETE
    with pytest.raises(ChfftError) as ei:
        bad.forward(xh)
    assert ei.value.status == _lib.ERR_INCOMPLETE
    bad.plan.check()  # the record was cleared by the reads above
    good = MultiFFT(lens, np.float32)
    y = good.forward(torch.from_numpy(xh).cuda()).cpu().numpy()
    for seg, o0, o1 in zip(lens, good.offsets[:-1], good.offsets[1:]):
        check(y[o0:o1], np.fft.fft(xh[o0:o1].astype(np.complex128)), False, seg)


def test_plan_exec_rejects_mismatched_buffers():
    import torch
    p = Plan("c2c", False, 4, 256)
    x = torch.zeros(4, 256, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p.exec(x, y, "forward")
    with pytest.raises(ValueError, match="dtype"):
        p.exec(x.to(torch.complex128), y, "forward")
    with pytest.raises(ValueError, match="contiguous"):
        p.exec(torch.zeros(256, 8, dtype=torch.complex64, device="cuda").t(), y, "forward")
    with pytest.raises(ValueError, match="invalid length"):
        p.exec(x[:2], y, "forward")
    with pytest.raises(ValueError, match="dtype"):
        p.exec(np.zeros((4, 256), np.complex128), np.zeros((4, 256), np.complex64), "forward")
