"""GPU parity: the HIP engine (through the C ABI, via the Python mirror of the
reference API) against the CPU oracle on the same seeded inputs.

Bars: FP64 plain method rel-L2 <= 1e-12 against the FP64 oracle (the
restatement of chfft f64); FP32 form rel-L2 <= 1e-6 * log2(N) against the FP64
oracle evaluated on the same FP32-rounded inputs. Length sweeps follow the
reference tests (cfft1d.rs:566-603, rfft1d.rs:390-402, dct1d.rs:483-525,
mdct1d.rs:384-466, cfft2d.rs:454-488); config sizes follow the benchmark
workloads at reduced batch.
"""
import numpy as np
import pytest

import oracle
from chfft_amd import CFft1D, CFft2D, Dct1D, DctType, Mdct1D, RFft1D
from tests.gpu_util import TOL_F64, crand, rel_l2, rng, tol_f32

pytestmark = pytest.mark.gpu

C2C_METHODS = ["forward", "forward0", "forwardu", "forwardn", "backward", "backward0", "backwardu", "backwardn"]
PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]


def check(gpu, ref, f64, n):
    e = rel_l2(gpu, ref)
    tol = TOL_F64 if f64 else tol_f32(n)
    assert e <= tol, f"rel-L2 {e:.3e} > {tol:.1e} (n={n})"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c2c_sweep_all_methods(dtype):
    # cfft1d.rs:566-592: N = 1..99, every exec variant, 10 vectors per length
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    r = rng(1)
    fft = CFft1D.new(dtype)
    for n in range(1, 100):
        x = crand(r, (10, n), cdt)
        for m in C2C_METHODS:
            got = getattr(fft, m)(x)
            ref = oracle.c2c(x.astype(np.complex128), m)
            check(got, ref, f64, n)
        # in-place variants (cfft1d.rs:526-542)
        for m, mi in [("forward0", "forward0i"), ("backward0", "backward0i"), ("forwardu", "forwardui"),
                      ("backwardu", "backwardui")]:
            y = x.copy()
            getattr(fft, mi)(y)
            check(y, oracle.c2c(x.astype(np.complex128), m), f64, n)


def test_c2c_primes_one_plan():
    # cfft1d.rs:594-603: one plan object reused across prime lengths (replans)
    r = rng(2)
    fft = CFft1D.new(np.float64)
    for n in PRIMES:
        x = crand(r, (10, n))
        got = fft.forward(x)
        check(got, oracle.c2c(x, "forward"), True, n)
        back = fft.backward(got)
        check(back, x, True, n)


def test_readme_known_answer():
    # README.md:33-37 / cfft1d.rs:43-48: [2, 1+i, 3i, 2+4i]
    x = np.array([2, 1 + 1j, 3j, 2 + 4j])
    got = CFft1D.with_len(4).forward(x)
    np.testing.assert_allclose(got, [5 + 8j, -1 - 2j, -1 - 2j, 5 - 4j], atol=1e-14)


@pytest.mark.parametrize("n,batch", [(4096, 256), (512, 100), (2048, 33), (4093, 37), (4096, 3),
                                     (43200, 4), (4093, 8), (10007, 4), (65537, 2),
                                     (8192, 16), (16384, 4), (20000, 2), (1 << 20, 1), (3 ** 9, 2),
                                     (7 * 11 * 13 * 17, 2), (1048573, 1)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c2c_config_sizes(n, batch, dtype):
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    r = rng(n)
    x = (r.standard_normal((batch, n)) + 1j * r.standard_normal((batch, n))).astype(cdt)
    fft = CFft1D.with_len(n, dtype)
    y = fft.forward(x)
    xr = x.astype(np.complex128)
    check(y, np.fft.fft(xr) if n > 70000 else oracle.c2c(xr, "forward"), f64, n)
    z = fft.backward(y)
    check(z, xr, f64, n)


@pytest.mark.parametrize("n,cols,rows", [(65537, 256, 512), (16385, 128, 256), (40961, 256, 320)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bluestein_padding_2n_minus_2(n, cols, rows, dtype):
    # four-step Bluestein over M = 2N - 2 (2^17, 2^15, 2^14 x 5): the one
    # slot that the lags N-1 and -(N-1) share holds the symmetric chirp's
    # value for both, so the cyclic convolution equals the linear one
    from chfft_amd import Plan
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    d = Plan("c2c", f64, 2, n).describe().split("inverse:")[0]
    assert f"columns L={cols} " in d and f"rows L={rows} " in d and cols * rows == 2 * n - 2, d
    r = rng(n + 7)
    x = (r.standard_normal((2, n)) + 1j * r.standard_normal((2, n))).astype(cdt)
    fft = CFft1D.with_len(n, dtype)
    xr = x.astype(np.complex128)
    for m in ("forward", "backward", "forwardn", "backwardu"):
        check(getattr(fft, m)(x), np.fft.fft(xr) * {"forward": 1, "forwardn": 1 / n}[m] if m.startswith("f") else
              np.fft.ifft(xr) * {"backward": 1, "backwardu": np.sqrt(n)}[m], f64, n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rfft_sweep(dtype):
    # rfft1d.rs:390-402: even N = 2..198, all scalings, round trips
    f64 = dtype == np.float64
    r = rng(3)
    for n in range(2, 200, 2):
        x = r.random((10, n)).astype(dtype)
        f = RFft1D.new(n, dtype)
        for m in ["forward", "forward0", "forwardu", "forwardn"]:
            got = f.__getattribute__(m)(x)
            check(got, oracle.rfft(x.astype(np.float64), m), f64, n)
        X = oracle.rfft(x.astype(np.float64), "forward")
        for m in ["backward", "backward0", "backwardu", "backwardn"]:
            got = getattr(f, m)(X.astype(np.complex128 if f64 else np.complex64))
            check(got, oracle.rfft(X, m, n=n), f64, n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rfft_config_size(dtype):
    f64 = dtype == np.float64
    r = rng(4)
    n = 8192
    x = r.standard_normal((64, n)).astype(dtype)
    f = RFft1D.new(n, dtype)
    X = f.forward(x)
    check(X, oracle.rfft(x.astype(np.float64), "forward"), f64, n)
    xb = f.backward(X)
    check(xb, x.astype(np.float64), f64, n)


def test_rfft_invalid_length():
    # rfft1d.rs:404-422
    with pytest.raises(ValueError, match="invalid length"):
        RFft1D.new(11)
    f = RFft1D.new(4)
    with pytest.raises(ValueError, match="invalid length"):
        f.forward(np.arange(5.0))
    with pytest.raises(ValueError, match="invalid length"):
        f.backward(np.arange(5.0).astype(np.complex128))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dct_sweep(dtype):
    # dct1d.rs:436-503: DCT-II forwardu/forwardn/forward0 against DCT-III pairs
    f64 = dtype == np.float64
    r = rng(5)
    for n in range(2, 200, 2):
        x = r.random((10, n)).astype(dtype)
        d2 = Dct1D.new(DctType.Two, n, False, dtype)
        d3 = Dct1D.new(DctType.Three, n, False, dtype)
        for m in ["forward", "forward0", "forwardu", "forwardn"]:
            check(getattr(d2, m)(x), oracle.dct(x.astype(np.float64), 2, False, m), f64, n)
            check(getattr(d3, m)(x), oracle.dct(x.astype(np.float64), 3, False, m), f64, n)
        check(d3.forwardu(d2.forwardu(x)), x, f64, n)
        check(d3.forward0(d2.forwardn(x)), x, f64, n)


def test_dct_ortho_and_setup():
    # is_ortho scales input element 0 (dct1d.rs:220-224, 343-347); setup paths
    # dct1d.rs:505-525
    r = rng(6)
    for n in [2, 4, 6, 10, 64, 126]:
        x = r.random((4, n))
        for t in (2, 3):
            d = Dct1D.new(DctType(t), n, True)
            check(d.forward(x), oracle.dct(x, t, True, "forward"), True, n)
        d = Dct1D.new(DctType.Two, 2 * n, True)
        d.setup(DctType.Three, n, False)
        check(d.forwardn(x), oracle.dct(x, 3, False, "forwardn"), True, n)


def test_dct_config_size():
    r = rng(7)
    n = 1024
    x = (2 * r.random((256, n)) - 1)
    for dtype in (np.float64, np.float32):
        xs = x.astype(dtype)
        for t in (2, 3):
            got = Dct1D.new(DctType(t), n, False, dtype).forward(xs)
            check(got, oracle.dct(xs.astype(np.float64), t, False, "forward"), dtype == np.float64, n)


def test_dct_invalid_length():
    with pytest.raises(ValueError, match="invalid length"):
        Dct1D.new(DctType.Two, 11, False)
    d = Dct1D.new(DctType.Two, 8, False)
    with pytest.raises(ValueError, match="invalid length"):
        d.forward(np.arange(10.0))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["sine", "vorbis"])
def test_mdct_sweep(dtype, kind):
    # mdct1d.rs:384-448: N = 4..396 step 4, forward/backward and unitary pairs
    f64 = dtype == np.float64
    r = rng(8)
    for n in range(4, 400, 4):
        x = r.random((10, n)).astype(dtype)
        md = Mdct1D.with_sine(n, dtype) if kind == "sine" else Mdct1D.with_vorbis(n, dtype)
        w = oracle.window(n, kind)
        for m in ["forward", "forwardu"]:
            check(getattr(md, m)(x), oracle.mdct(x.astype(np.float64), w, m), f64, n)
        c = oracle.mdct(x.astype(np.float64), w, "forward")
        for m in ["backward", "backwardu"]:
            check(getattr(md, m)(c.astype(dtype)), oracle.mdct(c, w, m), f64, n)


def test_mdct_setup_and_invalid():
    # mdct1d.rs:450-486
    md = Mdct1D.with_sine(64)
    md.setup(16)
    r = rng(9)
    x = r.random((3, 16))
    check(md.forward(x), oracle.mdct(x, oracle.window(16), "forward"), True, 16)
    with pytest.raises(ValueError, match="invalid length"):
        Mdct1D.with_sine(10)
    md = Mdct1D.with_sine(8)
    with pytest.raises(ValueError, match="invalid length"):
        md.forward(np.arange(10.0))
    with pytest.raises(ValueError, match="invalid length"):
        md.backward(np.arange(8.0))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mdct_overlapped_signal(dtype):
    # the MDCT workload: N = 2048, hop 1024 over a signal (reduced length)
    n, hop = 2048, 1024
    r = rng(10)
    sig = r.standard_normal(hop * 64 + n).astype(dtype)
    md = Mdct1D.with_sine(n, dtype)
    got = md.forward_signal(sig, hop)
    frames = (len(sig) - n) // hop + 1
    ref = oracle.mdct(sig.astype(np.float64), oracle.window(n), "forward", n=n, hop=hop, frames=frames)
    check(got, ref, dtype == np.float64, n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cfft2d_sweep(dtype):
    # cfft2d.rs:454-488: 1..9 x 1..9, forward then backward
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    r = rng(11)
    fft = CFft2D.new(dtype)
    for m in range(1, 10):
        for n in range(1, 10):
            x = crand(r, (m, n), cdt)
            for meth in ["forward", "forward0", "forwardu", "forwardn", "backward", "backward0", "backwardu"]:
                got = getattr(fft, meth)(x)
                check(got, oracle.c2c2d(x.astype(np.complex128), meth), f64, m * n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cfft2d_config_size(dtype):
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    r = rng(12)
    m = n = 2048
    x = (r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))).astype(cdt)
    f = CFft2D.with_len(m, n, dtype)
    y = f.forward(x)
    check(y, np.fft.fft2(x.astype(np.complex128)), f64, m * n)
    check(f.backward(y), x.astype(np.complex128), f64, m * n)


@pytest.mark.parametrize("f64", [False, True])
def test_benchmark_schedules(f64):
    # the benchmark shapes run on the length-specialised kernels with the
    # schedules they were tuned with: 43200 = 180 x 240 as prime-factor
    # passes (12 x 15, 16 x 15) in equal chunks whose scratch stays below
    # 96 MB, the 2-D 2048-point columns on their own tile kernel
    from chfft_amd import Plan
    d = Plan("c2c", f64, 512, 43200).describe()
    fwd = d.split("inverse:")[0].strip().splitlines()[1:]
    chunks, per = (4, 128) if f64 else (2, 256)
    assert len(fwd) == 2 * chunks, d
    cols = "specialised four-step columns L=180 radix[12,15] + twiddle (transforms %d+)"
    rows = "specialised four-step rows L=240 radix[16,15] + transpose (transforms %d+)"
    for c in range(chunks):
        assert fwd[2 * c].strip().startswith(cols % (c * per)), fwd[2 * c]
        assert fwd[2 * c + 1].strip().startswith(rows % (c * per)), fwd[2 * c + 1]
    d2 = Plan("c2c2d", f64, 1, 2048, 2048).describe()
    assert "specialised 2-D columns L=2048" in d2, d2
    for n, b in ((10007, 512), (65537, 64)):
        lines = Plan("c2c", f64, b, n).describe().split("inverse:")[0].strip().splitlines()[1:]
        if n == 10007 and not f64:
            # one fused launch: the padded 20160-point sequence fits one CU's LDS
            assert len(lines) == 1 and lines[0].strip().startswith("specialised fused bluestein N=10007 M L=20160"), \
                lines
        else:
            assert len(lines) == 3 and all(l.strip().startswith("specialised bluestein") for l in lines), lines


@pytest.mark.parametrize("schedule", ["persistent", "split"])
@pytest.mark.parametrize("n,batch", [(43200, 96), (10007, 64), (20000, 40), (3 ** 9, 33)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fourstep_schedules(monkeypatch, schedule, n, batch, dtype):
    # both four-step schedules (one persistent launch vs. split launches)
    monkeypatch.setenv("CHFFT_FOURSTEP", schedule)
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    r = rng(n + batch)
    x = (r.standard_normal((batch, n)) + 1j * r.standard_normal((batch, n))).astype(cdt)
    fft = CFft1D.with_len(n, dtype)
    xr = x.astype(np.complex128)
    for m in ("forward", "backward", "forwardu"):
        got = getattr(fft, m)(x)
        check(got, oracle.c2c(xr, m), f64, n)


@pytest.mark.parametrize("chunk", [None, "200", "-1"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_headline_full_batch(dtype, chunk, monkeypatch):
    # N = 43200 at the full batch of 512 on device buffers, split into the
    # planner's chunks (FP32: 2 x 256, FP64: 4 x 128), uneven chunks of
    # 200 + 200 + 112 and one launch pair over the whole batch; spot-check
    # transforms across the batch (first, both ends of chunks, last) against
    # the oracle
    import torch
    from chfft_amd import Plan
    if chunk is not None:
        monkeypatch.setenv("CHFFT_FS_CHUNK", chunk)
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    n, batch = 43200, 512
    r = rng(43)
    x = (r.standard_normal((batch, n)) + 1j * r.standard_normal((batch, n))).astype(cdt)
    p = Plan("c2c", f64, batch, n)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    for m in ("forward", "backward"):
        p.exec(xd, yd, m)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        for b in (0, 1, 127, 128, 199, 200, 255, 256, 399, 400, 511):
            check(y[b], oracle.c2c(x[b].astype(np.complex128), m), f64, n)
        # in place (the *i methods): every chunk's row launch overwrites only
        # its own transforms, whose input its column launch has consumed
        zd = xd.clone()
        p.exec(zd, zd, m)
        torch.cuda.synchronize()
        assert np.array_equal(zd.cpu().numpy(), y)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_operator_paths_config_size(dtype):
    # DCT-II/III (N = 1024, with and without is_ortho) and MDCT (N = 2048,
    # hop 1024) at their benchmark sizes on device buffers, once aligned and
    # once offset by one real (the kernels then take their per-element path):
    # both against the oracle
    import torch
    from chfft_amd import Plan
    f64 = dtype == np.float64
    tdt = torch.float64 if f64 else torch.float32
    r = rng(60)
    n, batch = 1024, 64
    x = (2 * r.random((batch, n)) - 1).astype(dtype)
    for t in (2, 3):
        for ortho in (False, True):
            p = Plan("dct", f64, batch, t, n, ortho)
            ref = oracle.dct(x.astype(np.float64), t, ortho, "forward")
            for off in (0, 1):
                xd = torch.zeros(batch * n + 1, dtype=tdt, device="cuda")
                xd[off:off + batch * n] = torch.from_numpy(x.reshape(-1)).cuda()
                yd = torch.zeros(batch * n + 1, dtype=tdt, device="cuda")
                p.exec_ptr(xd[off:].data_ptr(), yd[off:].data_ptr(), "forward",
                           torch.cuda.current_stream().cuda_stream)
                torch.cuda.synchronize()
                got = yd[off:off + batch * n].cpu().numpy().reshape(batch, n)
                check(got, ref, f64, n)
    nm, hop, frames = 2048, 1024, 40
    sig = r.standard_normal(hop * (frames - 1) + nm).astype(dtype)
    win = oracle.window(nm, "sine")
    p = Plan("mdct", f64, frames, nm, win, hop)
    ref = oracle.mdct(sig.astype(np.float64), win, "forward", n=nm, hop=hop, frames=frames)
    for off in (0, 1):
        xd = torch.zeros(len(sig) + 1, dtype=tdt, device="cuda")
        xd[off:off + len(sig)] = torch.from_numpy(sig).cuda()
        yd = torch.zeros(frames * nm // 2 + 1, dtype=tdt, device="cuda")
        p.exec_ptr(xd[off:].data_ptr(), yd[off:].data_ptr(), "forward", torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got = yd[off:off + frames * nm // 2].cpu().numpy().reshape(frames, nm // 2)
        check(got, ref, f64, nm)
