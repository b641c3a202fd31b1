"""GPU tests of the uneven multi-length batch (one persistent launch), its
hipGraph capture and replay, and the autotuner. Every transform is checked
against the oracle (CFft1D of its own length, cfft1d.rs:208-230) or, above
70000 points, against numpy's FFT of the same input; same bars as the other
parity tests (FP64 rel-L2 <= 1e-12, FP32 <= 1e-6 log2 N)."""
import numpy as np
import pytest

import oracle
from chfft_amd import MultiFFT
from chfft_amd.workloads import uneven_lengths
from tests.gpu_util import TOL_F64, rel_l2, rng, tol_f32

pytestmark = pytest.mark.gpu

MIXED = [1, 2, 3, 5, 16, 97, 100, 4096, 4093, 8192, 8191, 9000, 10007, 20011, 65536, 3 ** 9, 7 * 11 * 13, 44100]


def ref_segments(fft, x, method):
    out = []
    for seg in fft.split(x.astype(np.complex128)):
        n = len(seg)
        if n > 70000:
            s = {"forward": 1, "forwardu": n ** -0.5, "forwardn": 1 / n, "backward": 1 / n, "backward0": 1,
                 "backwardu": n ** -0.5}[method]
            r = (np.fft.fft(seg) if method.startswith("f") else np.fft.ifft(seg) * n) * s
        else:
            r = oracle.c2c(seg, method)
        out.append(r)
    return out


def check_segments(fft, got, ref, f64):
    worst = 0.0
    for g, r in zip(fft.split(np.asarray(got)), ref):
        n = len(r)
        e = rel_l2(g, r)
        tol = TOL_F64 if f64 else tol_f32(n)
        assert e <= tol, f"n={n}: rel-L2 {e:.3e} > {tol:.1e}"
        worst = max(worst, e)
    return worst


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("method", ["forward", "backward", "forwardu", "backward0"])
def test_multi_mixed_lengths(dtype, method):
    f64 = dtype == np.float64
    cdt = np.complex128 if f64 else np.complex64
    lens = MIXED + list(rng(50).integers(2, 3000, 40))
    fft = MultiFFT(lens, dtype)
    r = rng(51)
    x = (r.standard_normal(fft.total) + 1j * r.standard_normal(fft.total)).astype(cdt)
    got = getattr(fft, method)(x)
    check_segments(fft, got, ref_segments(fft, x, method), f64)


# column and row lengths of the register-resident tiles for smooth four-step
# items (FP32): products of pairs of them, so that the planner's splits put
# both phases of most items on those tiles. A wrong four-step twiddle keeps
# the transform unitary (round trips and Parseval still hold), so every item
# is compared with the reference directly.
SMOOTH_COLS = [125, 150, 225, 243, 250, 270, 300, 324, 360, 375, 400, 405, 450, 480, 486, 500]
SMOOTH_ROWS = [240, 256, 288, 320, 375, 384, 400, 432, 480, 512, 576, 625, 640, 720, 768, 800]
# lengths whose preferred split is 225, 243, 300 or 450 columns, or 512 or 640
# rows, which the pair products below do not reach
SMOOTH_EXTRA = [54000, 58320, 76800, 165888, 168750, 307200]


@pytest.mark.parametrize("method", ["forward", "backward"])
def test_multi_smooth_tile_shapes(method):
    lens = sorted({c * r for i, c in enumerate(SMOOTH_COLS)
                   for r in (SMOOTH_ROWS[i], SMOOTH_ROWS[len(SMOOTH_ROWS) - 1 - i])} | set(SMOOTH_EXTRA))
    fft = MultiFFT(lens, np.float32)
    assert "four-step phases on register-resident tiles" in fft.plan.describe()
    r = rng(53)
    x = (r.standard_normal(fft.total) + 1j * r.standard_normal(fft.total)).astype(np.complex64)
    got = getattr(fft, method)(x)
    check_segments(fft, got, ref_segments(fft, x, method), False)


def test_multi_graph_replay():
    # capture one exec in a HIP graph, replay it several times with eager work
    # in between; every replay must equal the eager result
    import torch
    lens = [4096, 4093, 9000, 10007, 97, 1, 20011] * 3
    fft = MultiFFT(lens, np.float32)
    r = rng(52)
    xh = (r.standard_normal(fft.total) + 1j * r.standard_normal(fft.total)).astype(np.complex64)
    x = torch.from_numpy(xh).cuda()
    eager = fft.forward(x)
    torch.cuda.synchronize()
    y = torch.empty_like(x)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        fft.plan.exec_ptr(x.data_ptr(), y.data_ptr(), "forward", stream.cuda_stream)  # warm-up off the graph
    torch.cuda.current_stream().wait_stream(stream)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fft.plan.exec_ptr(x.data_ptr(), y.data_ptr(), "forward", torch.cuda.current_stream().cuda_stream)
    for k in range(4):
        y.zero_()
        z = (x * (k + 1)).sum()  # unrelated eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), f"replay {k} differs from the eager result"
        del z
    check_segments(fft, y.cpu().numpy(), ref_segments(fft, xh, "forward"), False)


def test_multi_autotune_keeps_results():
    import torch
    lens = uneven_lengths(count=400, cap=1 << 22, seed=7)
    fft = MultiFFT(lens, np.float32)
    r = rng(53)
    xh = (r.standard_normal(fft.total) + 1j * r.standard_normal(fft.total)).astype(np.complex64)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty_like(x)
    choice = fft.plan.autotune(x, y, reps=2)
    assert choice >= 0
    got = fft.forward(x).cpu().numpy()
    check_segments(fft, got, ref_segments(fft, xh, "forward"), False)
    assert "autotune" in fft.plan.describe()


def test_uneven_workload_buckets():
    # a reduced uneven workload (same generator as the benchmark), FP32,
    # rel-L2 per octave bucket within the FP32 budget
    lens = uneven_lengths(count=600, cap=1 << 23, seed=11)
    fft = MultiFFT(lens, np.float32)
    r = rng(54)
    x = (r.standard_normal(fft.total) + 1j * r.standard_normal(fft.total)).astype(np.complex64)
    got = fft.forward(x)
    check_segments(fft, got, ref_segments(fft, x, "forward"), False)


def test_fourstep_autotune():
    import torch
    from chfft_amd import Plan
    n, batch = 43200, 8
    p = Plan("c2c", False, batch, n)
    r = rng(55)
    xh = (r.standard_normal((batch, n)) + 1j * r.standard_normal((batch, n))).astype(np.complex64)
    x = torch.from_numpy(xh).cuda()
    y = torch.empty_like(x)
    best = p.autotune(x, y, reps=2)
    spec, report = p.tuning()
    assert 0 <= best and f"[{best}] {spec}:" in report, (best, spec, report)
    p.exec(x, y, "forward")
    torch.cuda.synchronize()
    e = rel_l2(y.cpu().numpy(), oracle.c2c(xh.astype(np.complex128), "forward"))
    assert e <= tol_f32(n)
