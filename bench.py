"""Benchmark of the transform engine on one MI355X.

Default (the driver's line): CFft1D forward on the mixed-radix workload
N = 43200 (2^6 3^3 5^2), batch 512, complex N(0,1) entries from a seed. One
step = one forward transform of the whole batch, inputs resident in HBM. The
value is the FP32 form's throughput in GFLOP/s (nominal 5 N log2 N per
transform); the FP64 plain method runs on the same inputs beside it, each with
its rel-L2 against the FP64 oracle. `roofline` covers the dominant kernel of
the FP32 schedule, timed with HIP events on the stream the kernels run on;
`cpu_baseline` is the oracle (the C++ restatement of chfft's algorithm) timed
on the host cores over a bounded sample of the same workload.

--suite prints one extra JSON line per configuration (all five workloads:
c2c 4096 FP64/FP32, the headline, Bluestein primes, RFft/DCT/MDCT, the uneven
multi-length batch replayed through a HIP graph plus CFft2D 2048x2048).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--suite] [--no-cpu]
"""
import argparse
import ctypes
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FFT GFLOP/s (5Nlog2N) + % HBM roofline: FP32-batched vs FP64-plain @ rel-L2"
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
ROTATE_BYTES = 1 << 30  # buffer pools rotated so that reads come from HBM, not the 256 MB MALL


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--suite", action="store_true", help="also print one line per configuration")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only time a plain device-to-device copy of this many bytes (the reference kernel for "
                         "the HBM traffic counters and the achievable copy bandwidth)")
    return ap.parse_args()


def copy_calibration(torch, nbytes, reps=10):
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"copy_bytes": a.numel() * 4, "us": round(ms * 1e3, 2),
            "gb_s": round(2 * a.numel() * 4 / (ms * 1e6), 1)}


class Bufs:
    """A pool of input/output buffer pairs (rotated when one pair is small
    enough to stay in the MALL)."""

    def __init__(self, torch, in_bytes, out_bytes, fill, rotate=True):
        pair = in_bytes + out_bytes
        k = max(1, min(16, ROTATE_BYTES // max(pair, 1))) if rotate and pair < ROTATE_BYTES // 2 else 1
        self.ins, self.outs = [], []
        for i in range(k):
            a = torch.empty(in_bytes // 4 + 16, dtype=torch.float32, device="cuda")
            fill(a, i)
            self.ins.append(a)
            self.outs.append(torch.empty(out_bytes // 4 + 16, dtype=torch.float32, device="cuda"))
        self.k = k

    def pair(self, i):
        j = i % self.k
        return self.ins[j].data_ptr(), self.outs[j].data_ptr()


def time_plan(torch, plan, bufs, steps, warmup, method="forward", blocks=1):
    """ms per step over `steps` steps bracketed by events; with blocks > 1
    the median of that many such blocks (the per-configuration lines, whose
    microsecond steps are sensitive to one-off host stalls)."""
    st = torch.cuda.current_stream().cuda_stream
    for i in range(warmup):
        a, b = bufs.pair(i)
        plan.exec_ptr(a, b, method, st)
    torch.cuda.synchronize()
    res = []
    for _ in range(blocks):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(torch.cuda.current_stream())
        for i in range(steps):
            a, b = bufs.pair(i)
            plan.exec_ptr(a, b, method, st)
        e1.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        res.append(e0.elapsed_time(e1) / steps)
    return float(np.median(res))


def per_kernel_ms(torch, plan, bufs, steps, method="forward"):
    """Average duration of each launch of the schedule, bracketed by events
    recorded on the stream the kernels are launched on."""
    from chfft_amd import _lib
    L = _lib.lib()
    m = _lib.METHODS[method]
    n = L.chfft_plan_launch_count(plan.handle, m)
    s = torch.cuda.current_stream()
    st = s.cuda_stream
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n + 1)] for _ in range(steps)]
    torch.cuda.synchronize()
    for k in range(steps):
        a, b = bufs.pair(k)
        for i in range(n):
            ev[k][i].record(s)
            _lib.check(L.chfft_exec_range(plan.handle, ctypes.c_void_p(a), ctypes.c_void_p(b), m, i, 1,
                                          ctypes.c_void_p(st)))
        ev[k][n].record(s)
    torch.cuda.synchronize()
    return [sum(ev[k][i].elapsed_time(ev[k][i + 1]) for k in range(steps)) / steps for i in range(n)]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def cplx_fill(torch, seed):
    def fill(a, i):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed + i)
        a.normal_(generator=g)
    return fill


def traffic_for(key):
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            return json.load(f).get(key)
    return None


def c2c_case(torch, n, batch, f64, steps, warmup, seed, check=8, uniform=False, blocks=1):
    """Throughput, per-kernel times and rel-L2 (first `check` transforms vs
    the FP64 oracle) of CFft1D forward on [batch, n]."""
    import oracle
    from chfft_amd import Plan
    p = Plan("c2c", f64, batch, n)
    ib, ob = p.io_bytes("forward")

    def fill(a, i):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed + i)
        if uniform:
            a.uniform_(-1, 1, generator=g)
        else:
            a.normal_(generator=g)
    bufs = Bufs(torch, ib, ob, fill)
    ms = time_plan(torch, p, bufs, steps, warmup, blocks=blocks)
    kms = per_kernel_ms(torch, p, bufs, max(3, steps // 2))
    cdt = np.complex128 if f64 else np.complex64
    k = min(check, batch)
    fl = k * n * (4 if f64 else 2)  # float32 words of k complex rows
    x = bufs.ins[0][:fl].view(torch.complex128 if f64 else torch.complex64).cpu().numpy().reshape(k, n)
    p.exec_ptr(bufs.ins[0].data_ptr(), bufs.outs[0].data_ptr(), "forward", torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    y = bufs.outs[0][:fl].view(torch.complex128 if f64 else torch.complex64).cpu().numpy().reshape(k, n)
    ref = oracle.c2c(x.astype(np.complex128), "forward") if n <= 70000 else np.fft.fft(x.astype(np.complex128))
    err = rel_l2(y.astype(np.complex128), ref)
    flop = 5.0 * n * math.log2(n) * batch
    alg = (32.0 if f64 else 16.0) * n * batch
    return {"n": n, "batch": batch, "dtype": "f64" if f64 else "f32", "ms": round(ms, 5),
            "gflops": round(flop / (ms * 1e6), 2), "hbm_frac": round(alg / (ms * 1e6) / HBM_PEAK_GBS, 4),
            "kernel_ms": [round(v, 5) for v in kms], "rel_l2": err, "buffers_rotated": bufs.k,
            "schedule": p.describe().split("inverse:")[0].replace("forward:", "").strip()}, p, bufs, cdt


def headline(torch, args):
    import oracle
    N, BATCH, SEED = 43200, 512, 0x5EED0002
    r32, p32, b32, _ = c2c_case(torch, N, BATCH, False, args.steps, args.warmup, SEED)
    r64, p64, b64, _ = c2c_case(torch, N, BATCH, True, args.steps, args.warmup, SEED)
    k32 = r32["kernel_ms"]
    dom = int(np.argmax(k32))
    # the split four-step runs in equal chunks of transforms, two launches
    # (columns, rows) per chunk; each launch reads and writes its chunk once:
    # algorithmic bytes per launch = 16 N per transform of the chunk (FP32)
    chunks = max(1, len(k32) // 2)
    bytes_launch = 16.0 * N * BATCH / chunks
    achieved = bytes_launch / (k32[dom] * 1e6)
    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        xs = b32.ins[0][:64 * N * 2].view(torch.complex64).cpu().numpy().reshape(64, N)
        sec = oracle.time_c2c(xs, "forward", threads=threads, reps=3)
        cpu = {"value": round(5.0 * N * math.log2(N) * 64 / sec / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle (C++ restatement of chfft's mixed-radix CFft1D<f32>::forward) on 64 of the "
                         f"512 transforms of N={N} on {threads} threads, best of 3 passes, plan built before timing"}
    return {
        "metric": METRIC,
        "value": r32["gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: complex entries i.i.d. N(0,1), torch generator seed 0x5EED0002",
        "config": {"workload": "CFft1D forward, mixed radix N=43200 (2^6 3^3 5^2), batch 512",
                   "n": N, "batch": BATCH, "schedule": r32["schedule"]},
        "hbm_frac_whole_transform": r32["hbm_frac"],
        "rel_l2_f32_vs_f64_oracle": r32["rel_l2"],
        "fp64_plain": {"value": r64["gflops"], "unit": "GFLOP/s", "ms_per_step": r64["ms"],
                       "hbm_frac_whole_transform": r64["hbm_frac"], "rel_l2_vs_f64_oracle": r64["rel_l2"],
                       "kernel_ms": r64["kernel_ms"]},
        "speedup_f32_over_f64": round(r64["ms"] / r32["ms"], 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for("c2c_43200x512_f32_launch%d" % dom),
                     "kernel": "launch %d of %d (%s)" % (dom, len(k32), r32["schedule"].splitlines()[dom].strip()),
                     "kernel_ms": k32, "bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
    }


def suite(torch, args):
    """One line per configuration (printed after the headline)."""
    import oracle
    from chfft_amd import MultiFFT, Plan
    from chfft_amd.workloads import uneven_lengths
    lines = []
    st = torch.cuda.current_stream().cuda_stream

    def line(cfg, recs, extra=None):
        d = {"config": cfg, "results": recs}
        if extra:
            d.update(extra)
        lines.append(d)
        print(json.dumps(d), flush=True)

    # 1: plain method, CFft1D<f64> N=4096 forward then backward, batch 256 (+ FP32 form)
    recs = []
    for f64 in (True, False):
        r, p, b, _ = c2c_case(torch, 4096, 256, f64, args.steps, args.warmup, 0x5EED0001, uniform=True, blocks=5)
        msb = time_plan(torch, p, b, args.steps, args.warmup, "backward", blocks=5)
        r["ms_backward"] = round(msb, 5)
        recs.append(r)
    line("c1: CFft1D N=4096 batch 256, forward then backward", recs)
    # 3: Bluestein primes
    recs = []
    for n, batch in ((4093, 2048), (10007, 512), (65537, 64)):
        for f64 in (False, True):
            recs.append(c2c_case(torch, n, batch, f64, args.steps, args.warmup, 0x5EED0003, uniform=True,
                                 blocks=5)[0])
    line("c3: Bluestein chirp-z, N in {4093, 10007, 65537}", recs)
    # 4: RFft1D 8192 x 4096, Dct1D II/III 1024 x 16384, Mdct1D 2048 hop 1024 over 2^24 samples
    recs = []
    for kind, n, batch in (("r2c", 8192, 4096), ("dct2", 1024, 16384), ("dct3", 1024, 16384)):
        for f64 in (False, True):
            if kind == "r2c":
                p = Plan("r2c", f64, batch, n)
                alg = (4 * n + 8 * (n // 2 + 1)) * batch * (2 if f64 else 1)
            else:
                p = Plan("dct", f64, batch, 2 if kind == "dct2" else 3, n, False)
                alg = 8 * n * batch * (2 if f64 else 1)
            ib, ob = p.io_bytes("forward")
            b = Bufs(torch, ib, ob, cplx_fill(torch, 0x5EED0004))
            ms = time_plan(torch, p, b, args.steps, args.warmup, blocks=5)
            # accuracy on 8 rows
            rows = 8
            xr = b.ins[0][:rows * n * (2 if f64 else 1)].view(torch.float64 if f64 else torch.float32)
            xr = xr.cpu().numpy()[:rows * n].reshape(rows, n).astype(np.float64)
            p.exec_ptr(b.ins[0].data_ptr(), b.outs[0].data_ptr(), "forward", st)
            torch.cuda.synchronize()
            if kind == "r2c":
                ref = oracle.rfft(xr, "forward")
                got = b.outs[0].view(torch.complex128 if f64 else torch.complex64)[:rows * (n // 2 + 1)]
                got = got.cpu().numpy().reshape(rows, n // 2 + 1)
            else:
                ref = oracle.dct(xr, 2 if kind == "dct2" else 3)
                got = b.outs[0].view(torch.float64 if f64 else torch.float32)[:rows * n].cpu().numpy().reshape(rows, n)
            recs.append({"kind": kind, "n": n, "batch": batch, "dtype": "f64" if f64 else "f32", "ms": round(ms, 5),
                         "gflops": round(2.5 * n * math.log2(n) * batch / (ms * 1e6), 2),
                         "hbm_frac": round(alg / (ms * 1e6) / HBM_PEAK_GBS, 4),
                         "rel_l2": rel_l2(got.astype(np.complex128), ref.astype(np.complex128)),
                         "schedule": p.describe().split("inverse:")[0].replace("forward:", "").strip()})
    # MDCT over a 2^24-sample signal: 8 sinusoids + N(0, 0.01) noise, hop 1024
    n, hop, S = 2048, 1024, 1 << 24
    frames = (S - n) // hop + 1
    rng = np.random.default_rng(0x5EED0004)
    t = np.arange(S)
    sig = sum(rng.uniform(0.1, 1) * np.sin(2 * np.pi * rng.uniform(0, 0.5) * t + rng.uniform(0, 6.3))
              for _ in range(8)) + rng.normal(0, 0.01, S)
    for f64 in (False, True):
        win = oracle.window(n, "sine")
        p = Plan("mdct", f64, frames, n, win, hop)
        xs = torch.from_numpy(sig.astype(np.float64 if f64 else np.float32)).cuda()
        ys = torch.empty(frames * n // 2, dtype=xs.dtype, device="cuda")
        # copies of the signal rotated like the other cases, so that the
        # signal is read from HBM rather than the 256 MB MALL
        kr = max(1, min(16, ROTATE_BYTES // (xs.numel() * xs.element_size() + ys.numel() * ys.element_size())))
        pool = [(xs, ys)] + [(xs.clone(), torch.empty_like(ys)) for _ in range(kr - 1)]

        class Rot:
            k = kr

            def pair(self, i):
                a, b = pool[i % kr]
                return a.data_ptr(), b.data_ptr()
        ms = time_plan(torch, p, Rot(), args.steps, args.warmup, blocks=5)
        p.exec_ptr(xs.data_ptr(), ys.data_ptr(), "forward", st)
        torch.cuda.synchronize()
        chk = 64
        ref = oracle.mdct(sig[:(chk - 1) * hop + n].astype(np.float64), win, "forward", n=n, hop=hop, frames=chk)
        got = ys[:chk * n // 2].cpu().numpy().reshape(chk, n // 2)
        alg = (S + frames * n // 2) * (8 if f64 else 4)
        del pool
        recs.append({"kind": "mdct", "n": n, "hop": hop, "frames": frames, "dtype": "f64" if f64 else "f32",
                     "ms": round(ms, 5), "gflops": round(2.5 * n * math.log2(n) * frames / (ms * 1e6), 2),
                     "hbm_frac": round(alg / (ms * 1e6) / HBM_PEAK_GBS, 4), "rel_l2": rel_l2(got, ref),
                     "buffers_rotated": kr,
                     "schedule": p.describe().split("inverse:")[0].replace("forward:", "").strip()})
    line("c4: RFft1D 8192 x 4096, Dct1D II/III 1024 x 16384, Mdct1D 2048 hop 1024 over 2^24 samples", recs)

    # 5: uneven batch in one persistent launch, replayed through a HIP graph,
    # grid autotuned; plus CFft2D 2048 x 2048, all FP32
    lens = uneven_lengths()
    fft = MultiFFT(lens, np.float32)
    x = torch.empty(fft.total, dtype=torch.complex64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(0x5EED0005)
    x.view(torch.float32).normal_(generator=g)
    y = torch.empty_like(x)
    choice = fft.plan.autotune(x, y, reps=2)
    p2 = Plan("c2c2d", False, 1, 2048, 2048)
    x2 = torch.empty(2048 * 2048, dtype=torch.complex64, device="cuda")
    x2.view(torch.float32).normal_(generator=g)
    y2 = torch.empty_like(x2)
    s2 = torch.cuda.Stream()
    s2.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s2):
        fft.plan.exec_ptr(x.data_ptr(), y.data_ptr(), "forward", s2.cuda_stream)
        p2.exec_ptr(x2.data_ptr(), y2.data_ptr(), "forward", s2.cuda_stream)
    torch.cuda.current_stream().wait_stream(s2)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        cs = torch.cuda.current_stream().cuda_stream
        fft.plan.exec_ptr(x.data_ptr(), y.data_ptr(), "forward", cs)
        p2.exec_ptr(x2.data_ptr(), y2.data_ptr(), "forward", cs)
    for _ in range(2):
        graph.replay()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps // 4)
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    ms_graph = e0.elapsed_time(e1) / reps
    # the two parts separately (eager)

    class P1:
        k = 1

        def pair(self, i):
            return x.data_ptr(), y.data_ptr()

    # the 2-D transform alone: 67 MB per pair, rotated over 1 GB of pairs so
    # that it reads HBM rather than the MALL
    k2 = max(1, ROTATE_BYTES // (2 * x2.numel() * x2.element_size()))
    pool2 = [(x2, y2)] + [(x2.clone(), torch.empty_like(y2)) for _ in range(k2 - 1)]

    class P2:
        k = k2

        def pair(self, i):
            a, b = pool2[i % k2]
            return a.data_ptr(), b.data_ptr()
    ms_multi = time_plan(torch, fft.plan, P1(), max(3, args.steps // 4), 1)
    ms_2d = time_plan(torch, p2, P2(), args.steps, args.warmup, blocks=5)
    del pool2
    # rel-L2 per octave bucket on a sample of transforms (numpy FP64 reference)
    xs = x.cpu().numpy()
    ys = y.cpu().numpy()
    buckets = {}
    for i in range(0, len(lens), 7):
        o0, o1 = fft.offsets[i], fft.offsets[i + 1]
        n = lens[i]
        e = rel_l2(ys[o0:o1].astype(np.complex128), np.fft.fft(xs[o0:o1].astype(np.complex128)))
        b = int(math.log2(n))
        buckets.setdefault(b, []).append(e)
    bucket_err = {f"2^{b}..2^{b + 1}": {"max": max(v), "count": len(v), "budget": 1e-6 * (b + 1)}
                  for b, v in sorted(buckets.items())}
    x2h = x2.cpu().numpy().reshape(2048, 2048).astype(np.complex128)
    e2d = rel_l2(y2.cpu().numpy().reshape(2048, 2048), np.fft.fft2(x2h))
    flop_multi = sum(5.0 * n * math.log2(n) for n in lens if n > 1)
    alg_multi = 16.0 * fft.total
    flop_2d = 5.0 * 2048 * 2048 * math.log2(2048 * 2048)
    line("c5: 10,000 uneven FP32 transforms (log-uniform 16..2^20, 50/50 2*3*5-smooth/prime, <= 2^28 points) "
         "in one persistent launch + CFft2D 2048x2048, replayed through one HIP graph",
         [{"part": "multi", "transforms": len(lens), "points": fft.total, "ms": round(ms_multi, 4),
           "gflops": round(flop_multi / (ms_multi * 1e6), 2),
           "hbm_frac": round(alg_multi / (ms_multi * 1e6) / HBM_PEAK_GBS, 4), "autotune_choice": choice,
           "schedule": fft.plan.describe().strip(), "rel_l2_per_bucket": bucket_err},
          {"part": "cfft2d", "m": 2048, "n": 2048, "ms": round(ms_2d, 4),
           "gflops": round(flop_2d / (ms_2d * 1e6), 2),
           "hbm_frac": round(16.0 * 2048 * 2048 / (ms_2d * 1e6) / HBM_PEAK_GBS, 4), "rel_l2": e2d,
           "schedule": p2.describe().split("inverse:")[0].replace("forward:", "").strip()},
          {"part": "graph (multi + 2-D)", "ms": round(ms_graph, 4),
           "gflops": round((flop_multi + flop_2d) / (ms_graph * 1e6), 2)}])
    return lines


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the batch dimension "
              "would be split across GPUs with no exchange", file=sys.stderr)
        sys.exit(2)
    import torch
    import chfft_amd  # noqa: F401  (loads the HIP engine after torch's runtime)
    torch.cuda.init()
    if args.copy_bytes:
        print(json.dumps(copy_calibration(torch, args.copy_bytes)), flush=True)
        return
    print(json.dumps(headline(torch, args)), flush=True)
    if args.suite:
        suite(torch, args)


if __name__ == "__main__":
    main()
