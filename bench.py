"""Benchmark of the transform engine on one MI355X.

One JSON line. Its top-level fields are the headline: CFft1D forward on the
mixed-radix workload N = 43200 (2^6 3^3 5^2), batch 512, complex N(0,1)
entries from a seed; one step = one forward transform of the whole batch,
inputs resident in HBM. `value` is the FP32 form's throughput in GFLOP/s
(nominal 5 N log2 N per transform); the FP64 plain method runs on the same
inputs beside it, each with its rel-L2 against the FP64 oracle.

`configs` holds every benchmark configuration (c1 plain FP64 C2C 4096 x 256,
c2 the headline in both precisions, c3 Bluestein primes, c4 RFft/DCT/MDCT,
c5 the uneven batch in one persistent launch and CFft2D 2048 x 2048, replayed
through a HIP graph), one record per case with:
  * `roofline`: algorithmic bytes of the whole transform (one read of the
    input, one write of the output) over the step time, HIP events on the
    stream the kernels are launched on around K back-to-back steps;
    `traffic` = bytes per step counted by rocprofv3 below the L2 (FETCH_SIZE
    + WRITE_SIZE, scaled by a plain copy) where a counter run has been
    recorded, else null; `launches_ms` brackets each launch on its own;
  * `cpu_baseline`: the oracle (a C++ restatement of chfft's algorithm,
    built with -march=native on this host where g++ is present) timed on a
    bounded sample of the same workload, plans built before the clock, on
    min(cores, 16) threads and on 1 core.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only ID ...] [--copy-bytes B]
                [--uneven-class all|fused|smooth4|blue4_small|blue4_big]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FFT GFLOP/s (5Nlog2N) + % HBM roofline: FP32-batched vs FP64-plain @ rel-L2"
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
ROTATE_BYTES = 1 << 30  # buffer pools rotated so that reads come from HBM, not the 256 MB MALL
CPU_REP_SECONDS = 0.6   # CPU baseline sample sized for about this long per pass (3 passes, best)
CONFIG_IDS = ("c1", "c2", "c3", "c4", "c5")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--only", nargs="*", default=None, choices=CONFIG_IDS,
                    help="configurations to run (default: all; c2 is always the headline)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only time a plain device-to-device copy of this many bytes (the reference kernel for "
                         "the HBM traffic counters and the achievable copy bandwidth)")
    ap.add_argument("--uneven-class", default=None,
                    choices=("all", "fused", "smooth4", "blue4_small", "blue4_big"),
                    help="only time the FP32 multi-length plan over one size class of the 10,000-length uneven "
                         "batch (for per-class kernel counters)")
    return ap.parse_args()


def uneven_class_timing(torch, name, reps=5):
    """One persistent multi-length launch over a size class of the uneven
    batch, timed with HIP events over `reps` back-to-back executions."""
    from chfft_amd import MultiFFT
    from chfft_amd.workloads import uneven_class, uneven_lengths
    lens = uneven_class(uneven_lengths(), name)
    fft = MultiFFT(lens, np.float32)
    x = torch.empty(fft.total, dtype=torch.complex64, device="cuda")
    x.view(torch.float32).normal_()
    y = torch.empty_like(x)
    st = torch.cuda.current_stream().cuda_stream
    fft.plan.exec_ptr(x.data_ptr(), y.data_ptr(), "forward", st)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fft.plan.exec_ptr(x.data_ptr(), y.data_ptr(), "forward", st)
    e1.record()
    torch.cuda.synchronize()
    fft.plan.check()
    ms = e0.elapsed_time(e1) / reps
    return {"uneven_class": name, "transforms": len(lens), "points": fft.total, "ms": round(ms, 4),
            "schedule": fft.plan.describe().strip().splitlines()[0]}


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
    """A pool of input/output buffer pairs, rotated when one pair is small
    enough to stay in the 256 MB memory-side cache."""

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


class Fixed:
    """Explicit (input, output) tensor pairs, rotated."""

    def __init__(self, pairs):
        self.pairs = pairs
        self.k = len(pairs)

    def pair(self, i):
        a, b = self.pairs[i % self.k]
        return a.data_ptr(), b.data_ptr()


def time_plan(torch, plan, bufs, steps, warmup, method="forward", blocks=1):
    """ms per step over `steps` steps bracketed by events on the launch
    stream; with blocks > 1 the median of that many such blocks."""
    s = torch.cuda.current_stream()
    st = s.cuda_stream
    for i in range(warmup):
        a, b = bufs.pair(i)
        plan.exec_ptr(a, b, method, st)
    torch.cuda.synchronize()
    res = []
    for _ in range(blocks):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for i in range(steps):
            a, b = bufs.pair(i)
            plan.exec_ptr(a, b, method, st)
        e1.record(s)
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
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def traffic_for(key):
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            return json.load(f).get(key)
    return None


def launch_shapes(plan, method="forward"):
    """[grid, block, lds, load op, store op] of each launch of `method`'s
    schedule (the key that ties rocprofv3's per-dispatch rows to this case)."""
    import re
    d = plan.describe()
    part = d.split("inverse:")[0] if method.startswith("forward") else d.split("inverse:")[-1]
    return [[int(g), int(b), int(l), int(a), int(c)]
            for b, g, l, a, c in re.findall(r"block=(\d+) grid=(\d+) lds=(\d+) ops=(\d+)/(\d+)", part)]


def roofline(alg_bytes, kms, key, ms_step):
    """Whole-transform roofline of one step: algorithmic bytes over the step
    time (ms_step: HIP events on the launch stream around K back-to-back
    steps, which matches the summed kernel durations of a rocprofv3 trace;
    `launches_ms` brackets each launch on its own and runs 10-30 % above the
    trace, from the gaps the brackets add)."""
    ach = alg_bytes / (ms_step * 1e6)
    tr = traffic_for(key)
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": tr, "algorithmic_bytes": alg_bytes,
            "kernel": f"whole schedule ({len(kms)} launch{'es' if len(kms) > 1 else ''} per step)", "ms": ms_step,
            "launches_ms": [round(v, 5) for v in kms], "traffic_key": key}


def cpu_rate(run, flop_per_unit, full_units, threads, label, note):
    """Times run(units, threads) -> seconds per pass on a sample of `units`
    grown until one pass takes about CPU_REP_SECONDS (capped at the full
    workload); GFLOP/s over the sample."""
    units = max(1, min(full_units, 4))
    while True:
        sec = run(units, threads)
        if sec >= CPU_REP_SECONDS * 0.5 or units >= full_units:
            break
        units = min(full_units, max(units + 1, int(units * CPU_REP_SECONDS / max(sec, 1e-4))))
    return {"value": round(flop_per_unit * units / sec / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port", "sample": f"{label}: {units} of {full_units}, best of 3 passes, plans built before "
                                      f"the clock; oracle {note}"}


def cpu_both(run, flop_per_unit, full_units, label, note):
    threads = min(os.cpu_count() or 1, 16)
    d = cpu_rate(run, flop_per_unit, full_units, threads, label, note)
    one = cpu_rate(run, flop_per_unit, max(1, full_units), 1, label, note)
    d["value_1core"] = one["value"]
    d["sample_1core"] = one["sample"]
    return d


# ---------------------------------------------------------------- cases
def c2c_case(torch, n, batch, f64, args, seed, uniform=False, blocks=1, check=8, cpu_note=None, traffic_key=None):
    """CFft1D forward on [batch, n]: throughput, per-launch times, rel-L2 of
    `check` transforms against the FP64 oracle, roofline, CPU baseline."""
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
    ms = time_plan(torch, p, bufs, args.steps, args.warmup, blocks=blocks)
    kms = per_kernel_ms(torch, p, bufs, max(3, args.steps // 2))
    tdt = torch.complex128 if f64 else torch.complex64
    k = min(check, batch)
    fl = k * n * (4 if f64 else 2)
    x = bufs.ins[0][:fl].view(tdt).cpu().numpy().reshape(k, n)
    p.exec_ptr(bufs.ins[0].data_ptr(), bufs.outs[0].data_ptr(), "forward", torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    y = bufs.outs[0][:fl].view(tdt).cpu().numpy().reshape(k, n)
    xr = x.astype(np.complex128)
    ref = oracle.c2c(xr, "forward") if n <= 70000 else np.fft.fft(xr)
    flop = 5.0 * n * math.log2(n)
    alg = (32.0 if f64 else 16.0) * n * batch
    rec = {"n": n, "batch": batch, "dtype": "f64" if f64 else "f32", "ms_per_step": round(ms, 5),
           "gflops": round(flop * batch / (ms * 1e6), 2),
           "rel_l2": rel_l2(y, ref), "buffers_rotated": bufs.k, "launch_shapes": launch_shapes(p),
           "roofline": roofline(alg, kms, traffic_key or f"c2c_{n}x{batch}_{'f64' if f64 else 'f32'}", ms)}
    if cpu_note is not None:
        big = bufs.ins[0].view(tdt)[:batch * n].cpu().numpy().reshape(batch, n)

        def run(units, threads):
            return oracle.time_c2c(big[:units], "forward", threads=threads, reps=3)
        rec["cpu_baseline"] = cpu_both(run, flop, batch, f"CFft1D<{'f64' if f64 else 'f32'}>::forward, N={n}, "
                                                         f"transforms", cpu_note)
    return rec, p, bufs


def headline_and_c2(torch, args, cpu_note):
    N, BATCH, SEED = 43200, 512, 0x5EED0002
    r32, p32, b32 = c2c_case(torch, N, BATCH, False, args, SEED, cpu_note=cpu_note)
    r64, _, _ = c2c_case(torch, N, BATCH, True, args, SEED, cpu_note=cpu_note)
    sched = p32.describe().split("inverse:")[0].replace("forward:", "").strip()
    kms = r32["roofline"]["launches_ms"]
    dom = int(np.argmax(kms))
    # the dominant launch on its own: it moves one read and one write of its
    # chunk's data, half of each transform's passes
    chunks = max(1, len(kms) // 2)
    launch_bytes = 16.0 * N * BATCH / chunks
    head = {
        "metric": METRIC,
        "value": r32["gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: complex entries i.i.d. N(0,1), torch generator seed 0x5EED0002",
        "config": {"workload": "CFft1D forward, mixed radix N=43200 (2^6 3^3 5^2), batch 512", "n": N,
                   "batch": BATCH, "schedule": sched},
        "rel_l2_f32_vs_f64_oracle": r32["rel_l2"],
        "fp64_plain": {"value": r64["gflops"], "unit": "GFLOP/s", "ms_per_step": r64["ms_per_step"],
                       "rel_l2_vs_f64_oracle": r64["rel_l2"], "roofline_frac": r64["roofline"]["frac"],
                       "launches_ms": r64["roofline"]["launches_ms"]},
        "speedup_f32_over_f64": round(r64["ms_per_step"] / r32["ms_per_step"], 3),
        "roofline": dict(r32["roofline"], dominant_launch={
            "index": dom, "ms": kms[dom], "bytes": launch_bytes,
            "achieved": round(launch_bytes / (kms[dom] * 1e6), 1),
            "what": sched.splitlines()[dom].strip() if dom < len(sched.splitlines()) else ""}),
        "cpu_baseline": r32.get("cpu_baseline"),
    }
    cfg = {"id": "c2", "workload": "CFft1D forward, N=43200 (2^6 3^3 5^2) x 512, complex N(0,1); FP32 form and "
                                   "FP64 plain method", "cases": [r32, r64]}
    return head, cfg


def cfg_c1(torch, args, cpu_note):
    recs = []
    for f64 in (True, False):
        r, p, b = c2c_case(torch, 4096, 256, f64, args, 0x5EED0001, uniform=True, blocks=5, cpu_note=cpu_note)
        r["ms_backward"] = round(time_plan(torch, p, b, args.steps, args.warmup, "backward", blocks=5), 5)
        recs.append(r)
    return {"id": "c1", "workload": "CFft1D N=4096 x 256, uniform[-1,1); forward (timed) then backward (1/N); "
                                    "FP64 plain method and FP32 form", "cases": recs}


def cfg_c3(torch, args, cpu_note):
    recs = []
    for n, batch in ((4093, 2048), (10007, 512), (65537, 64)):
        for f64 in (False, True):
            r, p, _ = c2c_case(torch, n, batch, f64, args, 0x5EED0003, uniform=True, blocks=5,
                               cpu_note=cpu_note if not f64 else None)
            d = p.describe()
            r["padded_M"] = int(d.split(" M L=")[1].split()[0]) if " M L=" in d else bluestein_m(d)
            recs.append(r)
    return {"id": "c3", "workload": "Bluestein chirp-z, N in {4093 x 2048, 10007 x 512, 65537 x 64}, "
                                    "uniform[-1,1); chirp index n^2 mod 2N in exact integers; padded_M is the "
                                    "inner FFT length each case ran", "cases": recs}


def bluestein_m(desc):
    """Padded length of a four-step Bluestein schedule: product of its column
    and row lengths (from the plan description)."""
    import re
    ls = [int(v) for v in re.findall(r"L=(\d+)", desc.split("inverse:")[0])]
    return ls[0] * ls[1] if len(ls) >= 2 else None


def cfg_c4(torch, args, cpu_note):
    import oracle
    from chfft_amd import Plan
    st = torch.cuda.current_stream().cuda_stream
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

            def fill(a, i, seed=0x5EED0004):
                g = torch.Generator(device="cuda")
                g.manual_seed(seed + i)
                if kind == "r2c":
                    a.normal_(generator=g)
                else:
                    a.uniform_(-1, 1, generator=g)
            b = Bufs(torch, ib, ob, fill)
            ms = time_plan(torch, p, b, args.steps, args.warmup, blocks=5)
            kms = per_kernel_ms(torch, p, b, max(3, args.steps // 2))
            rdt = torch.float64 if f64 else torch.float32
            rows = 8
            xr = b.ins[0].view(rdt)[:rows * n].cpu().numpy().reshape(rows, n).astype(np.float64)
            p.exec_ptr(b.ins[0].data_ptr(), b.outs[0].data_ptr(), "forward", st)
            torch.cuda.synchronize()
            if kind == "r2c":
                ref = oracle.rfft(xr, "forward")
                got = b.outs[0].view(torch.complex128 if f64 else torch.complex64)[:rows * (n // 2 + 1)]
                got = got.cpu().numpy().reshape(rows, n // 2 + 1)
            else:
                ref = oracle.dct(xr, 2 if kind == "dct2" else 3)
                got = b.outs[0].view(rdt)[:rows * n].cpu().numpy().reshape(rows, n)
            flop = 2.5 * n * math.log2(n)
            rec = {"kind": kind, "n": n, "batch": batch, "dtype": "f64" if f64 else "f32",
                   "ms_per_step": round(ms, 5), "gflops": round(flop * batch / (ms * 1e6), 2),
                   "rel_l2": rel_l2(got, ref), "buffers_rotated": b.k, "launch_shapes": launch_shapes(p),
                   "roofline": roofline(alg, kms, f"{kind}_{n}x{batch}_{'f64' if f64 else 'f32'}", ms)}
            if cpu_note is not None and not f64:
                full = b.ins[0].view(rdt)[:batch * n].cpu().numpy().reshape(batch, n)

                def run(units, threads, full=full, kind=kind, n=n):
                    return oracle.time_op("rfft" if kind == "r2c" else kind, full[:units], n, threads=threads)
                label = {"r2c": "RFft1D<f32>::forward", "dct2": "Dct1D<f32> type II forward",
                         "dct3": "Dct1D<f32> type III forward"}[kind] + f", N={n}, rows"
                rec["cpu_baseline"] = cpu_both(run, flop, batch, label, cpu_note)
            recs.append(rec)
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
        sd = sig.astype(np.float64 if f64 else np.float32)
        xs = torch.from_numpy(sd).cuda()
        ys = torch.empty(frames * n // 2, dtype=xs.dtype, device="cuda")
        # copies of the signal rotated like the other cases, so that the
        # signal is read from HBM rather than the 256 MB cache
        kr = max(1, min(16, ROTATE_BYTES // (xs.numel() * xs.element_size() + ys.numel() * ys.element_size())))
        bufs = Fixed([(xs, ys)] + [(xs.clone(), torch.empty_like(ys)) for _ in range(kr - 1)])
        ms = time_plan(torch, p, bufs, args.steps, args.warmup, blocks=5)
        kms = per_kernel_ms(torch, p, bufs, max(3, args.steps // 2))
        p.exec_ptr(xs.data_ptr(), ys.data_ptr(), "forward", st)
        torch.cuda.synchronize()
        chk = 64
        ref = oracle.mdct(sig[:(chk - 1) * hop + n].astype(np.float64), win, "forward", n=n, hop=hop, frames=chk)
        got = ys[:chk * n // 2].cpu().numpy().reshape(chk, n // 2)
        alg = (S + frames * n // 2) * (8 if f64 else 4)
        flop = 2.5 * n * math.log2(n)
        rec = {"kind": "mdct", "n": n, "hop": hop, "frames": frames, "dtype": "f64" if f64 else "f32",
               "ms_per_step": round(ms, 5), "gflops": round(flop * frames / (ms * 1e6), 2),
               "rel_l2": rel_l2(got, ref), "buffers_rotated": kr, "launch_shapes": launch_shapes(p),
               "roofline": roofline(alg, kms, f"mdct_{n}x{frames}_{'f64' if f64 else 'f32'}", ms)}
        if cpu_note is not None and not f64:
            w32 = oracle.window(n, "sine", np.float32)

            def run(units, threads):
                return oracle.time_op("mdct", sd, n, threads=threads, batch=units, hop=hop, window=w32)
            rec["cpu_baseline"] = cpu_both(run, flop, frames, "Mdct1D<f32>::forward (sine), N=2048 hop 1024, "
                                                               "frames", cpu_note)
        del bufs
        recs.append(rec)
    return {"id": "c4", "workload": "RFft1D 8192 x 4096 (real N(0,1)); Dct1D II and III 1024 x 16384 "
                                    "(uniform[-1,1)); Mdct1D 2048 (sine window, hop 1024) over a 2^24-sample "
                                    "signal of 8 sinusoids + N(0, 0.01) noise, 16383 frames", "cases": recs}


def cfg_c5(torch, args, cpu_note):
    import oracle
    from chfft_amd import MultiFFT, Plan
    from chfft_amd.workloads import uneven_lengths
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
    fft.plan.check()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps // 4)
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    fft.plan.check()
    ms_graph = e0.elapsed_time(e1) / reps
    one = Fixed([(x, y)])
    ms_multi = time_plan(torch, fft.plan, one, reps, 1)
    kms_multi = per_kernel_ms(torch, fft.plan, one, 3)
    fft.plan.check()
    k2 = max(1, ROTATE_BYTES // (2 * x2.numel() * x2.element_size()))
    pool2 = Fixed([(x2, y2)] + [(x2.clone(), torch.empty_like(y2)) for _ in range(k2 - 1)])
    ms_2d = time_plan(torch, p2, pool2, args.steps, args.warmup, blocks=5)
    kms_2d = per_kernel_ms(torch, p2, pool2, max(3, args.steps // 2))
    del pool2
    # rel-L2 per octave bucket on a sample of transforms (numpy FP64 reference)
    xs = x.cpu().numpy()
    ys = y.cpu().numpy()
    buckets = {}
    for i in range(0, len(lens), 7):
        o0, o1 = fft.offsets[i], fft.offsets[i + 1]
        n = lens[i]
        e = rel_l2(ys[o0:o1], np.fft.fft(xs[o0:o1].astype(np.complex128)))
        buckets.setdefault(int(math.log2(n)), []).append(e)
    bucket_err = {f"2^{b}..2^{b + 1}": {"max": max(v), "count": len(v), "budget": 1e-6 * (b + 1)}
                  for b, v in sorted(buckets.items())}
    x2h = x2.cpu().numpy().reshape(2048, 2048).astype(np.complex128)
    e2d = rel_l2(y2.cpu().numpy().reshape(2048, 2048), np.fft.fft2(x2h))
    flop_multi = sum(5.0 * n * math.log2(n) for n in lens if n > 1)
    flop_2d = 5.0 * 2048 * 2048 * math.log2(2048 * 2048)
    rm = {"part": "multi", "transforms": len(lens), "points": fft.total, "dtype": "f32",
          "ms_per_step": round(ms_multi, 4), "gflops": round(flop_multi / (ms_multi * 1e6), 2),
          "autotune_choice": choice, "rel_l2_per_bucket": bucket_err,
          "schedule": fft.plan.describe().strip().splitlines()[0],
          "launch_shapes": [[int(fft.plan.describe().split("grid=")[1].split()[0]), 1024, 0, -1, -1]],
          "roofline": roofline(16.0 * fft.total, kms_multi, "multi_uneven10000_f32", ms_multi)}
    r2 = {"part": "cfft2d", "m": 2048, "n": 2048, "dtype": "f32", "ms_per_step": round(ms_2d, 4),
          "gflops": round(flop_2d / (ms_2d * 1e6), 2), "rel_l2": e2d, "launch_shapes": launch_shapes(p2),
          "roofline": roofline(16.0 * 2048 * 2048, kms_2d, "c2c2d_2048x2048_f32", ms_2d)}
    rg = {"part": "graph (multi + 2-D)", "ms_per_step": round(ms_graph, 4),
          "gflops": round((flop_multi + flop_2d) / (ms_graph * 1e6), 2)}
    if cpu_note is not None:
        # a sample of the uneven batch: every 50th transform (the same size mix)
        idx = list(range(0, len(lens), 50))
        sl = [lens[i] for i in idx]
        sx = np.concatenate([xs[fft.offsets[i]:fft.offsets[i + 1]] for i in idx])
        sflop = sum(5.0 * n * math.log2(n) for n in sl if n > 1)

        def run_m(units, threads):
            k = min(units, len(sl))
            return oracle.time_lens(sx[:sum(sl[:k])], sl[:k], threads=threads, reps=3)
        d = cpu_both(run_m, sflop / len(sl), len(sl), "CFft1D<f32>::forward over every 50th transform of the "
                                                      "uneven batch (one plan per transform)", cpu_note)
        rm["cpu_baseline"] = d
        x2s = x2.cpu().numpy().reshape(1, 2048, 2048)

        def run_2d(units, threads):
            return oracle.time_op("c2c2d", x2s, 2048, m=2048, threads=threads)
        r2["cpu_baseline"] = cpu_both(run_2d, flop_2d, 1, "CFft2D<f32>::forward 2048 x 2048, matrices", cpu_note)
    return {"id": "c5", "workload": "10,000 uneven FP32 transforms (log-uniform 16..2^20, 50/50 2*3*5-smooth / "
                                    "prime, <= 2^28 points, N(0,1)) in one persistent launch, plus CFft2D 2048 x "
                                    "2048; both replayed through one HIP graph", "cases": [rm, r2, rg]}


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
    if args.uneven_class:
        print(json.dumps(uneven_class_timing(torch, args.uneven_class)), flush=True)
        return
    t0 = time.time()
    cpu_note = None
    tmp = None
    if not args.no_cpu:
        import oracle
        tmp = tempfile.TemporaryDirectory(prefix="chfft_oracle_")
        cpu_note = oracle.use_native(tmp.name)
    only = set(args.only) if args.only else set(CONFIG_IDS)
    head, c2 = headline_and_c2(torch, args, cpu_note)
    configs = []
    for cid, fn in (("c1", cfg_c1), ("c2", None), ("c3", cfg_c3), ("c4", cfg_c4), ("c5", cfg_c5)):
        if cid not in only:
            continue
        print(f"bench: {cid} ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
        configs.append(c2 if cid == "c2" else fn(torch, args, cpu_note))
        torch.cuda.empty_cache()
    head["configs"] = configs
    head["bench_seconds"] = round(time.time() - t0, 1)
    print(json.dumps(head), flush=True)
    if tmp is not None:
        tmp.cleanup()


if __name__ == "__main__":
    main()
