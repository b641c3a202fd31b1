"""Benchmark: CFft1D forward on the mixed-radix workload N = 43200 (2^6 3^3 5^2),
batch 512, complex N(0,1) entries from a seed, on one MI355X.

One step = one forward transform of the whole batch, inputs already resident
in HBM. The headline value is the FP32 form's throughput in GFLOP/s (nominal
5 N log2 N per transform); the FP64 plain method runs on the same inputs and
is reported beside it, each with its rel-L2 against the FP64 oracle. The
roofline object covers the dominant kernel of the FP32 schedule, timed with HIP
events on the stream the kernels are launched on. The cpu_baseline object is
the oracle (the C++ restatement of chfft's algorithm) timed on the host cores
over the same workload.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, BATCH, SEED = 43200, 512, 0x5EED0002
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_c2c_43200x512.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


def timed_steps(torch, plan, src, dst, steps, stream):
    """Whole-job time of `steps` forward transforms (ms per step)."""
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(torch.cuda.current_stream())
    for _ in range(steps):
        plan.exec_ptr(src.data_ptr(), dst.data_ptr(), "forward", stream)
    e1.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def per_kernel_ms(torch, plan, src, dst, steps, stream):
    """Average duration of each launch of the schedule, bracketed by events
    recorded on the same stream the kernels run on."""
    from chfft_amd import _lib
    import ctypes
    L = _lib.lib()
    n = L.chfft_plan_launch_count(plan.handle, 0)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n + 1)] for _ in range(steps)]
    s = torch.cuda.current_stream()
    torch.cuda.synchronize()
    for k in range(steps):
        for i in range(n):
            ev[k][i].record(s)
            _lib.check(L.chfft_exec_range(plan.handle, ctypes.c_void_p(src.data_ptr()),
                                          ctypes.c_void_p(dst.data_ptr()), 0, i, 1, ctypes.c_void_p(stream)))
        ev[k][n].record(s)
    torch.cuda.synchronize()
    return [sum(ev[k][i].elapsed_time(ev[k][i + 1]) for k in range(steps)) / steps for i in range(n)]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the batch dimension would be split across GPUs with no exchange", file=sys.stderr)
        sys.exit(2)
    import torch
    from chfft_amd import Plan

    torch.manual_seed(SEED)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    x64 = torch.complex(torch.randn(BATCH, N, dtype=torch.float64, device=dev),
                        torch.randn(BATCH, N, dtype=torch.float64, device=dev))
    x32 = x64.to(torch.complex64)
    y32 = torch.empty_like(x32)
    y64 = torch.empty_like(x64)
    p32 = Plan("c2c", False, BATCH, N)
    p64 = Plan("c2c", True, BATCH, N)

    for _ in range(args.warmup):
        p32.exec_ptr(x32.data_ptr(), y32.data_ptr(), "forward", stream)
        p64.exec_ptr(x64.data_ptr(), y64.data_ptr(), "forward", stream)
    torch.cuda.synchronize()

    ms32 = timed_steps(torch, p32, x32, y32, args.steps, stream)
    ms64 = timed_steps(torch, p64, x64, y64, args.steps, stream)
    k32 = per_kernel_ms(torch, p32, x32, y32, max(3, args.steps // 2), stream)
    k64 = per_kernel_ms(torch, p64, x64, y64, max(3, args.steps // 2), stream)

    flop = 5.0 * N * math.log2(N) * BATCH
    gf32 = flop / (ms32 * 1e6)
    gf64 = flop / (ms64 * 1e6)
    bytes32 = 16.0 * N * BATCH  # one read of the input + one write of the output, FP32
    bytes64 = 32.0 * N * BATCH

    # accuracy on a sample of the batch against the FP64 oracle (same inputs)
    import oracle
    nchk = 8
    xs = x32[:nchk].cpu().numpy().astype(np.complex128)
    ref32 = oracle.c2c(xs, "forward")
    err32 = rel_l2(y32[:nchk].cpu().numpy().astype(np.complex128), ref32)
    ref64 = oracle.c2c(x64[:nchk].cpu().numpy(), "forward")
    err64 = rel_l2(y64[:nchk].cpu().numpy(), ref64)

    # dominant kernel of the FP32 schedule: each launch of the four-step reads
    # and writes the whole batch once (8 N bytes in + 8 N bytes out per transform)
    dom = int(np.argmax(k32))
    achieved = bytes32 / (k32[dom] * 1e6)
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            t = json.load(f)
        traffic = t.get("per_launch_bytes", {}).get(str(dom))

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        xh = x32[:64].cpu().numpy()
        sec = oracle.time_c2c(xh, "forward", threads=threads, reps=3)
        cpu = {"value": round(5.0 * N * math.log2(N) * 64 / sec / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle (C++ restatement of chfft's mixed-radix CFft1D<f32>::forward) on 64 of the "
                         f"512 transforms of N={N}, best of 3 passes, plan built before timing"}

    line = {
        "metric": "FFT GFLOP/s (5Nlog2N) + % HBM roofline: FP32-batched vs FP64-plain @ rel-L2",
        "value": round(gf32, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: complex entries i.i.d. N(0,1), torch seed 0x5EED0002",
        "config": {"workload": "CFft1D forward, mixed radix N=43200 (2^6 3^3 5^2), batch 512",
                   "n": N, "batch": BATCH, "schedule": p32.describe().split("inverse:")[0].strip()},
        "hbm_frac_whole_transform": round(bytes32 / (ms32 * 1e6) / HBM_PEAK_GBS, 4),
        "rel_l2_f32_vs_f64_oracle": err32,
        "fp64_plain": {"value": round(gf64, 2), "unit": "GFLOP/s", "ms_per_step": round(ms64, 4),
                       "hbm_frac_whole_transform": round(bytes64 / (ms64 * 1e6) / HBM_PEAK_GBS, 4),
                       "rel_l2_vs_f64_oracle": err64, "kernel_ms": [round(v, 4) for v in k64]},
        "speedup_f32_over_f64": round(ms64 / ms32, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"launch {dom} of {len(k32)}", "kernel_ms": [round(v, 4) for v in k32]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
