"""Tuning table: autotuned planner overrides per (plan kind, precision, shape).

The table (``tuning.json`` beside this file) maps a plan key to the
chfft_plan_tune spec the autotuner found fastest on an MI355X, with every
candidate's time. ``Plan`` applies an entry's spec when the entry is marked
``applied`` (its best candidate beat the planner's own choice by at least
``MIN_GAIN``); otherwise the planner's choice stands. The autotuner's role is
the one the reference's fixed factor order plays
(``prime_factorization.rs:66-113``): choosing how a length is split into
passes, here together with the GPU-side knobs (workgroup shape, four-step
split, chunking, kernel family).

``CHFFT_TUNING=off`` disables the table; ``CHFFT_TUNING=<path>`` reads another
file.

Regenerate on a GPU (the device buffers are overwritten):

    python -m chfft_amd.tuning --out chfft_amd/tuning.json
"""
import argparse
import hashlib
import json
import os
import warnings

import numpy as np

MIN_GAIN = 1.03
_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
_TABLE = None


def plan_key(kind, f64, batch, args):
    """Key of a plan in the table, or None for kinds the tuner does not cover."""
    prec = "f64" if f64 else "f32"
    if kind == "c2c":
        return f"c2c/{prec}/{int(args[0])}x{int(batch)}"
    if kind == "c2c_multi":
        lens = np.ascontiguousarray(args[0], dtype=np.uint64)
        digest = hashlib.sha1(lens.tobytes()).hexdigest()[:12]
        return f"multi/{prec}/{len(lens)}/{int(lens.sum())}/{digest}"
    return None


def _table():
    global _TABLE
    if _TABLE is None:
        src = os.environ.get("CHFFT_TUNING", "")
        _TABLE = {}
        if src.lower() != "off":
            path = src or _PATH
            if os.path.exists(path):
                with open(path) as f:
                    _TABLE = json.load(f).get("entries", {})
    return _TABLE


def spec_for(kind, f64, batch, args):
    """The spec to apply to a new plan, or None (planner's own choice)."""
    key = plan_key(kind, f64, batch, args)
    if key is None:
        return None
    e = _table().get(key)
    if not e or not e.get("applied"):
        return None
    return e["spec"]


def apply(plan):
    """Apply the table's spec for `plan` (a chfft_amd.Plan); a spec this
    device's plan cannot take leaves the planner's choice, with a warning."""
    spec = spec_for(plan.kind, plan.f64, plan.batch, plan.args)
    if spec is None:
        return None
    from . import _lib
    try:
        plan.tune(spec)
    except _lib.ChfftError as e:
        warnings.warn(f"tuning spec {spec!r} not applied to {plan.kind}{plan.args[:1]}: {e}")
        return None
    return spec


def parse_report(report):
    """[(index, spec, ms)] from the report of chfft_plan_tuning."""
    out = []
    for line in report.splitlines():
        line = line.strip()
        if not line.startswith("[") or "] " not in line or ":" not in line:
            continue
        idx, rest = line[1:].split("] ", 1)
        spec, ms = rest.rsplit(":", 1)
        try:
            out.append((int(idx), spec.strip(), float(ms)))
        except ValueError:
            continue
    return out


def tune_case(kind, f64, batch, args, reps=20):
    """Autotune one plan on fresh device buffers; the table entry."""
    import torch
    from .api import Plan
    plan = Plan(kind, f64, batch, *args, tuned=False)
    ib, ob = plan.io_bytes("forward")
    src = torch.randn(max(ib, ob) // 4 + 16, device="cuda").contiguous()
    dst = torch.empty_like(src)
    best = plan.autotune(src.data_ptr(), dst.data_ptr(), reps=reps,
                         stream=torch.cuda.current_stream().cuda_stream)
    spec, report = plan.tuning()
    cands = parse_report(report)
    default_ms = next((ms for i, s, ms in cands if i == 0), None)
    best_ms = next((ms for i, s, ms in cands if i == best), None)
    gain = (default_ms / best_ms) if default_ms and best_ms else 1.0
    return {
        "spec": spec,
        "ms": best_ms,
        "default_ms": default_ms,
        "gain": round(gain, 4),
        "applied": bool(best != 0 and gain >= MIN_GAIN),
        "candidates": [[s, ms] for _, s, ms in cands],
    }


def cases():
    """The benchmark shapes (both precisions where the benchmark runs both)
    and lengths without a length-specialised kernel."""
    from .workloads import uneven_lengths
    out = []
    for f64 in (False, True):
        for n, b in ((4096, 256), (43200, 512), (4093, 2048), (10007, 512), (65537, 64)):
            out.append(("c2c", f64, b, (n,)))
    for n, b in ((1000, 4096), (3000, 1024), (6000, 1024), (12000, 512), (100000, 64)):
        for f64 in (False, True):
            out.append(("c2c", f64, b, (n,)))
    out.append(("c2c_multi", False, 1, (uneven_lengths(),)))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=_PATH)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--only", default="", help="substring of the plan keys to tune")
    a = ap.parse_args(argv)
    import torch
    entries = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            entries = json.load(f).get("entries", {})
    for kind, f64, batch, args in cases():
        key = plan_key(kind, f64, batch, args)
        if a.only and a.only not in key:
            continue
        e = tune_case(kind, f64, batch, args, reps=a.reps)
        entries[key] = e
        print(f"{key}: {e['spec']} {e['ms']:.4f} ms (default {e['default_ms']:.4f}, "
              f"{len(e['candidates'])} candidates, applied={e['applied']})", flush=True)
    doc = {
        "device": torch.cuda.get_device_name(0),
        "reps": a.reps,
        "min_gain": MIN_GAIN,
        "entries": dict(sorted(entries.items())),
    }
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
