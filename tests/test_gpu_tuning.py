"""GPU tests of the autotuner and the planner overrides (chfft_plan_autotune,
chfft_plan_tune).

Every candidate the autotuner times is a schedule the library may run, so
each one is checked against the oracle, not only the winner: fused lengths
(every radix order of the runtime-radix kernel, workgroup shapes around the
planner's, the length-specialised kernel), four-step lengths (persistent
schedule, runtime-radix kernels, chunk sizes, other splits, radix orders),
Bluestein lengths, and the multi-length launch (workgroup counts, work-unit
sizes, runtime-radix tiles). The committed tuning table's specs are built
at their own shapes and checked the same way, and specs a plan cannot take
are refused with the plan left working. Bars as in the other parity tests:
FP64 rel-L2 <= 1e-12 against the FP64 oracle, FP32 <= 1e-6 log2 N.
"""
import json
import os

import numpy as np
import pytest

import oracle
from chfft_amd import Plan, tuning
from chfft_amd._lib import ChfftError
from tests.gpu_util import TOL_F64, rel_l2, tol_f32

pytestmark = pytest.mark.gpu


def cuda_input(batch, n, f64, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    dt = torch.float64 if f64 else torch.float32
    x = torch.empty(batch * n * 2, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)
    return x.view(torch.complex128 if f64 else torch.complex64).view(batch, n)


def ref(x):
    x = np.asarray(x, dtype=np.complex128)
    return oracle.c2c(x, "forward") if x.shape[-1] <= 70000 else np.fft.fft(x)


def run_and_check(p, x, f64, n, rows, what):
    import torch
    y = torch.empty_like(x)
    p.exec(x, y, "forward")
    torch.cuda.synchronize()
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    tol = TOL_F64 if f64 else tol_f32(n)
    for b in rows:
        e = rel_l2(ys[b], ref(xs[b]))
        assert e <= tol, f"{what}: transform {b} rel-L2 {e:.3e} > {tol:.1e}"


# fused (specialised and runtime-radix), four-step, Bluestein fused / four-step
CASES = [(1000, 8), (4096, 8), (480, 64), (6144, 4), (43200, 6), (100000, 2), (4093, 4), (10007, 4), (65537, 2)]


@pytest.mark.parametrize("n,batch", CASES)
@pytest.mark.parametrize("f64", [False, True])
def test_every_candidate_passes_parity(n, batch, f64):
    import torch
    x = cuda_input(batch, n, f64, 0x70E0 + n)
    p = Plan("c2c", f64, batch, n, tuned=False)
    y = torch.empty_like(x)
    best = p.autotune(x.clone(), y, reps=2)
    spec, report = p.tuning()
    cands = tuning.parse_report(report)
    assert len(cands) >= 2, report
    assert cands[0][0] == 0 and cands[0][1] == "default"
    assert 0 <= best < len(cands) and spec == cands[best][1], (best, spec, report)
    rows = sorted({0, batch // 2, batch - 1})
    run_and_check(p, x, f64, n, rows, f"autotuned [{best}] {spec}")
    for idx, s, ms in cands:
        p.tune(s)
        assert p.tuning()[0] == s
        run_and_check(p, x, f64, n, rows, f"candidate [{idx}] {s}")


def test_multi_candidates_pass_parity():
    import torch
    from chfft_amd.workloads import uneven_lengths
    lens = uneven_lengths(count=300, cap=1 << 21, seed=0x70E1)
    total = int(np.sum(lens))
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.empty(2 * total, dtype=torch.float32, device="cuda").uniform_(-1, 1, generator=g)
    x = x.view(torch.complex64)
    p = Plan("c2c_multi", False, 1, lens, tuned=False)
    y = torch.empty_like(x)
    p.autotune(x.clone(), y, reps=2)
    cands = tuning.parse_report(p.tuning()[1])
    assert len(cands) >= 4
    offs = np.concatenate([[0], np.cumsum(lens)])
    pick = sorted({0, len(lens) // 2, len(lens) - 1, int(np.argmax(lens))})
    xs = x.cpu().numpy()
    refs = {i: np.fft.fft(xs[offs[i]:offs[i + 1]].astype(np.complex128)) for i in pick}
    assert any(s.startswith("tpb=") for _, s, _ in cands)  # generic tile widths are among them
    for idx, s, ms in cands:
        p.tune(s)
        p.exec(x, y, "forward")
        p.check()
        ys = y.cpu().numpy()
        for i in pick:
            e = rel_l2(ys[offs[i]:offs[i + 1]], refs[i])
            assert e <= tol_f32(lens[i]), f"[{idx}] {s}: item {i} (N={lens[i]}) rel-L2 {e:.3e}"
    # generic tile widths: powers of two up to 128 only; the plan keeps working
    for bad in ("tpb=3", "tpb=256", "tpt=4"):
        with pytest.raises(ChfftError):
            p.tune(bad)
    p.exec(x, y, "forward")
    p.check()
    ys = y.cpu().numpy()
    for i in pick:
        assert rel_l2(ys[offs[i]:offs[i + 1]], refs[i]) <= tol_f32(lens[i])


def test_invalid_specs_leave_the_plan_working():
    import torch
    n, batch = 4096, 4
    x = cuda_input(batch, n, False, 0x70E2)
    p = Plan("c2c", False, batch, n, tuned=False)
    for bad in ("tpb=3 tpt=1", "bogus=1", "order=99", "n1=7", "grid=4", "generic=2", "tpt=abc"):
        with pytest.raises(ChfftError):
            p.tune(bad)
        run_and_check(p, x, False, n, [0, batch - 1], f"after refused {bad!r}")
    q = Plan("c2c", False, 2, 43200, tuned=False)
    with pytest.raises(ChfftError):
        q.tune("tpb=2")  # tile knobs are for fused lengths
    with pytest.raises(ChfftError):
        q.tune("n1=7")  # does not divide 43200
    q.tune("n1=240 chunk=1")
    run_and_check(q, cuda_input(2, 43200, False, 0x70E3), False, 43200, [0, 1], "n1=240 chunk=1")


def table_entries():
    path = os.path.join(os.path.dirname(tuning.__file__), "tuning.json")
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return sorted(json.load(f)["entries"].items())


@pytest.mark.parametrize("key,entry", [e for e in table_entries() if e[0].startswith("c2c/")],
                         ids=[e[0] for e in table_entries() if e[0].startswith("c2c/")])
def test_tuning_table_specs_pass_parity(key, entry):
    _, prec, shape = key.split("/")
    n, batch = (int(v) for v in shape.split("x"))
    f64 = prec == "f64"
    x = cuda_input(batch, n, f64, 0x70E4 + n)
    p = Plan("c2c", f64, batch, n, tuned=False)
    p.tune(entry["spec"])
    run_and_check(p, x, f64, n, sorted({0, batch // 2, batch - 1}), f"{key} {entry['spec']}")
    q = Plan("c2c", f64, batch, n)  # the table applied by the constructor
    assert q.tuned_spec == (entry["spec"] if entry["applied"] else None)
