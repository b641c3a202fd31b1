"""CPU tests of the tuning table logic (chfft_amd.tuning) and of the
committed table's consistency; the GPU side (every candidate passes parity)
is tests/test_gpu_tuning.py."""
import ctypes
import json
import os

import numpy as np
import pytest

from chfft_amd import _lib, tuning

TABLE = os.path.join(os.path.dirname(tuning.__file__), "tuning.json")
KEYS = {"generic", "order", "tpb", "tpt", "n1", "chunk", "sched", "grid", "unit"}


def test_plan_keys():
    assert tuning.plan_key("c2c", False, 512, (43200,)) == "c2c/f32/43200x512"
    assert tuning.plan_key("c2c", True, 1, (7,)) == "c2c/f64/7x1"
    k = tuning.plan_key("c2c_multi", False, 1, (np.array([3, 5, 8]),))
    assert k.startswith("multi/f32/3/16/") and len(k.split("/")[-1]) == 12
    assert k != tuning.plan_key("c2c_multi", False, 1, (np.array([5, 3, 8]),))
    assert tuning.plan_key("r2c", False, 4, (64,)) is None


def test_parse_report():
    rep = ("autotune (2 execs per candidate, ms per exec):\n  [0] default: 0.1400\n"
           "  [1] sched=1: 0.4490\n  [2] generic=1 order=2: -1.0000\n  -> [0] default")
    assert tuning.parse_report(rep) == [(0, "default", 0.14), (1, "sched=1", 0.449), (2, "generic=1 order=2", -1.0)]


def test_spec_for_applies_only_marked_entries(monkeypatch):
    table = {"c2c/f32/1000x8": {"spec": "generic=1 order=1", "applied": True},
             "c2c/f32/2000x8": {"spec": "generic=1", "applied": False}}
    monkeypatch.setattr(tuning, "_TABLE", table)
    assert tuning.spec_for("c2c", False, 8, (1000,)) == "generic=1 order=1"
    assert tuning.spec_for("c2c", False, 8, (2000,)) is None
    assert tuning.spec_for("c2c", True, 8, (1000,)) is None
    assert tuning.spec_for("dct", False, 8, (2, 1000, False)) is None


def test_tune_entry_points_reject_null_without_a_gpu():
    L = _lib.lib()
    assert L.chfft_plan_tune(None, b"default") == _lib.ERR_INVALID_ARGUMENT
    buf = ctypes.create_string_buffer(16)
    assert L.chfft_plan_tuning(None, buf, 16) == _lib.ERR_INVALID_ARGUMENT


@pytest.mark.skipif(not os.path.exists(TABLE), reason="no tuning table")
def test_committed_table_is_consistent():
    doc = json.load(open(TABLE))
    assert doc["entries"]
    for key, e in doc["entries"].items():
        assert key.startswith(("c2c/", "multi/")), key
        cands = e["candidates"]
        assert cands[0][0] == "default"
        specs = [c[0] for c in cands]
        assert e["spec"] in specs, key
        for s in specs:
            if s == "default":
                continue
            for kv in s.split():
                k, v = kv.split("=")
                assert k in KEYS and int(v) == int(v), (key, s)
        timed = [c for c in cands if c[1] > 0]
        best = min(timed, key=lambda c: c[1])
        assert e["ms"] == pytest.approx(best[1]), key
        assert e["default_ms"] == pytest.approx(cands[0][1]), key
        gain = e["default_ms"] / e["ms"]
        assert e["applied"] == (e["spec"] != "default" and gain >= doc["min_gain"]), key
