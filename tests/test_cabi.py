"""CPU tests of the C ABI library: it loads, exports every function that
include/chfft_hip.h declares, and rejects invalid arguments with the status
codes the reference's panics map to — all without a GPU (argument checks run
before any HIP call)."""
import ctypes
import os
import re

import numpy as np
import pytest

import chfft_amd
from chfft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "chfft_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(chfft_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib._signatures())


def test_status_strings():
    L = _lib.lib()
    assert L.chfft_status_string(1) == b"invalid length"
    assert L.chfft_status_string(0) == b"ok"


def plan_rc(fn, *args):
    h = ctypes.c_void_p()
    return getattr(_lib.lib(), fn)(ctypes.byref(h), *args), h


def test_invalid_lengths_before_any_gpu_call():
    # rfft1d.rs:74-76, dct1d.rs:54-56, mdct1d.rs:102-104: "invalid length"
    assert plan_rc("chfft_plan_r2c_1d", 11, 1, _lib.F64)[0] == _lib.ERR_INVALID_LENGTH
    assert plan_rc("chfft_plan_dct_1d", 2, 11, 0, 1, _lib.F64)[0] == _lib.ERR_INVALID_LENGTH
    w = np.ones(10)
    assert plan_rc("chfft_plan_mdct_1d", 10, w.ctypes.data, 1, 0, _lib.F64)[0] == _lib.ERR_INVALID_LENGTH
    assert plan_rc("chfft_plan_c2c_1d", 0, 1, _lib.F64)[0] == _lib.ERR_INVALID_LENGTH
    assert plan_rc("chfft_plan_c2c_2d", 0, 4, 1, _lib.F64)[0] == _lib.ERR_INVALID_LENGTH


def test_invalid_arguments():
    L = _lib.lib()
    assert plan_rc("chfft_plan_c2c_1d", 16, 1, 7)[0] == _lib.ERR_INVALID_ARGUMENT  # precision
    assert plan_rc("chfft_plan_c2c_1d", 16, 0, _lib.F32)[0] == _lib.ERR_INVALID_ARGUMENT  # batch
    assert plan_rc("chfft_plan_dct_1d", 4, 16, 0, 1, _lib.F64)[0] == _lib.ERR_INVALID_ARGUMENT  # dct type
    assert plan_rc("chfft_plan_mdct_1d", 16, None, 1, 0, _lib.F64)[0] == _lib.ERR_INVALID_ARGUMENT  # window
    assert L.chfft_plan_c2c_1d(None, 16, 1, _lib.F64) == _lib.ERR_INVALID_ARGUMENT
    assert L.chfft_exec(None, None, None, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.chfft_exec_host(None, None, None, 0) == _lib.ERR_INVALID_ARGUMENT
    assert L.chfft_plan_destroy(None) == _lib.OK


def test_python_mirror_errors_match_reference_panics():
    # rfft1d.rs:404-422, dct1d.rs:527-538, mdct1d.rs:468-486
    with pytest.raises(ValueError, match="invalid length"):
        chfft_amd.RFft1D.new(11)
    with pytest.raises(ValueError, match="invalid length"):
        chfft_amd.Dct1D.new(chfft_amd.DctType.Two, 11, False)
    with pytest.raises(ValueError, match="invalid length"):
        chfft_amd.Mdct1D.with_sine(10)
    f = chfft_amd.RFft1D.new(4)
    with pytest.raises(ValueError, match=r"invalid length \(soure: 5, rdft.len: 4\)"):
        f.forward(np.arange(5.0))
    d = chfft_amd.Dct1D.new(chfft_amd.DctType.Two, 8, False)
    with pytest.raises(ValueError, match=r"invalid length \(soure: 10, dct.len: 8\)"):
        d.forward(np.arange(10.0))
    m = chfft_amd.Mdct1D.with_sine(8)
    with pytest.raises(ValueError, match=r"invalid length \(soure: 8, mdct.len: 8\)"):
        m.backward(np.arange(8.0))


def test_length_one_is_identity_without_gpu():
    # cfft1d.rs:213-214: length <= 1 returns the input unchanged (no plan)
    f = chfft_amd.CFft1D.new()
    x = np.array([3 + 4j])
    assert np.array_equal(f.forwardn(x), x)
    assert f.backward(np.zeros(0, complex)).shape == (0,)


def test_new_entry_points_reject_null_plans():
    L = _lib.lib()
    assert L.chfft_plan_check(None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.chfft_exec_mdct_ola(None, None, None, 4, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.chfft_status_string(_lib.ERR_INCOMPLETE).startswith(b"incomplete")
