"""The C ABI library loads, exports every entry point of include/altar_amd.h,
and its host-only functions match the oracle (CPU only, no GPU calls)."""
import ctypes
import os
import re

import numpy as np

from altar_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "altar_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(altar_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding declares every one of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_argument_errors_are_reported():
    h = ctypes.c_void_p()
    rc = _lib.lib().altar_source_create(ctypes.byref(h), 7, 5, 10, 0.25)
    assert rc == _lib.EINVAL
    assert b"unknown model" in _lib.lib().altar_last_error()
    rc = _lib.lib().altar_source_create(ctypes.byref(h), _lib.MOGI, 5, 10, 0.25)
    assert rc == _lib.OK
    bad = np.array([0, 2, 3, 9, 1], dtype=np.int32)   # offsets beyond P
    assert _lib.lib().altar_source_layout(h, _lib.ptr(bad), 5) == _lib.EINVAL
    llk = ctypes.c_void_p(8)
    assert _lib.lib().altar_source_data_llk(h, llk, 4, llk, _lib.FP32, None) == _lib.ESTATE
    assert _lib.lib().altar_source_destroy(h) == _lib.OK


def test_dbeta_grid_matches_oracle_bit_for_bit(oracle):
    rng = np.random.default_rng(5)
    for beta0, scale in ((0.0, 40.0), (0.3, 5.0), (0.9, 0.5), (0.0, 1e4)):
        llk = rng.normal(-500, scale, 4096)
        med = float(np.median(llk))
        w = np.empty_like(llk)
        b, c = ctypes.c_double(), ctypes.c_double()
        _lib.call("altar_cov_dbeta_grid", _lib.ptr(llk), len(llk), med, beta0, 0.01, 1.0, _lib.ptr(w),
                  ctypes.byref(b), ctypes.byref(c))
        rb, rc, rw = oracle.dbeta_grid(llk, beta0)
        assert b.value == rb and c.value == rc
        assert np.array_equal(w, rw)


def test_multiplicities_match_oracle(oracle):
    rng = np.random.default_rng(6)
    C = 5000
    w = rng.random(C) ** 4
    w /= w.sum()
    count = np.empty(C, dtype=np.int32)
    _lib.call("altar_cov_multiplicities", _lib.ptr(w), C, 42, 3, _lib.ptr(count))
    r = oracle.uniforms(42, 3, 0, _lib.TAG_RESAMPLE, C)
    assert np.array_equal(count, oracle.multiplicities(w, r))


def test_scaling_matches_oracle(oracle):
    for a, r, u in ((10, 5, 85), (0, 0, 10), (100, 0, 0)):
        assert _lib.lib().altar_mh_scaling(8.0, 1.0, a, r, u) == oracle.mh_adjust_scaling(a, r, u)


def test_dbeta_brent_matches_oracle_bit_for_bit(oracle):
    """The host Brent solver against the oracle's restatement: same steps of
    Brent's method, so the same dbeta, COV and weights bit for bit."""
    import ctypes
    rng = np.random.default_rng(61)
    for C, scale, beta0 in ((3000, 40.0, 0.0), (1001, 3.0, 0.3), (4096, 0.2, 0.8), (500, 15.0, 0.95)):
        llk = rng.normal(-700, scale, C)
        med = float(np.median(llk))
        w = np.empty(C)
        b, c = ctypes.c_double(), ctypes.c_double()
        _lib.call("altar_cov_dbeta_brent", _lib.ptr(llk), C, med, beta0, 0.01, 1.0, 1000, _lib.ptr(w),
                  ctypes.byref(b), ctypes.byref(c))
        rb, rc, rw = oracle.dbeta_brent(llk, beta0)
        assert b.value == rb and c.value == rc and np.array_equal(w, rw), (C, scale, beta0)
