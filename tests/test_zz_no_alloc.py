"""The allocation-free resampling entry point (altar_cov_resample_dev_noalloc,
scratch in a fixed device array) against altar_cov_resample_dev (scratch
from the caller) through the C ABI: multiplicities and resampling maps
identical, bit for bit, on two calls in a row, the second with fewer chains
(the array keeps the first call's contents), and both equal to the oracle's
multiplicities on the same Philox uniforms.  A call beyond the array's
capacity is refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from altar_amd import _lib

pytestmark = pytest.mark.gpu

MAX_CHAINS = 262144     # ALTAR_NOALLOC_MAX_CHAINS


def resample_both(cuda, w_host, seed, stage):
    C = len(w_host)
    w = torch.from_numpy(w_host).to(cuda)
    edges = torch.empty(C + 1, dtype=torch.float64, device=cuda)
    offset = torch.empty(C, dtype=torch.int32, device=cuda)
    count_a = torch.full((C,), -7, dtype=torch.int32, device=cuda)
    src_a = torch.full((C,), -7, dtype=torch.int32, device=cuda)
    _lib.call("altar_cov_resample_dev", _lib.ptr(w), C, seed, stage, _lib.ptr(edges), _lib.ptr(count_a),
              _lib.ptr(offset), _lib.ptr(src_a), None)
    count_b = torch.full((C,), -9, dtype=torch.int32, device=cuda)
    src_b = torch.full((C,), -9, dtype=torch.int32, device=cuda)
    _lib.call("altar_cov_resample_dev_noalloc", _lib.ptr(w), C, seed, stage, _lib.ptr(count_b), _lib.ptr(src_b),
              None)
    torch.cuda.synchronize()
    return count_a.cpu().numpy(), src_a.cpu().numpy(), count_b.cpu().numpy(), src_b.cpu().numpy()


def test_noalloc_resample_matches_existing_entry_point(cuda, oracle):
    rng = np.random.default_rng(71)
    for C, seed, stage in ((200_000, 77, 5), (3_001, 78, 6)):
        print(f"case: allocation-free resampling C={C} seed={seed} stage={stage}", flush=True)
        w = rng.gamma(0.3, 1.0, C)
        w /= w.sum()
        ca, sa, cb, sb = resample_both(cuda, w, seed, stage)
        assert np.array_equal(ca, cb) and np.array_equal(sa, sb)
        want = oracle.multiplicities(w, oracle.uniforms(seed, stage, 0, 2, C))
        assert np.array_equal(cb, want)
        assert cb.sum() == C and np.array_equal(sb, np.repeat(np.arange(C), cb))
    # capacity: refused on the host, nothing launched
    big = torch.empty(1, dtype=torch.float64, device=cuda)
    rc = _lib.lib().altar_cov_resample_dev_noalloc(_lib.ptr(big), MAX_CHAINS + 1, 1, 1, _lib.ptr(big),
                                                   _lib.ptr(big), None)
    assert rc == _lib.EINVAL
    assert b"exceeds the scratch capacity" in _lib.lib().altar_last_error()
