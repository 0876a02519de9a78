"""clifft_svm_sample_noalloc against clifft_svm_sample, through the C ABI.

The no-allocation entry point keeps global-memory amplitudes in the
library's fixed device array; its records and expectation values must
equal the allocating entry point's, call after call (the array keeps the
last call's contents), with device outputs only."""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from randprog import random_program

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    from clifft_amd import _lib, svm
    assert _lib.backend() == "hip:gfx950", _lib.backend()
    return torch, _lib, svm


def run(env, fn_name, prog, shots, seed):
    torch, L, svm = env
    dp = svm.device_program(prog)
    dev = torch.device("cuda:0")
    m = torch.full((shots, prog.num_measurements), 7, dtype=torch.uint8, device=dev)
    d = torch.full((shots, prog.num_detectors), 7, dtype=torch.uint8, device=dev)
    o = torch.full((shots, prog.num_observables), 7, dtype=torch.uint8, device=dev)
    e = torch.full((shots, prog.num_exp_vals), 7.0, dtype=torch.float64, device=dev)
    ptr = lambda t: t.data_ptr() if t.numel() else None
    rec = L.Records(ptr(m), ptr(d), ptr(o), ptr(e), None)
    torch.cuda.synchronize()
    rc = getattr(L.lib(), fn_name)(dp.handle, shots, seed, L.REC_DEVICE, ctypes.byref(rec), None)
    L.check(rc, fn_name)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (m, d, o, e)]


# (num_qubits, peak, shots, seed): one shot per lane, LDS teams,
# global-memory teams (fused runs use a second slice), and the largest
# accepted rank (one 64 MiB slice); seeds whose programs reach the peak
CASES = [(10, 4, 700, 7004), (16, 9, 400, 7009), (20, 14, 300, 7014), (24, 16, 120, 7016), (28, 22, 3, 7028)]


@pytest.mark.parametrize("nq,peak,shots,seed", CASES)
def test_noalloc_matches_sample(env, nq, peak, shots, seed):
    announce("noalloc", nq, peak, shots, seed)
    p = random_program(seed, num_qubits=nq, peak=peak, length=200 + 10 * peak, p_array=0.5)
    assert p.active_k_history().max() == peak
    for call, (n, sd) in enumerate([(shots, 11), (max(1, shots // 3), 12)]):  # second call smaller
        ref = run(env, "clifft_svm_sample", p, n, sd)
        got = run(env, "clifft_svm_sample_noalloc", p, n, sd)
        for j in range(3):
            assert np.array_equal(got[j], ref[j]), (call, j)
        np.testing.assert_allclose(got[3], ref[3], atol=1e-12, rtol=0)


def test_noalloc_refusals(env):
    torch, L, svm = env
    from randprog import random_program as rp
    big = rp(7100, num_qubits=30, peak=23, length=60)
    assert big.peak_rank == 23
    dp = svm.device_program(big)
    rec = L.Records(None, None, None, None, None)
    rc = L.lib().clifft_svm_sample_noalloc(dp.handle, 2, 1, L.REC_DEVICE, ctypes.byref(rec), None)
    assert rc == L.E_UNSUPPORTED and b"peak_rank <= 22" in L.lib().clifft_svm_last_error()
    small = rp(7101, num_qubits=8, peak=3, length=60)
    dp2 = svm.device_program(small)
    rc = L.lib().clifft_svm_sample_noalloc(dp2.handle, 2, 1, 0, ctypes.byref(rec), None)
    assert rc == L.E_INVALID and b"CLIFFT_REC_DEVICE" in L.lib().clifft_svm_last_error()
