"""Device k-fault importance sampling vs the oracle's sample_k with the
per-shot stream convention: records bit-exact, in both fault-set samplers
(DP walk and uniform Fisher-Yates) and every team tier."""

from __future__ import annotations

import numpy as np
import pytest

from randprog import random_program
from test_oracle_sample_k import fault_program

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def svm():
    import clifft_amd.svm as svm
    from clifft_amd import _lib
    assert _lib.backend() == "hip:gfx950", _lib.backend()
    return svm


def same(res, ref):
    m, d, o, e, _ = ref
    assert np.array_equal(res.measurements, m)
    assert np.array_equal(res.detectors, d)
    assert np.array_equal(res.observables, o)
    np.testing.assert_allclose(res.exp_vals, e, atol=1e-12, rtol=0)


@pytest.mark.parametrize("nq,peak", [(8, 3), (12, 6), (16, 9), (20, 14)])
@pytest.mark.parametrize("k", [1, 3])
def test_sample_k_random_programs(svm, orc, nq, peak, k):
    print("CASE sample_k", nq, peak, k, flush=True)
    p = random_program(300 + peak, num_qubits=nq, peak=peak, length=200, readout=True)
    n_sites = len(orc.noise_site_probabilities(p))
    if n_sites < k:
        pytest.skip("too few fault sites drawn")
    shots = 64 if peak >= 12 else 300
    same(svm.sample_k(p, shots, k, seed=17), orc.sample_k(p, shots, k, seed=17))


@pytest.mark.parametrize("k", [1, 2, 4])
def test_sample_k_surface_code(svm, orc, k):
    """The surface-code workload (its DEPOLARIZE2 channel sums differ from
    p = 1e-3 in the last bit, so the DP sampler runs)."""
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as cc
    print("CASE sample_k surface", k, flush=True)
    p = cc(W.surface_code_d3())
    same(svm.sample_k(p, 2000, k, seed=3), orc.sample_k(p, 2000, k, seed=3))


@pytest.mark.parametrize("k", [1, 3, 6])
def test_sample_k_uniform_sampler(svm, orc, k):
    """Equal probabilities on every site: the partial Fisher-Yates sampler."""
    print("CASE sample_k uniform", k, flush=True)
    p = fault_program([0.03] * 7, readout=[0.03] * 3)
    assert np.all(orc.noise_site_probabilities(p) == 0.03)
    res = svm.sample_k(p, 3000, k, seed=k + 1)
    assert np.all(res.measurements.sum(axis=1) == k)
    same(res, orc.sample_k(p, 3000, k, seed=k + 1))


def test_sample_k_fault_sets(svm, orc):
    probs = [0.01, 0.05, 0.2, 0.02, 0.1]
    p = fault_program(probs, readout=[0.03, 0.3])
    for k in (1, 2, 5):
        print("CASE sample_k faults", k, flush=True)
        res = svm.sample_k(p, 4000, k, seed=k)
        assert np.all(res.measurements.sum(axis=1) == k)
        same(res, orc.sample_k(p, 4000, k, seed=k))


def test_sample_k_survivors_and_errors(svm, orc):
    print("CASE sample_k survivors", flush=True)
    p = random_program(1, num_qubits=12, peak=5, length=250, postselect=True, readout=True)
    assert p.has_postselection
    g = svm.sample_k_survivors(p, 500, 2, seed=4, keep_records=True)
    m, d, o, e, disc = orc.sample_k(p, 500, 2, seed=4)
    keep = disc == 0
    assert g.passed_shots == int(keep.sum())
    assert np.array_equal(g.measurements, m[keep])
    assert np.array_equal(g.detectors, d[keep])
    n = len(orc.noise_site_probabilities(p))
    with pytest.raises(ValueError, match="exceeds"):
        svm.sample_k(p, 10, n + 1, seed=1)
    np.testing.assert_array_equal(svm.noise_site_probabilities(p), orc.noise_site_probabilities(p))
