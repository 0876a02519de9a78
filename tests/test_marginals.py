"""Sampled record distributions against exact ones (the distribution
clause of the parity bar: chi-square, p > 0.01 at fixed seeds).

Exact law: for a noiseless program with m <= 8 visible measurements the
probability of every one of the 2^m records comes from forced execution
(record_probabilities).  CPU tests check the oracle's sampler; GPU tests
check the device sampler (one-sample chi-square against the exact law, and a
two-sample test against the oracle for noisy programs, whose law has no
closed form here)."""

from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy import stats

from randprog import random_measure_only_program, random_program


def exact_law(orc, p):
    m = p.num_measurements
    recs = np.array(list(itertools.product([0, 1], repeat=m)), np.uint8)[:, ::-1]
    lp = orc.record_log_probs(p.forced(), recs)
    prob = np.where(lp < -1e300, 0.0, np.exp(lp))
    assert abs(prob.sum() - 1.0) < 1e-9
    return recs, prob


def histogram(records, m):
    idx = records.astype(np.int64) @ (1 << np.arange(m))
    return np.bincount(idx, minlength=1 << m)


def chi2_fit(counts, prob):
    keep = prob > 0
    assert counts[~keep].sum() == 0, "sampled a record of probability 0"
    exp = prob[keep] * counts.sum()
    # merge bins with small expectation into one
    small = exp < 5
    obs = np.append(counts[keep][~small], counts[keep][small].sum())
    ex = np.append(exp[~small], exp[small].sum())
    if ex[-1] == 0:
        obs, ex = obs[:-1], ex[:-1]
    return stats.chisquare(obs, ex).pvalue


def measure_programs():
    out = []
    for seed in range(40):
        p = random_measure_only_program(seed, num_qubits=8, peak=4, length=100)
        if 2 <= p.num_measurements <= 8 and p.total_meas_slots == p.num_measurements:
            out.append((seed, p))
        if len(out) == 4:
            break
    return out


@pytest.mark.parametrize("case", range(4))
def test_oracle_sampler_matches_exact_law(orc, case):
    seed, p = measure_programs()[case]
    recs, prob = exact_law(orc, p)
    m, *_ = orc.sample(p, 40000, seed=100 + seed)
    assert chi2_fit(histogram(m, p.num_measurements), prob) > 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(4))
def test_device_sampler_matches_exact_law(orc, case):
    from clifft_amd import svm
    seed, p = measure_programs()[case]
    print("CASE marginals exact", case, seed, flush=True)
    recs, prob = exact_law(orc, p)
    res = svm.sample(p, 200000, seed=200 + seed)
    assert chi2_fit(histogram(res.measurements, p.num_measurements), prob) > 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("peak", [2, 5, 9])
def test_device_noisy_marginals_match_oracle(orc, peak):
    """Per-bit marginals of a noisy program: device vs oracle samples with
    different seeds (two-sample chi-square on each measured bit)."""
    from clifft_amd import svm
    print("CASE marginals noisy", peak, flush=True)
    p = random_program(600 + peak, num_qubits=12, peak=peak, length=220)
    n = 20000
    g = svm.sample(p, n, seed=5).measurements
    o, *_ = orc.sample(p, n, seed=6)
    pvals = []
    for j in range(p.num_measurements):
        a, b = g[:, j].sum(), o[:, j].sum()
        if a + b in (0, 2 * n):
            continue
        pvals.append(stats.chi2_contingency([[a, n - a], [b, n - b]])[1])
    # Bonferroni over the bits tested
    assert not pvals or min(pvals) > 0.01 / len(pvals)
