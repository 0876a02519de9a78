"""Oracle sample_k (src/clifft/svm/svm.cc:347-622): exactly k forced faults
per shot, the fault set drawn from the conditional Poisson-binomial law
(DP walk) or uniformly (partial Fisher-Yates) when all site probabilities
are equal.

Pins: (i) every shot has exactly k faults; (ii) the empirical law of the
fault set matches P(S) = prod_{i in S} w_i / e_k(w), w_i = p_i / (1 - p_i),
computed here by brute force over all k-subsets (chi-square, p > 0.01);
(iii) uniform sets are uniform; (iv) readout entries take part; (v) the
reference's errors for infeasible strata; (vi) shot 0 of the per-shot
convention equals the reference stream's first shot."""

from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy import stats

from clifft_amd.program import Op, build_program, make_meas, make_noise, make_noise_block, make_readout_noise


def fault_program(probs, readout=()):
    """Site i is X_ERROR(probs[i]) on dormant qubit i, measured at the end
    (record bit i = site fired); readout entries flip extra measurements of
    qubits that carry no noise."""
    n = len(probs)
    nr = len(readout)
    bc = [make_noise(i) for i in range(n)]
    bc += [make_meas(Op.MEAS_DORMANT_STATIC, q, q) for q in range(n + nr)]
    bc += [make_readout_noise(e) for e in range(nr)]
    sites = [[(1 << i, 0, p)] for i, p in enumerate(probs)]
    return build_program(bc, peak_rank=0, num_qubits=n + nr, num_measurements=n + nr, noise_sites=sites,
                         readout_noise=[(n + e, p) for e, p in enumerate(readout)])


def subset_law(probs, k):
    w = np.array([p / (1 - p) for p in probs])
    subsets = list(itertools.combinations(range(len(probs)), k))
    wt = np.array([np.prod(w[list(s)]) for s in subsets])
    return subsets, wt / wt.sum()


def chi2_subsets(rec, probs, k):
    subsets, law = subset_law(probs, k)
    index = {s: i for i, s in enumerate(subsets)}
    counts = np.zeros(len(subsets))
    for row in rec:
        counts[index[tuple(np.flatnonzero(row))]] += 1
    return stats.chisquare(counts, law * counts.sum()).pvalue


@pytest.mark.parametrize("k", [1, 2, 3])
def test_exactly_k_and_poisson_binomial_law(orc, k):
    probs = [0.01, 0.05, 0.2, 0.02, 0.1]
    p = fault_program(probs)
    m, *_ = orc.sample_k(p, 20000, k, seed=11 + k)
    assert np.all(m.sum(axis=1) == k)
    assert chi2_subsets(m, probs, k) > 0.01


def test_uniform_mode_law(orc):
    probs = [0.03] * 6
    p = fault_program(probs)
    for per_shot in (True, False):
        m, *_ = orc.sample_k(p, 15000, 2, seed=5, per_shot=per_shot)
        assert np.all(m.sum(axis=1) == 2)
        assert chi2_subsets(m, probs, 2) > 0.01


def test_readout_sites_take_part(orc):
    probs, ro = [0.02, 0.1], [0.05, 0.3]
    p = fault_program(probs, readout=ro)
    m, *_ = orc.sample_k(p, 20000, 2, seed=9)
    assert np.all(m.sum(axis=1) == 2)
    assert chi2_subsets(m, probs + ro, 2) > 0.01


def test_infeasible_strata_raise(orc):
    p = fault_program([0.1, 0.0, 0.2])
    with pytest.raises(ValueError, match="exceeds"):
        orc.sample_k(p, 1, 4)
    with pytest.raises(ValueError, match="zero probability mass"):
        orc.sample_k(p, 1, 3)


def test_shot0_matches_reference_stream(orc):
    p = fault_program([0.01, 0.05, 0.2, 0.02, 0.1])
    a = orc.sample_k(p, 1, 2, seed=77, per_shot=True)
    b = orc.sample_k(p, 1, 2, seed=77, per_shot=False)
    assert np.array_equal(a[0], b[0])


def test_noise_block_equals_noise_sequence(orc):
    probs = [0.01, 0.05, 0.2, 0.02]
    p = fault_program(probs)
    blk = build_program([make_noise_block(0, 4)] + [make_meas(Op.MEAS_DORMANT_STATIC, i, i) for i in range(4)],
                        peak_rank=0, num_qubits=4, num_measurements=4,
                        noise_sites=[[(1 << i, 0, pr)] for i, pr in enumerate(probs)])
    a = orc.sample_k(p, 3000, 2, seed=4)
    b = orc.sample_k(blk, 3000, 2, seed=4)
    assert np.array_equal(a[0], b[0])
