"""Pins for the oracle branches the dense-model checks leave open.

* FLAG_IDENTITY measurements (src/clifft/svm/svm_kernels.inl:2361-2369 and
  the forced shortcut at :2426-2432): the reference's known-answer tests
  tests/test_svm.cc:2026-2150 and tests/test_execute_forced.cc:81-106,
  restated, plus the defining property that the record is the sign bit
  whatever the Pauli frame holds.
* PRNG draw order of NOISE / READOUT_NOISE / MEAS_DORMANT_RANDOM in one shot
  of the reference's single stream (svm.h:231-248, svm_kernels.inl:1886-1959,
  1551-1566), re-derived in pure Python from the published xoshiro256++ and
  SplitMix64 and compared record for record: a skipped or extra draw, a
  channel chosen by u instead of u * sum(p), or a flipped readout comparison
  changes the records.
* READOUT_NOISE sampling mode: 5-sigma binomial firing rates
  (tests/python/conftest.py:38-77) for entries whose measurement indices are
  a permutation of their entry indices, and two entries on one record.
* Channel choice on a site with unequal channel probabilities: chi-square
  of the (no fault, channel 0, 1, 2) counts against (1 - sum p, p0, p1, p2),
  with the same counts rejected against the equiprobable law.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats

from clifft_amd.program import (FLAG_IDENTITY, Op, build_program, make_detector, make_meas, make_noise,
                                make_observable, make_postselect, make_readout_noise)
from test_oracle_kats import splitmix64_stream, xoshiro_next


def identity_meas(sign, forced=False):
    op = Op.MEAS_DORMANT_STATIC_FORCED if forced else Op.MEAS_DORMANT_STATIC
    return make_meas(op, 0, 0, sign, flags=FLAG_IDENTITY)


# ----------------------------------------------------------------------------
# FLAG_IDENTITY (tests/test_svm.cc:2026-2150)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("sign,meas,det", [(True, 1, 0), (False, 0, 1)])
def test_identity_detector_expected_one(orc, sign, meas, det):
    p = build_program([identity_meas(sign), make_detector(0, 0, expected_one=True)], peak_rank=0, num_qubits=1,
                      num_measurements=1, num_detectors=1, detector_targets=[[0]])
    m, d, *_ = orc.sample(p, 1, seed=0, per_shot=False)
    assert m[0, 0] == meas and d[0, 0] == det


@pytest.mark.parametrize("sign,passed", [(True, 10), (False, 0)])
def test_identity_postselect_expected_one(orc, sign, passed):
    p = build_program([identity_meas(sign), make_postselect(0, 0, expected_one=True)], peak_rank=0,
                      num_qubits=1, num_measurements=1, num_detectors=1, detector_targets=[[0]])
    assert orc.sample_survivors(p, 10, seed=0)["passed"] == passed


def test_identity_observable_normalisation(orc):
    p = build_program([identity_meas(True), make_observable(0, 0)], peak_rank=0, num_qubits=1,
                      num_measurements=1, num_observables=1, observable_targets=[[0]], expected_observables=[1])
    _, _, o, *_ = orc.sample(p, 1, seed=0, per_shot=False)
    assert o[0, 0] == 0
    surv = orc.sample_survivors(p, 10, seed=0)
    assert surv["passed"] == 10 and surv["logical_errors"] == 0


@pytest.mark.parametrize("sign", [False, True])
@pytest.mark.parametrize("px", [0, 1])
def test_identity_ignores_the_frame(orc, sign, px):
    """The record is the sign bit; a plain static measurement would give
    p_x ^ sign (svm_kernels.inl:1540-1545)."""
    prog = build_program([identity_meas(sign)], peak_rank=1, num_qubits=1, num_measurements=1)
    st = orc.State(1, 1, num_qubits=1)
    st.p_x[0] = px
    orc.execute(prog, st)
    assert int(st.meas_record[0]) == int(sign)
    plain = build_program([make_meas(Op.MEAS_DORMANT_STATIC, 0, 0, sign)], peak_rank=1, num_qubits=1,
                          num_measurements=1)
    st = orc.State(1, 1, num_qubits=1)
    st.p_x[0] = px
    orc.execute(plain, st)
    assert int(st.meas_record[0]) == px ^ int(sign)


@pytest.mark.parametrize("px", [0, 1])
def test_identity_forced(orc, px):
    """tests/test_execute_forced.cc:81-106: identity + SIGN forced to 1 is
    reachable with log-probability 0; forced to 0 it is unreachable, for
    either frame bit."""
    prog = build_program([identity_meas(True, forced=True)], peak_rank=2, num_qubits=2, num_measurements=1)
    for rec, reachable in ((1, True), (0, False)):
        st = orc.State(2, 1, num_qubits=2)
        st.p_x[0] = px
        st.set_forced_record(np.array([rec], np.uint8))
        orc.execute(prog, st)
        assert st.forced_reachable == reachable
        if reachable:
            assert st.forced_log_probability == 0.0 and int(st.meas_record[0]) == 1


# ----------------------------------------------------------------------------
# PRNG draw order over noise, readout noise and a random measurement
# ----------------------------------------------------------------------------
CHANNELS = [(0b01, 0, 0.3), (0b10, 0, 0.1)]  # X on q0 (p 0.3), X on q1 (p 0.1)
RO_P = 0.4


def draw_order_program():
    bc = [make_noise(0),
          make_meas(Op.MEAS_DORMANT_STATIC, 0, 0), make_meas(Op.MEAS_DORMANT_STATIC, 1, 1),
          make_meas(Op.MEAS_DORMANT_STATIC, 2, 2), make_readout_noise(0),
          make_meas(Op.MEAS_DORMANT_RANDOM, 3, 3)]
    return build_program(bc, peak_rank=1, num_qubits=4, num_measurements=4, noise_sites=[CHANNELS],
                         readout_noise=[(2, RO_P)])


def python_shot(seed):
    """One shot of the reference's single stream, restated step by step."""
    s = splitmix64_stream(seed, 4)
    u = lambda: (xoshiro_next(s) >> 11) * 2.0 ** -53
    psum = 0.0
    for _, _, p in CHANNELS:
        psum += p
    hazard = -math.log1p(-psum)
    rec = [0, 0, 0, 0]
    gap = -math.log(1.0 - u())  # svm.cc:224-225: draw_next_noise at shot start
    if hazard > gap:  # upper_bound(hazards, gap) == 0: site 0 fires
        r = u() * psum
        cum = 0.0
        for x, _, p in CHANNELS:
            cum += p
            if r < cum:
                rec[0] ^= x & 1
                rec[1] ^= (x >> 1) & 1
                break
        # next_noise_idx = 1 = number of sites: no further draw (svm.h:238)
    if u() < RO_P:  # READOUT_NOISE draws once per entry (svm_kernels.inl:1953)
        rec[2] ^= 1
    rec[3] = 0 if u() < 0.5 else 1  # MEAS_DORMANT_RANDOM
    return rec


def test_draw_order_single_stream(orc):
    prog = draw_order_program()
    seen = set()
    for seed in range(400):
        m, *_ = orc.sample(prog, 1, seed=seed, per_shot=False)
        exp = python_shot(seed)
        assert list(map(int, m[0])) == exp, seed
        seen.add(tuple(exp))
    # the seeds cover every outcome the two channels and the readout allow
    assert {(1, 0), (0, 1), (0, 0)} <= {e[:2] for e in seen}
    assert {0, 1} <= {e[2] for e in seen} and {0, 1} <= {e[3] for e in seen}


# ----------------------------------------------------------------------------
# READOUT_NOISE firing rates (sampling mode)
# ----------------------------------------------------------------------------
# (meas_idx, p): measurement indices permuted against entry order, and
# record 4 carries two entries (flip probability p1 (1 - p2) + p2 (1 - p1)).
RO_ENTRIES = [(3, 0.05), (0, 0.3), (2, 0.5), (1, 0.9), (4, 0.2), (4, 0.35)]


def readout_program():
    bc = [make_meas(Op.MEAS_DORMANT_STATIC, q, q, sign=(q == 1)) for q in range(5)]
    bc += [make_readout_noise(e) for e in range(len(RO_ENTRIES))]
    return build_program(bc, peak_rank=1, num_qubits=5, num_measurements=5, readout_noise=RO_ENTRIES)


def readout_flip_rates():
    rate = {}
    for mi, p in RO_ENTRIES:
        q = rate.get(mi, 0.0)
        rate[mi] = q * (1 - p) + p * (1 - q)
    return rate


def test_readout_noise_rates_binomial(orc):
    shots = 40000
    m, *_ = orc.sample(readout_program(), shots, seed=17, per_shot=True)
    base = np.array([0, 1, 0, 0, 0])  # sign bits of the static measurements
    flips = m ^ base
    for mi, p in readout_flip_rates().items():
        tol = 5 * math.sqrt(p * (1 - p) / shots)
        assert abs(flips[:, mi].mean() - p) < tol, (mi, flips[:, mi].mean(), p)


# ----------------------------------------------------------------------------
# Channel choice with unequal channel probabilities
# ----------------------------------------------------------------------------
UNEQUAL = [(0b01, 0, 0.05), (0b10, 0, 0.15), (0b11, 0, 0.3)]


def channel_program():
    bc = [make_noise(0), make_meas(Op.MEAS_DORMANT_STATIC, 0, 0), make_meas(Op.MEAS_DORMANT_STATIC, 1, 1)]
    return build_program(bc, peak_rank=1, num_qubits=2, num_measurements=2, noise_sites=[UNEQUAL])


def channel_counts(m):
    code = m[:, 0] + 2 * m[:, 1]  # 0: silent, 1: channel 0, 2: channel 1, 3: channel 2
    return np.bincount(code, minlength=4)


def test_unequal_channel_choice_chi_square(orc):
    shots = 60000
    m, *_ = orc.sample(channel_program(), shots, seed=23, per_shot=True)
    counts = channel_counts(m)
    psum = sum(c[2] for c in UNEQUAL)
    law = np.array([1 - psum] + [c[2] for c in UNEQUAL])
    assert stats.chisquare(counts, law * shots).pvalue > 0.01
    # power: the same counts reject channels chosen with equal probability
    eq = np.array([1 - psum] + [psum / 3] * 3)
    assert stats.chisquare(counts, eq * shots).pvalue < 1e-6
