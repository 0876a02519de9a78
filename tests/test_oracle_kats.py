"""Pins the CPU oracle (oracle/svm_oracle.c) before it is trusted as the
parity checker.

* PRNG: published SplitMix64 / xoshiro256++ values and a pure-Python
  re-derivation of the published algorithms.
* Frame opcodes: the reference's known-answer tests
  (tests/test_svm.cc:93-272, 839-1000, 981-1000) restated.
* Array opcodes, EXPAND*, measurements, APPLY_PAULI, EXP_VAL: against a dense
  2^n statevector model (tests/dense.py) with textbook gate matrices, random
  states and random Pauli frames, including the global phase.
* Noise: NOISE_BLOCK == NOISE sequence (tests/test_bytecode_passes.cc:136-203)
  and binomial 5-sigma firing rates (tests/python/conftest.py:38-77).
* Drivers: expected_observables XOR, POSTSELECT discards, survivor counts
  (tests/test_svm.cc:1613-1915, 2026-2150), record probabilities summing to 1.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import dense as D
from clifft_amd.program import (Op, build_program, make_apply_pauli, make_array_cnot, make_array_cz,
                                make_array_h, make_array_multi_cnot, make_array_multi_cz, make_array_rot,
                                make_array_s, make_array_s_dag, make_array_swap, make_array_t,
                                make_array_t_dag, make_detector, make_exp_val, make_expand, make_expand_rot,
                                make_expand_t, make_expand_t_dag, make_frame_cnot, make_frame_cz, make_frame_h,
                                make_frame_s, make_frame_s_dag, make_frame_swap, make_meas, make_noise,
                                make_noise_block, make_observable, make_postselect, make_swap_meas_interfere)

M64 = (1 << 64) - 1


# ----------------------------------------------------------------------------
# PRNG
# ----------------------------------------------------------------------------
def splitmix64_stream(seed, n):
    z, out = seed & M64, []
    for _ in range(n):
        z = (z + 0x9E3779B97F4A7C15) & M64
        x = z
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
        out.append(x ^ (x >> 31))
    return out


def xoshiro_next(s):
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M64
    result = (rotl((s[0] + s[3]) & M64, 23) + s[0]) & M64
    t = (s[1] << 17) & M64
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = rotl(s[3], 45)
    return result


def test_splitmix64_published_value():
    # splitmix64.c with x = 0: first output is 0xe220a8397b1dcdaf
    assert splitmix64_stream(0, 1)[0] == 0xE220A8397B1DCDAF


def test_xoshiro_closed_form_first_output():
    # state {1,2,3,4}: rotl(1 + 4, 23) + 1
    assert xoshiro_next([1, 2, 3, 4]) == (5 << 23) + 1


def test_oracle_rng_matches_python(orc):
    for seed in (0, 42, 2**63 + 5):
        st = orc.State(1, 0, seed=seed)
        s = splitmix64_stream(seed, 4)
        assert list(map(int, st.rng)) == s
        for _ in range(200):
            expect = (xoshiro_next(s) >> 11) * 2.0**-53
            assert st.random_double() == expect


def test_per_shot_stream_convention(orc):
    seed = 42
    st = orc.State(1, 0, seed=0)
    for shot in (0, 1, 7, 1000):
        st.seed_shot(seed, shot)
        assert list(map(int, st.rng)) == splitmix64_stream(seed, 4 * shot + 4)[4 * shot:]
    st.seed_shot(seed, 0)
    ref = orc.State(1, 0, seed=seed)
    assert list(map(int, st.rng)) == list(map(int, ref.rng))


# ----------------------------------------------------------------------------
# Frame KATs (tests/test_svm.cc:93-272)
# ----------------------------------------------------------------------------
def run1(orc, instr, peak=2, px=0, pz=0, gamma=1.0, n=2):
    st = orc.State(peak, 0, num_qubits=n)
    st.p_x[0], st.p_z[0] = px, pz
    st.gamma = gamma
    orc.execute(build_program([instr], peak_rank=peak, num_qubits=n), st)
    return st


@pytest.mark.parametrize("instr,px,pz,epx,epz,egamma", [
    (make_frame_cnot(0, 1), 0b01, 0b00, 0b11, 0b00, 1),   # X spreads forward
    (make_frame_cnot(0, 1), 0b00, 0b10, 0b00, 0b11, 1),   # Z spreads backward
    (make_frame_cz(0, 1), 0b11, 0b00, 0b11, 0b11, -1),    # both X: gamma *= -1
    (make_frame_cz(0, 1), 0b01, 0b00, 0b01, 0b10, 1),     # one X: no phase
    (make_frame_h(0), 0b01, 0b00, 0b00, 0b01, 1),         # H swaps
    (make_frame_h(0), 0b01, 0b01, 0b01, 0b01, -1),        # H on Y: -1
    (make_frame_s(0), 0b01, 0b00, 0b01, 0b01, 1j),        # S on X: *i
    (make_frame_s(0), 0b00, 0b01, 0b00, 0b01, 1),         # S on Z: nothing
    (make_frame_s_dag(0), 0b01, 0b00, 0b01, 0b01, -1j),   # S_dag on X: *-i
    (make_frame_s_dag(0), 0b00, 0b00, 0b00, 0b00, 1),
    (make_frame_swap(0, 1), 0b01, 0b10, 0b10, 0b01, 1),
    (make_frame_swap(0, 1), 0b11, 0b00, 0b11, 0b00, 1),
])
def test_frame_kats(orc, instr, px, pz, epx, epz, egamma):
    st = run1(orc, instr, px=px, pz=pz)
    assert int(st.p_x[0]) == epx and int(st.p_z[0]) == epz
    assert abs(st.gamma - egamma) < 1e-15


def test_apply_pauli_kats(orc):
    # tests/test_svm.cc:839-980: Z error on X frame negates gamma; sign negates
    for (ex, ez, sgn, px0, pz0, epx, epz, eg) in [
        (0b1, 0, False, 0, 0, 1, 0, 1),
        (0, 0b1, False, 0, 0, 0, 1, 1),
        (0b1, 0, False, 0, 1, 1, 1, 1),
        (0, 0b1, False, 1, 0, 1, 1, -1),
        (0b1, 0, True, 0, 0, 1, 0, -1),
    ]:
        p = build_program([make_meas(Op.MEAS_DORMANT_STATIC, 1, 0, True), make_apply_pauli(0, 0)], peak_rank=1,
                          num_qubits=2, num_measurements=1, pauli_masks=[(ex, ez, sgn)])
        st = orc.State(1, 1, num_qubits=2)
        st.p_x[0], st.p_z[0] = px0, pz0
        orc.execute(p, st)
        assert int(st.meas_record[0]) == 1
        assert (int(st.p_x[0]), int(st.p_z[0])) == (epx, epz)
        assert abs(st.gamma - eg) < 1e-15


def test_two_t_equal_s_and_t_tdag_cancel(orc):
    # tests/test_svm.cc:463-496
    p = build_program([make_expand(0), make_array_t(0), make_array_t(0)], peak_rank=1)
    st = orc.State(1, 0)
    orc.execute(p, st)
    assert abs(st.v[1] - 1j) < 1e-14 and abs(st.v[0] - 1) < 1e-14
    p = build_program([make_expand(0), make_array_t(0), make_array_t_dag(0)], peak_rank=1)
    st = orc.State(1, 0)
    orc.execute(p, st)
    assert abs(st.v[1] - 1) < 1e-14


# ----------------------------------------------------------------------------
# Dense checks of the unitary opcodes
# ----------------------------------------------------------------------------
N = 5


def prepared(orc, rng, k, peak=None, nq=N):
    peak = peak if peak is not None else nq
    st = orc.State(peak, 4, num_qubits=nq)
    phi = D.random_state(rng, k)
    st.v[:] = 0
    st.v[: 1 << k] = phi
    st.active_k = k
    st.p_x[0] = int(rng.integers(0, 1 << nq))
    st.p_z[0] = int(rng.integers(0, 1 << nq))
    st.gamma = np.exp(1j * rng.uniform(0, 2 * np.pi))
    return st


def unitary_cases(rng, k):
    a, b = (int(x) for x in rng.choice(k, 2, replace=False))
    v = int(rng.integers(0, k))
    z = np.exp(1j * rng.uniform(0, 2 * np.pi))
    mask_c = sum(1 << q for q in range(k) if q != v and rng.random() < 0.6)
    rot = np.diag([1, z])
    cases = [
        (make_array_cnot(a, b), lambda s: D.cnot(s, a, b, N)),
        (make_array_cz(a, b), lambda s: D.cz(s, a, b, N)),
        (make_array_swap(a, b), lambda s: D.swap(s, a, b, N)),
        (make_array_h(v), lambda s: D.apply_1q(s, D.H, v, N)),
        (make_array_s(v), lambda s: D.apply_1q(s, D.S, v, N)),
        (make_array_s_dag(v), lambda s: D.apply_1q(s, D.SDG, v, N)),
        (make_array_t(v), lambda s: D.apply_1q(s, D.T, v, N)),
        (make_array_t_dag(v), lambda s: D.apply_1q(s, D.TDG, v, N)),
        (make_array_rot(v, z.real, z.imag), lambda s: D.apply_1q(s, rot, v, N)),
    ]

    def multi_cnot(s):
        for c in range(k):
            if (mask_c >> c) & 1:
                s = D.cnot(s, c, v, N)
        return s

    def multi_cz(s):
        for t in range(k):
            if (mask_c >> t) & 1:
                s = D.cz(s, v, t, N)
        return s

    cases.append((make_array_multi_cnot(v, mask_c), multi_cnot))
    cases.append((make_array_multi_cz(v, mask_c), multi_cz))
    # dormant-axis T / ROT act on the |0> or X-flipped dormant qubit
    d = int(rng.integers(k, N))
    cases.append((make_array_t(d), lambda s: D.apply_1q(s, D.T, d, N)))
    cases.append((make_array_rot(d, z.real, z.imag), lambda s: D.apply_1q(s, rot, d, N)))
    return cases


@pytest.mark.parametrize("seed", range(6))
def test_unitary_opcodes_dense(orc, seed):
    rng = np.random.default_rng(seed)
    k = 3
    for case_i in range(len(unitary_cases(np.random.default_rng(seed), k))):
        rng_c = np.random.default_rng(seed)
        instr, U = unitary_cases(rng_c, k)[case_i]
        st = prepared(orc, rng, k)
        before = D.state_of(st, N)
        orc.execute(build_program([instr], peak_rank=N, num_qubits=N), st)
        after = D.state_of(st, N)
        np.testing.assert_allclose(after, U(before), atol=1e-13, err_msg=str(instr))


@pytest.mark.parametrize("seed", range(4))
def test_frame_opcodes_dense_on_dormant(orc, seed):
    """Frame opcodes act on axes whose array factor is |0>: the physical gate
    then only conjugates P (src/clifft/backend/backend.cc:302-350)."""
    rng = np.random.default_rng(seed)
    k = 2
    c_d = int(rng.integers(k, N))
    t_any = int(rng.choice([q for q in range(N) if q != c_d]))
    d1, d2 = (int(x) for x in rng.choice(range(k, N), 2, replace=False))
    for instr, U in [
        (make_frame_cnot(c_d, t_any), lambda s: D.cnot(s, c_d, t_any, N)),
        (make_frame_cz(c_d, t_any), lambda s: D.cz(s, c_d, t_any, N)),
        (make_frame_s(d1), lambda s: D.apply_1q(s, D.S, d1, N)),
        (make_frame_s_dag(d1), lambda s: D.apply_1q(s, D.SDG, d1, N)),
        (make_frame_swap(d1, d2), lambda s: D.swap(s, d1, d2, N)),
    ]:
        st = prepared(orc, rng, k)
        before = D.state_of(st, N)
        orc.execute(build_program([instr], peak_rank=N, num_qubits=N), st)
        np.testing.assert_allclose(D.state_of(st, N), U(before), atol=1e-13, err_msg=str(instr))


@pytest.mark.parametrize("seed", range(4))
def test_expand_family_dense(orc, seed):
    """FRAME_H(k) + EXPAND*(k) == H (then T/T_dag/diag(1,z)) on dormant qubit k
    (route_to_active_z, src/clifft/backend/backend.cc:360-395)."""
    rng = np.random.default_rng(seed)
    k = 2
    z = np.exp(1j * rng.uniform(0, 2 * np.pi))
    for instr, G in [(make_expand(k), D.I2), (make_expand_t(k), D.T), (make_expand_t_dag(k), D.TDG),
                     (make_expand_rot(k, z.real, z.imag), np.diag([1, z]))]:
        st = prepared(orc, rng, k)
        before = D.state_of(st, N)
        orc.execute(build_program([make_frame_h(k), instr], peak_rank=N, num_qubits=N), st)
        assert st.active_k == k + 1
        np.testing.assert_allclose(D.state_of(st, N), D.apply_1q(D.apply_1q(before, D.H, k, N), G, k, N),
                                   atol=1e-13)


# ----------------------------------------------------------------------------
# Measurements (forced outcomes) against projectors
# ----------------------------------------------------------------------------
def forced_meas(orc, st, instr, rec_bit):
    prog = build_program([instr], peak_rank=N, num_qubits=N, num_measurements=4).forced()
    st.set_forced_record(np.array([rec_bit, 0, 0, 0], np.uint8))
    orc.execute(prog, st)
    return st


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("sign", [False, True])
def test_active_measurements_dense(orc, seed, sign):
    rng = np.random.default_rng(seed)
    k = 4
    t = k - 1
    f = int(rng.integers(0, t))
    for m in (0, 1):
        r = m ^ int(sign)
        for instr, post in [
            (make_meas(Op.MEAS_ACTIVE_DIAGONAL, t, 0, sign), lambda s: D.project(s, t, "z", m, N)),
            (make_meas(Op.MEAS_ACTIVE_INTERFERE, t, 0, sign),
             lambda s: D.apply_1q(D.project(s, t, "x", m, N), D.H, t, N)),
            (make_swap_meas_interfere(f, t, 0, sign),
             lambda s: D.apply_1q(D.project(D.swap(s, f, t, N), t, "x", m, N), D.H, t, N)),
        ]:
            st = prepared(orc, rng, k)
            before = D.state_of(st, N)
            before /= np.linalg.norm(before)
            proj = post(before)
            p = np.vdot(proj, proj).real
            forced_meas(orc, st, instr, r)
            assert st.forced_reachable
            assert int(st.meas_record[0]) == r
            assert abs(st.forced_log_probability - math.log(p)) < 1e-12
            assert st.active_k == k - 1
            np.testing.assert_allclose(D.state_of(st, N), proj / math.sqrt(p), atol=1e-12)


@pytest.mark.parametrize("seed", range(3))
def test_dormant_measurements_dense(orc, seed):
    rng = np.random.default_rng(seed)
    k = 2
    v = int(rng.integers(k, N))
    for m in (0, 1):
        st = prepared(orc, rng, k)
        before = D.state_of(st, N)
        before /= np.linalg.norm(before)
        post = D.apply_1q(D.project(before, v, "x", m, N), D.H, v, N)
        forced_meas(orc, st, make_meas(Op.MEAS_DORMANT_RANDOM, v, 0, False), m)
        assert abs(st.forced_log_probability - math.log(0.5)) < 1e-15
        np.testing.assert_allclose(D.state_of(st, N), post / np.linalg.norm(post), atol=1e-12)
    st = prepared(orc, rng, k)
    expect = (int(st.p_x[0]) >> v) & 1
    before = D.state_of(st, N)
    forced_meas(orc, st, make_meas(Op.MEAS_DORMANT_STATIC, v, 0, False), expect)
    assert st.forced_reachable and int(st.meas_record[0]) == expect
    np.testing.assert_allclose(D.state_of(st, N), before, atol=1e-14)
    st = prepared(orc, rng, k)
    expect = (int(st.p_x[0]) >> v) & 1
    forced_meas(orc, st, make_meas(Op.MEAS_DORMANT_STATIC, v, 0, False), expect ^ 1)
    assert not st.forced_reachable


def test_zero_probability_branches_are_deterministic(orc):
    # tests/test_svm.cc:1445-1571: |+> interfere -> 0, |0> diagonal -> 0
    p = build_program([make_expand(0), make_meas(Op.MEAS_ACTIVE_INTERFERE, 0, 0)], peak_rank=1, num_measurements=1)
    m, *_ = orc.sample(p, 64, seed=3)
    assert not m.any()
    p = build_program([make_expand(0), make_array_h(0), make_meas(Op.MEAS_ACTIVE_DIAGONAL, 0, 0)], peak_rank=1,
                      num_measurements=1)
    m, *_ = orc.sample(p, 64, seed=3)
    assert not m.any()


def test_dust_clamp_t4_interference(orc):
    # tests/test_svm.cc:1975-2003: H T^4 H |0> = |1> up to dust
    p = build_program([make_expand(0), make_array_t(0), make_array_t(0), make_array_t(0), make_array_t(0),
                       make_array_h(0), make_meas(Op.MEAS_ACTIVE_DIAGONAL, 0, 0)], peak_rank=1, num_measurements=1)
    m, *_ = orc.sample(p, 32, seed=1)
    assert m.all()


# ----------------------------------------------------------------------------
# EXP_VAL against the dense Pauli expectation
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(5))
def test_exp_val_dense(orc, seed):
    rng = np.random.default_rng(seed)
    k = 3
    masks = []
    for i in range(12):
        x = int(rng.integers(0, 1 << N))
        if i % 3:
            x &= (1 << k) - 1
        masks.append((x, int(rng.integers(0, 1 << N)), bool(rng.random() < 0.5)))
    bc = [make_exp_val(i, i) for i in range(len(masks))]
    p = build_program(bc, peak_rank=N, num_qubits=N, num_exp_vals=len(masks), exp_val_masks=masks)
    st = orc.State(N, 0, num_qubits=N, num_exp_vals=len(masks))
    phi = D.random_state(rng, k)
    st.v[: 1 << k] = phi
    st.active_k = k
    st.p_x[0], st.p_z[0] = int(rng.integers(0, 1 << N)), int(rng.integers(0, 1 << N))
    st.gamma = np.exp(1j * 0.3)
    psi = D.state_of(st, N)
    orc.execute(p, st)
    for i, (x, z, sgn) in enumerate(masks):
        ref = np.vdot(psi, D.hermitian_pauli(psi, x, z, sgn, N)).real / np.vdot(psi, psi).real
        assert abs(st.exp_vals[i] - ref) < 1e-12, (i, x, z, sgn)


# ----------------------------------------------------------------------------
# Noise and drivers
# ----------------------------------------------------------------------------
def noisy_program(block):
    sites = [[(1 << (s % 3), 0, 0.2)] for s in range(6)]
    bc = []
    if block:
        bc.append(make_noise_block(0, 6))
    else:
        bc += [make_noise(s) for s in range(6)]
    bc += [make_meas(Op.MEAS_DORMANT_STATIC, q, q) for q in range(3)]
    return build_program(bc, peak_rank=1, num_qubits=3, num_measurements=3, noise_sites=sites)


def test_noise_block_equals_noise_sequence(orc):
    a = orc.sample(noisy_program(False), 2000, seed=9)[0]
    b = orc.sample(noisy_program(True), 2000, seed=9)[0]
    assert np.array_equal(a, b)


def test_noise_firing_rate_binomial(orc):
    # each qubit q gets two independent X flips with p=0.2: P(bit=1) = 2p(1-p)
    m, *_ = orc.sample(noisy_program(False), 20000, seed=5, per_shot=False)
    p = 2 * 0.2 * 0.8
    sigma = math.sqrt(p * (1 - p) / len(m))
    for q in range(3):
        assert abs(m[:, q].mean() - p) < 5 * sigma


def test_expected_observables_and_postselect(orc):
    bc = [make_meas(Op.MEAS_DORMANT_STATIC, 0, 0, True),  # record 1
          make_meas(Op.MEAS_DORMANT_STATIC, 1, 1, False),  # record 0
          make_detector(0, 0, expected_one=True),  # parity(1) ^ 1 = 0
          make_observable(0, 0),
          make_postselect(1, 1, expected_one=False)]  # parity(rec0, rec1) = 1 -> discard
    p = build_program(bc, peak_rank=1, num_qubits=2, num_measurements=2, num_detectors=2, num_observables=1,
                      detector_targets=[[0], [0, 1]], observable_targets=[[0]], expected_observables=[1])
    m, d, o, e, disc = orc.sample(p, 4, seed=1)
    assert disc.all() and (d[:, 0] == 0).all() and (o[:, 0] == 0).all()  # obs 1 ^ expected 1
    surv = orc.sample_survivors(p, 4, seed=1)
    assert surv["passed"] == 0


def test_record_probabilities_sum_to_one(orc):
    from randprog import random_measure_only_program
    for seed in range(3):
        p = random_measure_only_program(seed, num_qubits=5, peak=4, length=40)
        nm = p.num_measurements
        if nm == 0 or nm > 12:
            continue
        recs = ((np.arange(1 << nm)[:, None] >> np.arange(nm)) & 1).astype(np.uint8)
        lp = orc.record_log_probs(p.forced(), recs)
        probs = np.where(lp > -1e300, np.exp(lp), 0.0)
        assert abs(probs.sum() - 1.0) < 1e-9
