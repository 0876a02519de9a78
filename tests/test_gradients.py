"""CPU checks of the gradient machinery's ground truth (oracle only).

The device adjoint kernel is compared against these in test_gpu_adjoint.py;
here the references themselves are pinned: a parameter slot's payload edit
equals recompiling the circuit with the new angle, the parameter-shift rule
agrees with central finite differences of the oracle, and the compiled
expectation values agree with a dense state-vector model."""

from __future__ import annotations

import numpy as np
import pytest

from gradprog import oracle_exp_vals, oracle_objective, oracle_param_shift, rz_circuit, with_thetas


@pytest.fixture(scope="module")
def cc():
    from clifft_amd.compiler import compile as c
    return c


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_payload_edit_equals_recompile(orc, cc, seed):
    build, coeffs = rz_circuit(seed=seed)
    rng = np.random.default_rng(seed)
    th0 = rng.uniform(0, 2 * np.pi, 6)
    th1 = rng.uniform(0, 2 * np.pi, 6)
    p0 = cc(build(th0), keep_rotations=True)
    p1 = cc(build(th1), keep_rotations=True)
    assert np.array_equal(p0.param_instr, p1.param_instr)
    e_edit = oracle_exp_vals(orc, p0, th1)
    _, _, _, e_rec, _ = orc.sample(p1, 1, seed=0)
    np.testing.assert_allclose(e_edit, e_rec[0], atol=1e-13, rtol=0)


@pytest.mark.parametrize("seed", [0, 3])
def test_param_shift_matches_central_fd(orc, cc, seed):
    build, coeffs = rz_circuit(seed=seed)
    th = np.random.default_rng(seed + 10).uniform(0, 2 * np.pi, 6)
    p = cc(build(th), keep_rotations=True)
    ps = oracle_param_shift(orc, p, th, coeffs)
    h = 1e-6
    fd = np.zeros_like(ps)
    for i in range(len(th)):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (oracle_objective(orc, p, tp, coeffs) - oracle_objective(orc, p, tm, coeffs)) / (2 * h)
    np.testing.assert_allclose(ps, fd, atol=2e-8, rtol=0)
    assert np.abs(ps).max() > 1e-3  # not a vacuous comparison


def test_exp_vals_match_dense_model(orc, cc):
    from dense_circuit import run
    build, coeffs = rz_circuit(n=6, layers=6, seed=5)
    th = np.random.default_rng(5).uniform(0, 2 * np.pi, 6)
    text = build(th)
    p = cc(text, keep_rotations=True)
    _, evs, _ = run(text)
    np.testing.assert_allclose(oracle_exp_vals(orc, p, th), evs, atol=1e-12, rtol=0)


def test_parameter_slots_cover_both_rotation_kinds(cc):
    from clifft_amd.program import Op
    kinds = set()
    for seed in range(6):
        build, _ = rz_circuit(seed=seed)
        p = cc(build(np.zeros(6)), keep_rotations=True)
        kinds |= {int(p.bytecode[pc]["opcode"]) for pc in p.param_instr}
    assert kinds == {Op.ARRAY_ROT, Op.EXPAND_ROT}
