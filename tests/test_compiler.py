"""Pins the Program producer end to end against dense statevector physics.

* EXP_VAL probes of compiled unitary Clifford+T+R_Z circuits (run on the
  oracle VM) equal <psi|P|psi> of the dense simulation (1e-12).
* record_probabilities of compiled measurement circuits (oracle, forced
  outcomes) equal the dense Born probabilities of every record (1e-12), and
  sum to one.
* The static active-rank trajectory respects the VM contract (the compiled
  program passes the executor's host validation).
* Noise channels fire at their stated rate (5-sigma binomial).
"""

import itertools
import math

import numpy as np
import pytest

import dense_circuit as DC
from clifft_amd.compiler import compile as ccompile
from clifft_amd.program import Op


@pytest.fixture(scope="module", autouse=True)
def _built():
    from clifft_amd import compiler as c
    c.build()


def oracle_exp_vals(orc, prog):
    st = orc.State.for_program(prog, seed=1)
    st.next_noise_idx = 0xFFFFFFFF
    orc.execute(prog, st)
    return np.array(st.exp_vals)


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("passes", [True, False])
def test_exp_val_matches_dense(orc, seed, passes):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 7))
    text = DC.random_circuit(rng, n, depth=5, t_count=int(rng.integers(0, 5)), rz=int(rng.integers(0, 3)),
                             exp_vals=8)
    prog = ccompile(text, passes=passes)
    _, ev, _ = DC.run(text)
    got = oracle_exp_vals(orc, prog)
    np.testing.assert_allclose(got, ev, atol=1e-12, err_msg=text)


@pytest.mark.parametrize("seed", range(10))
def test_record_probabilities_match_dense(orc, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 5))
    text = DC.random_circuit(rng, n, depth=4, t_count=int(rng.integers(1, 4)), measure=True, mr=True)
    prog = ccompile(text)
    nm = prog.num_measurements
    assert nm <= 12
    recs = np.array(list(itertools.product([0, 1], repeat=nm)), dtype=np.uint8)
    lp = orc.record_log_probs(prog.forced(), recs)
    total = 0.0
    for r, l in zip(recs, lp):
        _, _, ref = DC.run(text, record=list(r))
        if ref == -np.inf or ref < math.log(1e-14):
            assert l < -1e300 or l < math.log(1e-13), (text, r, l, ref)
        else:
            assert abs(l - ref) < 1e-10, (text, r, l, ref)
            total += math.exp(l)
    assert abs(total - 1.0) < 1e-9


def test_rank_trajectory_and_validation():
    rng = np.random.default_rng(7)
    text = DC.random_circuit(rng, 8, depth=12, t_count=6, measure=True, mr=True)
    prog = ccompile(text)
    hist = prog.active_k_history()
    assert hist.max() <= prog.peak_rank
    import ctypes
    import clifft_amd._lib as L
    h = ctypes.c_void_p()
    v = prog.view()
    rc = L.lib().clifft_svm_program_create(ctypes.byref(v), ctypes.byref(h))
    assert rc in (0, L.E_NODEVICE), L.lib().clifft_svm_last_error()


def test_x_error_rate(orc):
    text = "X_ERROR(0.2) 0 1\nM 0 1\nDEPOLARIZE1(0.3) 2\nM 2"
    prog = ccompile(text)
    m, *_ = orc.sample(prog, 20000, seed=3)
    for q, p in ((0, 0.2), (1, 0.2), (2, 0.2)):  # DEPOLARIZE1 flips Z-basis result w.p. 2p/3
        sigma = math.sqrt(p * (1 - p) / len(m))
        assert abs(m[:, q].mean() - p) < 5 * sigma


def test_detectors_and_observables(orc):
    # Bell-pair parity check: deterministic detector, observable on qubit 0
    text = """H 0
CX 0 1
M 0 1
DETECTOR rec[-1] rec[-2]
OBSERVABLE_INCLUDE(0) rec[-2]
"""
    prog = ccompile(text)
    m, d, o, *_ = orc.sample(prog, 2000, seed=4)
    assert not d.any()
    assert np.array_equal(o[:, 0], m[:, 0])
    assert 0.4 < m[:, 0].mean() < 0.6


def test_parameter_slots_keep_rotations():
    text = "H 0\nR_Z(0.3) 0\nH 1\nCX 0 1\nR_Z(-0.7) 1\nEXP_VAL Z0*Z1"
    prog = ccompile(text, keep_rotations=True)
    assert len(prog.param_instr) == 2
    for i in prog.param_instr:
        assert prog.bytecode["opcode"][i] in (Op.ARRAY_ROT, Op.EXPAND_ROT)
    assert set(np.abs(prog.param_sign)) == {1}


def test_blob_round_trip():
    rng = np.random.default_rng(9)
    text = DC.random_circuit(rng, 6, depth=6, t_count=4, measure=True, exp_vals=3)
    prog = ccompile(text)
    from clifft_amd.program import Program
    back = Program.from_bytes(prog.to_bytes())
    assert back.bytecode.tobytes() == prog.bytecode.tobytes()
    assert back.u2_nodes.tobytes() == prog.u2_nodes.tobytes()
    assert (back.peak_rank, back.num_measurements) == (prog.peak_rank, prog.num_measurements)


def tile_circuit(rng, noisy):
    """Two active qubits hammered by a long run of one- and two-qubit gates
    on that pair (what the tile-axis fusion pass fuses), plus a spectator.
    With `noisy`, depolarizing noise right before the run makes the frame
    entering the fused node random (all 16 frame states occur)."""
    lines = ["H 0 1 2", "T 0 1"]
    if noisy:
        lines += ["DEPOLARIZE2(0.3) 0 1", "DEPOLARIZE1(0.3) 0 1"]
    for _ in range(8):
        lines.append(str(rng.choice(["CX 0 1", "CX 1 0", "CZ 0 1", "SWAP 0 1"])))
        lines.append(str(rng.choice(["T 0", "T 1", "T_DAG 0", "T_DAG 1"])))
        lines.append(str(rng.choice(["H 0", "H 1", "S 0", "S_DAG 1", "Z 0"])))
    lines.append("CX 2 0")
    return lines


def _tile_text(seed, noisy, tail):
    """First circuit of the seeded family whose compiled form has a U4 node."""
    from clifft_amd.compiler import compile as cc
    from clifft_amd.program import Op
    rng = np.random.default_rng(seed)
    for _ in range(50):
        text = "\n".join(tile_circuit(rng, noisy) + [tail])
        if (cc(text).bytecode["opcode"] == Op.ARRAY_U4).sum() >= 1:
            return text
    raise AssertionError("no circuit with a tile run")


@pytest.mark.parametrize("seed", range(6))
def test_tile_fusion_matches_dense_and_unfused(orc, seed):
    from clifft_amd.compiler import compile as cc
    from dense_circuit import run
    text = _tile_text(seed, False, "EXP_VAL X0*X1 Z0 Y1 Z0*Z1 X0*Y1*Z2")
    _, _, _, e, _ = orc.sample(cc(text), 1, seed=0)
    np.testing.assert_allclose(e[0], run(text)[1], atol=1e-12, rtol=0)
    # noisy variant: every incoming frame state of the node is exercised;
    # the fused and unfused programs draw the same random numbers
    text = _tile_text(seed, True, "M 0 1 2")
    f, u = cc(text), cc(text, fusion=False)
    a = orc.sample(f, 3000, seed=seed)
    b = orc.sample(u, 3000, seed=seed)
    assert np.array_equal(a[0], b[0])
