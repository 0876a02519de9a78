"""Device executor vs CPU oracle on the same programs and seeds.

Records (measurements, detectors, observables) must be bit-exact: under
forced outcomes exactly, and in sampled mode because both sides draw from
the same per-shot xoshiro256++ streams (a knife-edge comparison flip from a
different reduction order has probability ~1e-16 per draw).  Amplitudes,
gamma, expectation values and log-probabilities agree to 1e-12 absolute.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from randprog import random_measure_only_program, random_program

pytestmark = pytest.mark.gpu

ATOL = 1e-12


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def svm():
    import clifft_amd.svm as svm
    from clifft_amd import _lib
    assert _lib.backend() == "hip:gfx950", _lib.backend()
    return svm


# (num_qubits, peak): one shot per lane (peak <= 6, <= 64 qubits), wave-team
# LDS, workgroup-team LDS, global-memory tier up to 2^20 amplitudes (16 MiB
# states), and >64-qubit frames
SHAPES = [(5, 0), (6, 1), (10, 2), (8, 3), (64, 3), (12, 6), (16, 8), (20, 10), (24, 12), (24, 13), (24, 14),
          (30, 15), (24, 16), (26, 18), (28, 20), (100, 3), (100, 6), (130, 9)]


def shots_for(peak):
    return 300 if peak < 12 else (64 if peak <= 16 else (8 if peak <= 18 else 3))


@pytest.mark.parametrize("nq,peak", SHAPES)
@pytest.mark.parametrize("seed", [1, 2])
def test_sample_bit_exact(svm, orc, nq, peak, seed):
    shots = shots_for(peak)
    announce("sample", nq, peak, seed, shots)
    p = random_program(1000 * seed + peak, num_qubits=nq, peak=peak, length=160 + 10 * peak)
    res = svm.sample(p, shots, seed=seed)
    m, d, o, e, _ = orc.sample(p, shots, seed=seed, per_shot=True)
    assert np.array_equal(res.measurements, m)
    assert np.array_equal(res.detectors, d)
    assert np.array_equal(res.observables, o)
    np.testing.assert_allclose(res.exp_vals, e, atol=ATOL, rtol=0)


@pytest.mark.parametrize("nq,peak", [(8, 2), (40, 5), (20, 9)])
def test_wide_parity_lists(svm, orc, nq, peak):
    """Detector/observable lists of up to 8 targets with repeats, over
    records of more than 64 slots: parities held inline in the lane
    bytecode (one or two record words, repeats cancelled) and lists that
    keep the CSR tables (three or more words)."""
    announce("wide-lists", nq, peak)
    p = random_program(4242 + peak, num_qubits=nq, peak=peak, length=900, wide_lists=True)
    assert p.total_meas_slots > 64 and p.num_detectors > 0
    res = svm.sample(p, 256, seed=9)
    m, d, o, e, _ = orc.sample(p, 256, seed=9, per_shot=True)
    assert np.array_equal(res.measurements, m)
    assert np.array_equal(res.detectors, d)
    assert np.array_equal(res.observables, o)
    packed = svm.sample(p, 256, seed=9, packed=True)
    assert np.array_equal(np.packbits(d, axis=1, bitorder="little"), packed.detectors)


@pytest.mark.parametrize("peak", [2, 4, 9, 14])
def test_packed_records(svm, peak):
    announce("packed", peak)
    p = random_program(77 + peak, num_qubits=20, peak=peak, length=200)
    a = svm.sample(p, 100, seed=5)
    b = svm.sample(p, 100, seed=5, packed=True)
    for x, y in ((a.measurements, b.measurements), (a.detectors, b.detectors), (a.observables, b.observables)):
        assert np.array_equal(np.packbits(x, axis=1, bitorder="little"), y)


@pytest.mark.parametrize("nq,peak", [(6, 2), (9, 3), (8, 4), (12, 7), (16, 10), (20, 14)])
def test_record_log_probs(svm, orc, nq, peak):
    announce("forced", nq, peak)
    p = random_measure_only_program(peak, num_qubits=nq, peak=peak, length=120)
    if p.num_measurements == 0:
        pytest.skip("no measurements drawn")
    recs, *_ = orc.sample(p, 40, seed=3)
    rng = np.random.default_rng(peak)
    flipped = recs.copy()
    flipped[::3, rng.integers(0, p.num_measurements)] ^= 1  # some become unreachable
    allrec = np.concatenate([recs, flipped])
    gpu = svm.record_probabilities(p, allrec)
    ref = orc.record_log_probs(p.forced(), allrec)
    unreach_g, unreach_r = gpu < -1e300, ref < -1e300
    assert np.array_equal(unreach_g, unreach_r)
    np.testing.assert_allclose(gpu[~unreach_g], ref[~unreach_r], atol=1e-10, rtol=1e-12)


@pytest.mark.parametrize("nq,t_count", [(20, 9), (24, 14)])
def test_record_log_probs_compiled(svm, orc, nq, t_count):
    """Forced execution of compiled Clifford+T programs, whose final blocks
    hold frame fans and array fan-outs (with the MULTI_CZ that follows them)
    in the forced copy of the team bytecode: log-probabilities of sampled
    records, and of records with one flipped outcome, against the oracle."""
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile
    announce("forced-compiled", nq, t_count)
    p = ccompile(W.random_clifford_t(n=nq, layers=20, t_count=t_count, seed=nq * 3 + t_count, final_measure=True))
    recs, *_ = orc.sample(p, 24, seed=5)
    rng = np.random.default_rng(t_count)
    flipped = recs.copy()
    flipped[::2, rng.integers(0, p.num_measurements)] ^= 1
    allrec = np.concatenate([recs, flipped])
    gpu = svm.record_probabilities(p, allrec)
    ref = orc.record_log_probs(p.forced(), allrec)
    unreach_g, unreach_r = gpu < -1e300, ref < -1e300
    assert np.array_equal(unreach_g, unreach_r)
    np.testing.assert_allclose(gpu[~unreach_g], ref[~unreach_r], atol=1e-10, rtol=1e-12)


@pytest.mark.parametrize("nq,peak", [(10, 5), (16, 9), (18, 14)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_execute_preset_state(svm, orc, nq, peak, seed):
    """Single-shot execute from a preset state: amplitudes, gamma, frame,
    records and PRNG state after the shot (tests/test_svm.cc:28-67 style)."""
    announce("execute", nq, peak, seed)
    p = random_program(seed * 31 + peak, num_qubits=nq, peak=peak, length=150, noise=False)
    rng = np.random.default_rng(seed)
    o = orc.State.for_program(p, seed=seed)
    g = svm.State.for_program(p, seed=seed)
    o.next_noise_idx = 0xFFFFFFFF
    g.next_noise_idx = 0xFFFFFFFF
    svm.execute(p, g)
    import ctypes
    from oracle import oracle as orcmod
    orcmod.execute(p, o)
    assert g.active_k == o.active_k
    n = 1 << g.active_k
    np.testing.assert_allclose(g.v[:n], o.v[:n], atol=ATOL, rtol=0)
    assert abs(g.gamma - o.gamma) < ATOL
    assert int(g.p_x[0]) == int(o.p_x[0]) and int(g.p_z[0]) == int(o.p_z[0])
    assert np.array_equal(g.meas_record, o.meas_record)
    np.testing.assert_allclose(g.exp_vals, o.exp_vals, atol=ATOL, rtol=0)
    assert list(map(int, g.rng)) == list(map(int, o.rng))


@pytest.mark.parametrize("nq,t_count", [(12, 5), (20, 9), (24, 14)])
def test_execute_preset_frame(svm, orc, nq, t_count):
    """Single-shot execute from a state whose Pauli frame is not zero.  The
    sampling copies of the bytecode leave out the frame opcodes that are
    the identity from the zero frame (most of a Clifford+T program's frame
    opcodes before its first measurement); from this frame they are not, and
    the device must match the oracle that runs every opcode."""
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile
    from oracle import oracle as orcmod
    announce("execute-frame", nq, t_count)
    p = ccompile(W.random_clifford_t(n=nq, layers=20, t_count=t_count, seed=nq + t_count, final_measure=True))
    frame_ops = np.count_nonzero(p.bytecode["opcode"] <= 5)
    assert frame_ops >= 50, frame_ops
    rng = np.random.default_rng(t_count)
    o = orc.State.for_program(p, seed=3)
    g = svm.State.for_program(p, seed=3)
    o.next_noise_idx = 0xFFFFFFFF
    g.next_noise_idx = 0xFFFFFFFF
    fx, fz = (int(rng.integers(1, 1 << nq)) for _ in range(2))
    o.p_x[0] = fx
    o.p_z[0] = fz
    g.p_x[0] = fx
    g.p_z[0] = fz
    svm.execute(p, g)
    orcmod.execute(p, o)
    assert g.active_k == o.active_k
    n = 1 << g.active_k
    np.testing.assert_allclose(g.v[:n], o.v[:n], atol=ATOL, rtol=0)
    assert abs(g.gamma - o.gamma) < ATOL
    assert int(g.p_x[0]) == int(o.p_x[0]) and int(g.p_z[0]) == int(o.p_z[0])
    assert np.array_equal(g.meas_record, o.meas_record)
    # the same program sampled from the zero frame, through the pruned copies
    res = svm.sample(p, 200, seed=11)
    m, _, _, _, _ = orc.sample(p, 200, seed=11, per_shot=True)
    assert np.array_equal(res.measurements, m)


@pytest.mark.parametrize("peak", [2, 5])
def test_survivors_counts(svm, orc, peak):
    announce("survivors", peak)
    p = random_program(5, num_qubits=12, peak=peak, length=250, postselect=True)
    if not p.has_postselection:
        pytest.skip("no postselect drawn")
    g = svm.sample_survivors(p, 500, seed=11, keep_records=True)
    r = orc.sample_survivors(p, 500, seed=11)
    assert g.passed_shots == r["passed"]
    assert g.logical_errors == r["logical_errors"]
    assert np.array_equal(g.observable_ones, r["observable_ones"])
    assert np.array_equal(g.measurements, r["measurements"])
    assert np.array_equal(g.detectors, r["detectors"])


@pytest.mark.parametrize("peak", [3, 5])
def test_reference_stream_first_shot(svm, orc, peak):
    """Shot 0 of the device equals the first shot of the reference's single
    stream (svm.cc:210-226)."""
    announce("first-shot", peak)
    p = random_program(8, num_qubits=10, peak=peak, length=200)
    m, d, o, e, _ = orc.sample(p, 1, seed=1234, per_shot=False)
    res = svm.sample(p, 1, seed=1234)
    assert np.array_equal(res.measurements, m) and np.array_equal(res.detectors, d)


@pytest.mark.parametrize("peak", [3, 5, 9, 15])
def test_fanout_cnot_runs(svm, orc, peak):
    """Runs of CNOTs sharing a control, lane and team tiers (peak 15: the
    global tier applies each run in one fused pass), against the oracle's
    one-CNOT-at-a-time execution."""
    from randprog import fanout_program
    announce("fanout", peak)
    for seed in range(3):
        p = fanout_program(seed, peak)
        shots = 64 if peak >= 12 else 400
        res = svm.sample(p, shots, seed=seed)
        m, d, o, e, _ = orc.sample(p, shots, seed=seed, per_shot=True)
        assert np.array_equal(res.measurements, m)
        np.testing.assert_allclose(res.exp_vals, e, atol=ATOL, rtol=0)


def mono_ops_at_rank(p, kmin):
    """Permutation/phase opcodes executed at static rank >= kmin (the ones
    the global tier fuses into runs)."""
    from clifft_amd.program import Op
    mono = {Op.ARRAY_CNOT, Op.ARRAY_CZ, Op.ARRAY_SWAP, Op.ARRAY_MULTI_CNOT, Op.ARRAY_MULTI_CZ, Op.ARRAY_S,
            Op.ARRAY_S_DAG, Op.ARRAY_T, Op.ARRAY_T_DAG, Op.ARRAY_ROT}
    phase = {Op.ARRAY_T, Op.ARRAY_T_DAG, Op.ARRAY_ROT}
    h = p.active_k_history()
    ops, a1 = p.bytecode["opcode"], p.bytecode["axis_1"]
    return sum(1 for i, op in enumerate(ops) if op in mono and h[i] >= kmin and (op not in phase or a1[i] < h[i]))


@pytest.mark.parametrize("peak,shots,seed", [(14, 1100, 4270), (15, 600, 4202)])
def test_global_tier_team_sizes(svm, orc, peak, shots, seed):
    """Enough shots that the launch picks smaller global-memory teams (256
    threads at peak 14, 512 at peak 15); the programs reach their peak rank
    and hold more than 50 permutation/phase opcodes at rank >= 10 (fused
    runs): records bit-exact and expectation values within 1e-12 of the
    oracle."""
    announce("global-teams", peak, shots, seed)
    p = random_program(seed, num_qubits=peak + 6, peak=peak, length=220, p_array=0.5)
    assert p.active_k_history().max() == peak and mono_ops_at_rank(p, 10) > 50
    r = svm.sample(p, shots, seed=99)
    m, d, o, e, _ = orc.sample(p, shots, seed=99, per_shot=True)
    assert np.array_equal(r.measurements, m)
    assert np.array_equal(r.detectors, d)
    assert np.array_equal(r.observables, o)
    np.testing.assert_allclose(r.exp_vals, e, atol=ATOL, rtol=0)


def identity_and_unequal(p):
    """Whether a program holds a FLAG_IDENTITY measurement and a noise site
    whose channels have unequal probabilities."""
    from clifft_amd.program import FLAG_IDENTITY, Op
    ident = bool(((p.bytecode["opcode"] == Op.MEAS_DORMANT_STATIC) & ((p.bytecode["flags"] & FLAG_IDENTITY) != 0)).any())
    off, prob = p.site_chan_off, p.chan_prob
    unequal = any(len(set(prob[off[s]:off[s + 1]].tolist())) > 1 for s in range(len(off) - 1))
    return ident, unequal


@pytest.mark.parametrize("nq,peak", [(6, 1), (10, 4), (12, 6), (14, 7), (16, 9), (20, 12), (100, 5)])
def test_identity_meas_and_unequal_channels(svm, orc, nq, peak):
    """Programs holding FLAG_IDENTITY measurements (record = sign bit,
    svm_kernels.inl:2361-2369) and noise sites with unequal channel
    probabilities (channel chosen by the running sum, :1899-1911) on every
    tier: records bit-exact against the oracle, sampled and forced."""
    announce("identity-unequal", nq, peak)
    hits = [False, False]
    for seed in range(4):
        p = random_program(5000 + 17 * seed + peak, num_qubits=nq, peak=peak, length=300)
        ident, unequal = identity_and_unequal(p)
        hits[0] |= ident
        hits[1] |= unequal
        res = svm.sample(p, 500, seed=seed)
        m, d, o, e, _ = orc.sample(p, 500, seed=seed, per_shot=True)
        assert np.array_equal(res.measurements, m)
        assert np.array_equal(res.detectors, d)
        assert np.array_equal(res.observables, o)
    assert hits == [True, True]
    # forced execution of a measure-only program holding identity measurements
    q = next(q for q in (random_measure_only_program(s, num_qubits=max(nq, 8), peak=min(peak, 8), length=150)
                         for s in range(900 + peak, 960 + peak))
             if identity_and_unequal(q)[0] and 0 < q.num_measurements <= 64)
    recs, *_ = orc.sample(q, 30, seed=4)
    recs[::4] ^= 1  # some unreachable, identity records among them
    gpu = svm.record_probabilities(q, recs)
    ref = orc.record_log_probs(q.forced(), recs)
    assert np.array_equal(gpu < -1e300, ref < -1e300)
    ok = ref > -1e300
    np.testing.assert_allclose(gpu[ok], ref[ok], atol=1e-10, rtol=1e-12)


def decay_program(peak, rounds, seed):
    """Expansions to rank peak-1, then `rounds` rounds of: EXPAND of a new
    top axis (exactly |+>), CNOTs from older axes into it (which leave |+>
    unchanged) and CZs among the older axes -- a permutation/phase run over
    a pending expansion -- then an interfere measurement of the new axis,
    certain to give +.  Every round scales gamma by 1/sqrt 2 (and the
    amplitudes by sqrt 2: the fold of |+>), so gamma falls below 1e-100
    after about 660 rounds and the amplitudes are renormalised
    (svm.h:177-186)."""
    from clifft_amd.program import (Op, build_program, make_array_cnot, make_array_cz, make_expand,
                                    make_expand_t, make_frame_h, make_meas)
    rng = np.random.default_rng(seed)
    bc = []
    for v in range(peak - 1):
        bc += [make_frame_h(v), make_expand_t(v)]
    k = peak - 1
    for r in range(rounds):
        bc.append(make_expand(k))
        for c in rng.choice(k, size=3, replace=False):
            bc.append(make_array_cnot(int(c), k))
        a, b = rng.choice(k, size=2, replace=False)
        bc.append(make_array_cz(int(a), int(b)))
        bc.append(make_meas(Op.MEAS_ACTIVE_INTERFERE, k, r))
    return build_program(bc, peak_rank=peak, num_qubits=peak, num_measurements=rounds)


@pytest.mark.parametrize("peak", [9, 14])
def test_renormalisation(svm, orc, peak):
    """Gamma renormalisation in the middle of a shot (LDS tier at peak 9,
    global tier with pending expansions read through at peak 14): final
    amplitudes (relative to their largest), gamma (relative) and records
    against the oracle's execute, and sampled records."""
    announce("renormalisation", peak)
    rounds = 800
    p = decay_program(peak, rounds, seed=peak)
    o = orc.State.for_program(p, seed=5)
    g = svm.State.for_program(p, seed=5)
    svm.execute(p, g)
    from oracle import oracle as orcmod
    orcmod.execute(p, o)
    assert g.active_k == o.active_k == peak - 1
    n = 1 << g.active_k
    scale = np.abs(o.v[:n]).max()
    assert 0.0 < scale < 1e60  # rescaled: 800 folds alone would take them past 1e100
    np.testing.assert_allclose(g.v[:n] / scale, o.v[:n] / scale, atol=ATOL, rtol=0)
    assert abs(g.gamma / o.gamma - 1.0) < 1e-10
    assert np.array_equal(g.meas_record, o.meas_record) and not np.asarray(o.meas_record[:rounds]).any()
    r = svm.sample(p, 8, seed=3)
    m, d, ob, e, _ = orc.sample(p, 8, seed=3, per_shot=True)
    assert np.array_equal(r.measurements, m)


@pytest.mark.parametrize("peak,seed", [(12, 7), (14, 8)])
def test_global_wavefront_teams(svm, orc, peak, seed):
    """A launch with more shots than the device holds wavefront teams runs
    global-amplitude shots one per wavefront: its first 200 shots (each
    shot's stream depends only on the seed and the shot index) against the
    oracle's 200-shot run."""
    announce("global-wavefront-teams", peak, seed)
    p = random_program(seed, num_qubits=peak + 6, peak=peak, length=200, p_array=0.5)
    assert p.active_k_history().max() == peak
    r = svm.sample(p, 4500, seed=11)
    m, d, o, e, _ = orc.sample(p, 200, seed=11, per_shot=True)
    assert np.array_equal(r.measurements[:200], m)
    assert np.array_equal(r.detectors[:200], d)
    assert np.array_equal(r.observables[:200], o)
    np.testing.assert_allclose(r.exp_vals[:200], e, atol=ATOL, rtol=0)
