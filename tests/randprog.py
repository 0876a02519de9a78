"""Random well-formed programs exercising every opcode (test helper).

Programs respect the VM's static invariants (EXPAND on axis k, active
measurements on axis k-1, 2-axis array ops on distinct active axes, noise
sites visited in increasing order, records written before they are read),
so both the oracle and the device executor can run them.
"""

from __future__ import annotations

import numpy as np

from clifft_amd.program import (FLAG_IDENTITY, Op, U2_DTYPE, U4_DTYPE, build_program, make_apply_pauli,
                                make_array_cnot,
                                make_array_cz, make_array_h, make_array_multi_cnot, make_array_multi_cz,
                                make_array_rot, make_array_s, make_array_s_dag, make_array_swap,
                                make_array_t, make_array_t_dag, make_array_u2, make_array_u4,
                                make_detector, make_exp_val, make_expand, make_expand_rot, make_expand_t,
                                make_expand_t_dag, make_frame_cnot, make_frame_cz, make_frame_h,
                                make_frame_s, make_frame_s_dag, make_frame_swap, make_meas, make_noise,
                                make_noise_block, make_observable, make_postselect, make_readout_noise,
                                make_swap_meas_interfere)


def random_unitary(rng, n):
    z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


def random_phase(rng):
    return np.exp(1j * rng.uniform(0, 2 * np.pi))


def random_u2_node(rng):
    node = np.zeros(1, dtype=U2_DTYPE)[0]
    for s in range(4):
        node["matrices"][s] = random_unitary(rng, 2).reshape(4)
        node["gamma"][s] = random_phase(rng)
        node["out_states"][s] = rng.integers(0, 4)
    return node


def random_u4_node(rng):
    node = np.zeros(1, dtype=U4_DTYPE)[0]
    for s in range(16):
        node["entries"][s]["matrix"] = random_unitary(rng, 4)
        node["entries"][s]["gamma"] = random_phase(rng)
        node["entries"][s]["out_state"] = rng.integers(0, 16)
    return node


def random_pauli_bits(rng, n):
    x = int(rng.integers(0, 2, size=n) @ (1 << np.arange(n, dtype=object))) if n else 0
    z = int(rng.integers(0, 2, size=n) @ (1 << np.arange(n, dtype=object))) if n else 0
    return x, z


def random_program(seed, num_qubits=10, peak=6, length=200, noise=True, postselect=False,
                   exp_vals=True, u_nodes=True, readout=True, dormant_meas=True, p_array=0.5, classical=True,
                   wide_lists=False):
    """A random program of about `length` instructions on `num_qubits`
    qubits whose active rank never exceeds `peak`.  wide_lists: detector and
    observable target lists of 1-8 entries drawn with repetition (targets
    listed twice cancel in the parity)."""
    rng = np.random.default_rng(seed)
    n = num_qubits
    bc = []
    k = 0
    meas_written = []
    n_meas = 0
    pauli_masks, ev_masks, sites, readout_entries = [], [], [], []
    det_lists, obs_lists = [], []
    u2_nodes, u4_nodes = [], []
    n_det = n_obs = n_ev = 0

    def axis_any():
        return int(rng.integers(0, n))

    def pair_any():
        a, b = rng.choice(n, size=2, replace=False)
        return int(a), int(b)

    def pair_active():
        a, b = rng.choice(k, size=2, replace=False)
        return int(a), int(b)

    def new_meas_slot():
        nonlocal n_meas
        n_meas += 1
        return n_meas - 1

    i = 0
    while i < length:
        r = rng.random()
        if r < 0.25:  # frame ops
            c = rng.integers(0, 6)
            if c == 0:
                bc.append(make_frame_cnot(*pair_any()))
            elif c == 1:
                bc.append(make_frame_cz(*pair_any()))
            elif c == 2:
                bc.append(make_frame_h(axis_any()))
            elif c == 3:
                bc.append(make_frame_s(axis_any()))
            elif c == 4:
                bc.append(make_frame_s_dag(axis_any()))
            else:
                bc.append(make_frame_swap(*pair_any()))
        elif r < 0.25 + 0.12 and k < peak and k < n:  # expand
            c = rng.integers(0, 4)
            if c == 0:
                bc.append(make_expand(k))
            elif c == 1:
                bc.append(make_expand_t(k))
            elif c == 2:
                bc.append(make_expand_t_dag(k))
            else:
                ph = random_phase(rng)
                bc.append(make_expand_rot(k, ph.real, ph.imag))
            k += 1
        elif r < 0.25 + 0.12 + p_array and k >= 1:  # array ops
            c = rng.integers(0, 13)
            if c in (0, 1, 2) and k >= 2:
                a, b = pair_active()
                bc.append([make_array_cnot, make_array_cz, make_array_swap][c](a, b))
            elif c in (3, 4) and k >= 2:
                ax = int(rng.integers(0, k))
                mask = 0
                for q in range(k):
                    if q != ax and rng.random() < 0.5:
                        mask |= 1 << q
                # high bits beyond k must be ignored by both executors
                if rng.random() < 0.3:
                    mask |= 1 << int(rng.integers(k, 60))
                    if (mask >> ax) & 1:
                        mask ^= 1 << ax
                bc.append((make_array_multi_cnot if c == 3 else make_array_multi_cz)(ax, mask))
            elif c == 5:
                bc.append(make_array_h(int(rng.integers(0, k))))
            elif c == 6:
                bc.append((make_array_s if rng.random() < 0.5 else make_array_s_dag)(int(rng.integers(0, k))))
            elif c == 7:
                ax = int(rng.integers(0, k)) if rng.random() < 0.8 else axis_any()
                bc.append((make_array_t if rng.random() < 0.5 else make_array_t_dag)(ax))
            elif c == 8:
                ax = int(rng.integers(0, k)) if rng.random() < 0.8 else axis_any()
                ph = random_phase(rng)
                bc.append(make_array_rot(ax, ph.real, ph.imag))
            elif c == 9 and u_nodes:
                ax = int(rng.integers(0, k)) if rng.random() < 0.8 else axis_any()
                u2_nodes.append(random_u2_node(rng))
                bc.append(make_array_u2(ax, len(u2_nodes) - 1))
            elif c == 10 and u_nodes and k >= 2:
                lo, hi = sorted(pair_active())
                u4_nodes.append(random_u4_node(rng))
                bc.append(make_array_u4(lo, hi, len(u4_nodes) - 1))
            elif c in (11, 12):  # active measurement on axis k-1
                slot = new_meas_slot()
                sign = bool(rng.random() < 0.3)
                cc = rng.integers(0, 3)
                if cc == 0:
                    bc.append(make_meas(Op.MEAS_ACTIVE_DIAGONAL, k - 1, slot, sign))
                elif cc == 1:
                    bc.append(make_meas(Op.MEAS_ACTIVE_INTERFERE, k - 1, slot, sign))
                else:
                    f = int(rng.integers(0, k))
                    bc.append(make_swap_meas_interfere(f, k - 1, slot, sign))
                k -= 1
                meas_written.append(slot)
            else:
                continue
        elif r < 0.85 and dormant_meas and k < n:  # dormant measurements
            slot = new_meas_slot()
            ax = int(rng.integers(k, n))
            sign = bool(rng.random() < 0.3)
            if rng.random() < 0.5:
                flags = FLAG_IDENTITY if rng.random() < 0.2 else 0  # record = sign bit
                bc.append(make_meas(Op.MEAS_DORMANT_STATIC, ax, slot, sign, flags=flags))
            else:
                bc.append(make_meas(Op.MEAS_DORMANT_RANDOM, ax, slot, sign))
            meas_written.append(slot)
        elif r < 0.90 and noise:
            nch = int(rng.choice([1, 3, 15]))
            p = 10 ** rng.uniform(-3, -0.5)
            # equal channel probabilities (DEPOLARIZE-like) or unequal ones
            w = np.full(nch, 1.0 / nch) if rng.random() < 0.5 else rng.dirichlet(np.ones(nch))
            site = [(*random_pauli_bits(rng, n), float(p * w[c])) for c in range(nch)]
            if rng.random() < 0.3:  # a block of consecutive sites
                cnt = int(rng.integers(1, 4))
                start = len(sites)
                for _ in range(cnt):
                    sites.append(site)
                bc.append(make_noise_block(start, cnt))
            else:
                sites.append(site)
                bc.append(make_noise(len(sites) - 1))
        elif r < 0.93 and meas_written:
            c = rng.integers(0, 4) if classical else 0
            if c == 0:
                x, z = random_pauli_bits(rng, n)
                pauli_masks.append((x, z, bool(rng.random() < 0.5)))
                bc.append(make_apply_pauli(len(pauli_masks) - 1, int(rng.choice(meas_written))))
            elif c == 1 and readout:
                readout_entries.append((int(rng.choice(meas_written)), float(rng.uniform(0.01, 0.4))))
                bc.append(make_readout_noise(len(readout_entries) - 1))
            elif c == 2:
                if wide_lists:
                    tg = [int(t) for t in rng.choice(meas_written, size=int(rng.integers(1, 9)), replace=True)]
                else:
                    tg = [int(t) for t in rng.choice(meas_written, size=min(3, len(meas_written)), replace=False)]
                det_lists.append(tg)
                if postselect and rng.random() < 0.3:
                    bc.append(make_postselect(len(det_lists) - 1, n_det, bool(rng.random() < 0.5)))
                else:
                    bc.append(make_detector(len(det_lists) - 1, n_det, bool(rng.random() < 0.5)))
                n_det += 1
            else:
                if wide_lists:
                    tg = [int(t) for t in rng.choice(meas_written, size=int(rng.integers(1, 9)), replace=True)]
                else:
                    tg = [int(t) for t in rng.choice(meas_written, size=min(2, len(meas_written)), replace=False)]
                obs_lists.append(tg)
                bc.append(make_observable(len(obs_lists) - 1, int(rng.integers(0, 2))))
                n_obs = max(n_obs, 2)
        elif exp_vals:
            x, z = random_pauli_bits(rng, n)
            if rng.random() < 0.7:  # keep X support active most of the time
                x &= (1 << k) - 1
            ev_masks.append((x, z, bool(rng.random() < 0.3)))
            bc.append(make_exp_val(len(ev_masks) - 1, n_ev))
            n_ev += 1
        else:
            continue
        i += 1

    return build_program(
        bc, peak_rank=peak, num_qubits=n, num_measurements=n_meas, num_detectors=n_det,
        num_observables=n_obs, num_exp_vals=n_ev, pauli_masks=pauli_masks, exp_val_masks=ev_masks,
        noise_sites=sites, readout_noise=readout_entries, detector_targets=det_lists,
        observable_targets=obs_lists,
        u2_nodes=np.array(u2_nodes, dtype=U2_DTYPE) if u2_nodes else None,
        u4_nodes=np.array(u4_nodes, dtype=U4_DTYPE) if u4_nodes else None,
        expected_observables=[1, 0] if n_obs else [])


def random_measure_only_program(seed, num_qubits=8, peak=5, length=120):
    """Unitary gates + measurements only (valid input of
    record_probabilities: no noise, detectors, observables)."""
    return random_program(seed, num_qubits, peak, length, noise=False, exp_vals=True, readout=False,
                          postselect=False, u_nodes=True, classical=False)


def fanout_program(seed, peak, num_qubits=None, frames=True):
    """Expansions to rank `peak`, then runs of ARRAY_CNOTs sharing a control
    (in the global tier the device applies each run in one fused pass), mixed with H/T, noise
    on the frame, EXP_VAL probes and final measurements of every axis."""
    from clifft_amd.program import (Op, build_program, make_array_cnot, make_array_h, make_array_t,
                                    make_exp_val, make_expand_t, make_frame_h, make_meas, make_noise)
    rng = np.random.default_rng(seed)
    n = num_qubits or peak + 2
    bc = []
    for v in range(peak):
        bc += [make_frame_h(v), make_expand_t(v)]
    sites = []
    for _ in range(6):
        c = int(rng.integers(0, peak))
        targets = [int(t) for t in rng.permutation(peak) if t != c][: int(rng.integers(1, max(2, peak)))]
        for t in targets:
            bc.append(make_array_cnot(c, t))
        if rng.random() < 0.3 and targets:
            bc.append(make_array_cnot(c, targets[0]))  # repeated target ends the run
        bc.append(make_array_h(int(rng.integers(0, peak))))
        bc.append(make_array_t(int(rng.integers(0, peak))))
        if frames:
            sites.append([(int(rng.integers(1, 1 << peak)), int(rng.integers(0, 1 << peak)), 0.3)])
            bc.append(make_noise(len(sites) - 1))
    ev_masks = [(int(rng.integers(0, 1 << peak)), int(rng.integers(0, 1 << peak)), False) for _ in range(3)]
    for i in range(3):
        bc.append(make_exp_val(i, i))
    k = peak
    for j in range(peak):
        k -= 1
        bc.append(make_meas(Op.MEAS_ACTIVE_INTERFERE if j % 2 else Op.MEAS_ACTIVE_DIAGONAL, k, j))
    return build_program(bc, peak_rank=peak, num_qubits=n, num_measurements=peak, num_exp_vals=3,
                         exp_val_masks=ev_masks, noise_sites=sites)
