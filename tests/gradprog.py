"""Helpers for gradient tests: parameterised circuits and the oracle's
parameter-shift / finite-difference gradients (CPU)."""

from __future__ import annotations

import numpy as np


def rz_circuit(n=8, layers=4, n_params=6, n_terms=12, t_count=3, seed=0, repeat_qubits=True, probe_width=2):
    """Random Clifford layers with T gates and R_Z parameter slots; R_Z may
    hit already-active qubits (ARRAY_ROT) or fresh ones (EXPAND_ROT).
    Returns (build(thetas) -> text, coeffs)."""
    rng = np.random.default_rng(seed)
    ops = []
    slots = sorted(rng.choice(layers * n, size=n_params + t_count, replace=False))
    pq = [int(rng.integers(0, max(2, n // 2))) if repeat_qubits else int(q % n) for q in slots[:n_params]]
    tq = [int(s % n) for s in slots[n_params:]]
    tl = [int(s // n) for s in slots[n_params:]]
    pl = [int(s // n) for s in slots[:n_params]]
    for layer in range(layers):
        ones = rng.random(n) < 0.5
        perm = rng.permutation(n)
        ops.append(("H", [q for q in range(n) if ones[q]]))
        ops.append(("S", [q for q in range(n) if not ones[q]]))
        for q, l in zip(tq, tl):
            if l == layer:
                ops.append(("T", [q]))
        for i, (q, l) in enumerate(zip(pq, pl)):
            if l == layer:
                ops.append(("RZ", i, q))
        ops.append(("CX", [int(x) for x in perm[: 2 * (n // 2)]]))
    coeffs = rng.normal(size=n_terms)

    def body(thetas):
        lines = []
        for op in ops:
            if op[0] == "RZ":
                lines.append(f"R_Z({float(thetas[op[1]] / np.pi):.17g}) {op[2]}")
            elif op[1]:
                lines.append(op[0] + " " + " ".join(map(str, op[1])))
        return lines

    terms = sensitive_terms(lambda th: "\n".join(body(th)), n, n_params, n_terms, probe_width, rng)

    def build(thetas):
        return "\n".join(body(thetas) + ["EXP_VAL " + " ".join(terms)])

    return build, coeffs


def sensitive_terms(text_of, n, n_params, count, width, rng):
    """Pauli strings of weight <= width whose dense-model expectation moves
    when a parameter is shifted (random strings on a Clifford-dominated state
    are mostly exactly 0, which would make gradient checks vacuous)."""
    from itertools import combinations, product

    from dense_circuit import pauli_product_matrix_apply, run
    th = rng.uniform(0, 2 * np.pi, n_params)
    states = [run(text_of(th))[0]]
    for i in range(n_params):
        t = th.copy()
        t[i] += np.pi / 2
        states.append(run(text_of(t))[0])
    m = 1 << n
    good = []
    for w in range(1, n + 1):
        for qs in combinations(range(n), w):
            for letters in product("XYZ", repeat=w):
                t = "*".join(f"{c}{q}" for c, q in zip(letters, qs))
                vals = [np.vdot(psi[:m], pauli_product_matrix_apply(psi[:m], t, n)).real for psi in states]
                if max(abs(v - vals[0]) for v in vals[1:]) > 1e-2:
                    good.append(t)
        if w >= width and len(good) >= count:
            break
    idx = rng.choice(len(good), size=count, replace=len(good) < count)
    return [good[i] for i in idx]


def with_thetas(prog, thetas):
    """Copy of `prog` with each parameter slot's phase set to
    exp(i * param_sign * theta) (what recompiling with those angles emits)."""
    p = prog.copy()
    bc = p.bytecode.copy()
    for i, pc in enumerate(p.param_instr):
        a = float(p.param_sign[i]) * float(thetas[i])
        bc[pc]["weight_re"] = np.cos(a)
        bc[pc]["weight_im"] = np.sin(a)
    p.bytecode = bc
    return p


def oracle_exp_vals(orc, prog, thetas):
    _, _, _, e, _ = orc.sample(with_thetas(prog, thetas), 1, seed=0)
    return e[0]


def oracle_objective(orc, prog, thetas, coeffs):
    return float(oracle_exp_vals(orc, prog, thetas) @ coeffs)


def oracle_param_shift(orc, prog, thetas, coeffs):
    g = np.zeros(len(thetas))
    for i in range(len(thetas)):
        tp = np.array(thetas, dtype=np.float64)
        tm = tp.copy()
        tp[i] += np.pi / 2
        tm[i] -= np.pi / 2
        g[i] = 0.5 * (oracle_objective(orc, prog, tp, coeffs) - oracle_objective(orc, prog, tm, coeffs))
    return g
