"""Dense statevector simulation of the circuit subset (test helper), used to
pin the compiler end to end: EXP_VAL values and exact record probabilities
of compiled programs must equal the dense physics."""

from __future__ import annotations

import numpy as np

import dense as D


def parse(text):
    ops = []
    for line in text.splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        name, _, rest = line.partition(" ")
        args = []
        if "(" in name:
            name, a = name.split("(")
            args = [float(x) for x in a.rstrip(")").split(",")]
        ops.append((name.upper(), args, rest.split()))
    return ops


def num_qubits(ops):
    n = 0
    for name, _, tg in ops:
        for t in tg:
            if t[0].isdigit():
                n = max(n, int(t) + 1)
            elif name == "EXP_VAL":
                for term in t.split("*"):
                    n = max(n, int(term[1:]) + 1)
    return n


def pauli_product_matrix_apply(psi, term_str, n):
    out = psi
    for term in reversed(term_str.split("*")):
        c, q = term[0].upper(), int(term[1:])
        out = D.apply_1q(out, {"X": D.X, "Y": D.Y, "Z": D.Z}[c], q, n)
    return out


def run(text, record=None):
    """Simulate; returns (state, exp_vals, log_prob of `record`).

    Measurements follow `record` (a list of 0/1 per visible measurement);
    the returned log-probability is the Born probability of that record."""
    ops = parse(text)
    n = num_qubits(ops)
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    evs, logp, mi = [], 0.0, 0
    for name, args, tg in ops:
        qs = [int(t) for t in tg if t[0].isdigit()]
        if name == "H":
            for q in qs:
                psi = D.apply_1q(psi, D.H, q, n)
        elif name in ("S", "SQRT_Z"):
            for q in qs:
                psi = D.apply_1q(psi, D.S, q, n)
        elif name in ("S_DAG", "SQRT_Z_DAG"):
            for q in qs:
                psi = D.apply_1q(psi, D.SDG, q, n)
        elif name in ("X", "Y", "Z"):
            for q in qs:
                psi = D.apply_1q(psi, {"X": D.X, "Y": D.Y, "Z": D.Z}[name], q, n)
        elif name == "T":
            for q in qs:
                psi = D.apply_1q(psi, D.T, q, n)
        elif name == "T_DAG":
            for q in qs:
                psi = D.apply_1q(psi, D.TDG, q, n)
        elif name == "R_Z":
            a = args[0] * np.pi
            U = np.diag([np.exp(-1j * a / 2), np.exp(1j * a / 2)])
            for q in qs:
                psi = D.apply_1q(psi, U, q, n)
        elif name in ("CX", "CNOT", "ZCX"):
            for c, t in zip(qs[::2], qs[1::2]):
                psi = D.cnot(psi, c, t, n)
        elif name == "CZ":
            for a, b in zip(qs[::2], qs[1::2]):
                psi = D.cz(psi, a, b, n)
        elif name == "SWAP":
            for a, b in zip(qs[::2], qs[1::2]):
                psi = D.swap(psi, a, b, n)
        elif name in ("M", "MR"):
            for q in qs:
                m = record[mi] if record is not None else 0
                mi += 1
                proj = D.project(psi, q, "z", m, n)
                p = np.vdot(proj, proj).real
                if p <= 1e-300:
                    return psi, evs, -np.inf
                logp += np.log(p)
                psi = proj / np.sqrt(p)
                if name == "MR" and m == 1:
                    psi = D.apply_1q(psi, D.X, q, n)
        elif name == "EXP_VAL":
            for t in tg:
                evs.append(np.vdot(psi, pauli_product_matrix_apply(psi, t, n)).real)
        elif name == "TICK":
            pass
        else:
            raise ValueError(name)
    return psi, evs, logp


def random_circuit(rng, n, depth, t_count=3, measure=False, rz=0, exp_vals=0, mr=False):
    lines = []
    slots = [(l, q) for l in range(depth) for q in range(n)]
    t_slots = set(map(tuple, [slots[i] for i in rng.choice(len(slots), size=min(t_count, len(slots)), replace=False)]))
    rz_slots = set(map(tuple, [slots[i] for i in rng.choice(len(slots), size=min(rz, len(slots)), replace=False)]))
    for layer in range(depth):
        for q in range(n):
            g = rng.choice(["H", "S", "S_DAG", "X", "Z", "I"])
            if g != "I":
                lines.append(f"{g} {q}")
            if (layer, q) in t_slots:
                lines.append(f"{'T' if rng.random() < 0.6 else 'T_DAG'} {q}")
            if (layer, q) in rz_slots:
                lines.append(f"R_Z({rng.uniform(-1, 1):.6f}) {q}")
        perm = rng.permutation(n)
        for a, b in zip(perm[::2], perm[1::2]):
            lines.append(f"{rng.choice(['CX', 'CZ'])} {a} {b}")
        if measure and rng.random() < 0.3:
            q = int(rng.integers(0, n))
            lines.append(f"{'MR' if (mr and rng.random() < 0.5) else 'M'} {q}")
    for _ in range(exp_vals):
        terms = [f"{rng.choice(list('XYZ'))}{q}" for q in range(n) if rng.random() < 0.6]
        if terms:
            lines.append("EXP_VAL " + "*".join(terms))
    if measure:
        lines.append("M " + " ".join(str(q) for q in range(n)))
    return "\n".join(lines)
