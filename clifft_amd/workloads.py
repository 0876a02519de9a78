"""Synthetic benchmark circuits (seeded), one generator per workload.

    random_clifford_t   20 qubits x 40 layers, H/S per qubit + CNOT perfect
                        matching, T gates on random (layer, qubit) slots,
                        final M (the 1e5-shot plain-VM workload)
    surface_code_d3     distance-3 rotated surface-code memory, 9 data +
                        8 ancilla, 3 rounds x 4 CNOT steps, DEPOLARIZE1/2 and
                        X_ERROR at p, detectors on consecutive rounds, one
                        injected T per round (the 1e6-shot noisy workload)
    large_k             34 qubits, C0-style layers, 26 T on distinct qubits,
                        MR on 4 random qubits every 10 layers
    batch_programs      many random_clifford_t programs, n ~ U[16,40],
                        T-count ~ U[4,16]
    rz_exp_val          24 qubits, 30 Clifford layers interleaved with 16
                        R_Z(theta/pi) rotations, 32 random Pauli-string probes
                        (drawn among the strings with a nonzero expectation)
"""

from __future__ import annotations

import numpy as np


def random_clifford_t(n=20, layers=40, t_count=6, seed=42, final_measure=True, mr_every=0, mr_count=0,
                      t_distinct_qubits=False):
    rng = np.random.default_rng(seed)
    slots = [(l, q) for l in range(layers) for q in range(n)]
    if t_distinct_qubits:
        qubits = rng.choice(n, size=t_count, replace=False)
        t_slots = {(int(rng.integers(0, layers)), int(q)) for q in qubits}
    else:
        idx = rng.choice(len(slots), size=t_count, replace=False)
        t_slots = {slots[i] for i in idx}
    lines = []
    for layer in range(layers):
        ones = rng.random(n) < 0.5
        lines.append("H " + " ".join(str(q) for q in range(n) if ones[q]))
        lines.append("S " + " ".join(str(q) for q in range(n) if not ones[q]))
        ts = [q for q in range(n) if (layer, q) in t_slots]
        if ts:
            lines.append("T " + " ".join(map(str, ts)))
        perm = rng.permutation(n)
        lines.append("CX " + " ".join(f"{a} {b}" for a, b in zip(perm[::2], perm[1::2])))
        if mr_every and (layer + 1) % mr_every == 0:
            qs = rng.choice(n, size=mr_count, replace=False)
            lines.append("MR " + " ".join(map(str, sorted(qs))))
    if final_measure:
        lines.append("M " + " ".join(map(str, range(n))))
    return "\n".join(l for l in lines if l.split()[1:] or l.startswith("M"))


def surface_code_d3(rounds=3, p=1e-3, inject_t=True, seed=42):
    """Rotated d=3 memory experiment in the Z basis.

    Data qubit (r, c) -> 3r + c; ancillas 9..16.  X checks use the N-shaped
    CNOT order (NW, NE, SW, SE), Z checks the Z-shaped order (NW, SW, NE, SE).
    Z_L = Z on data row 0."""
    rng = np.random.default_rng(seed)
    d = lambda r, c: 3 * r + c
    faces = []  # (type, corners dict)
    for r in range(-1, 3):
        for c in range(-1, 3):
            corners = {k: (r + dr, c + dc) for k, (dr, dc) in
                       {"NW": (0, 0), "NE": (0, 1), "SW": (1, 0), "SE": (1, 1)}.items()}
            corners = {k: d(*v) for k, v in corners.items() if 0 <= v[0] < 3 and 0 <= v[1] < 3}
            typ = "X" if (r + c) % 2 == 0 else "Z"
            interior = 0 <= r < 2 and 0 <= c < 2
            if interior:
                faces.append((typ, corners))
            elif len(corners) == 2:
                if typ == "X" and r in (-1, 2):
                    faces.append((typ, corners))
                if typ == "Z" and c in (-1, 2):
                    faces.append((typ, corners))
    assert len(faces) == 8 and sum(f[0] == "X" for f in faces) == 4
    anc = {i: 9 + i for i in range(8)}
    xa = [anc[i] for i, f in enumerate(faces) if f[0] == "X"]
    order = {"X": ["NW", "NE", "SW", "SE"], "Z": ["NW", "SW", "NE", "SE"]}
    data = list(range(9))
    L = []

    def dep1(qs):
        L.append(f"DEPOLARIZE1({p}) " + " ".join(map(str, qs)))

    for rnd in range(rounds):
        if inject_t:
            q = int(rng.integers(0, 9))
            L.append(f"T {q}")
            dep1([q])
        L.append("H " + " ".join(map(str, xa)))
        dep1(xa)
        for step in range(4):
            pairs = []
            for i, (typ, corners) in enumerate(faces):
                key = order[typ][step]
                if key not in corners:
                    continue
                dq = corners[key]
                pairs += [anc[i], dq] if typ == "X" else [dq, anc[i]]
            L.append("CX " + " ".join(map(str, pairs)))
            L.append(f"DEPOLARIZE2({p}) " + " ".join(map(str, pairs)))
        L.append("H " + " ".join(map(str, xa)))
        dep1(xa)
        ancs = [anc[i] for i in range(8)]
        L.append(f"X_ERROR({p}) " + " ".join(map(str, ancs)))
        L.append("MR " + " ".join(map(str, ancs)))
        for i, (typ, _) in enumerate(faces):
            back = 8 - i
            if rnd == 0:
                if typ == "Z":
                    L.append(f"DETECTOR rec[-{back}]")
            else:
                L.append(f"DETECTOR rec[-{back}] rec[-{back + 8}]")
    L.append(f"X_ERROR({p}) " + " ".join(map(str, data)))
    L.append("M " + " ".join(map(str, data)))
    # final Z checks: data parity vs the last ancilla round
    for i, (typ, corners) in enumerate(faces):
        if typ != "Z":
            continue
        recs = [f"rec[-{9 - q}]" for q in sorted(corners.values())]
        recs.append(f"rec[-{9 + 8 - i}]")
        L.append("DETECTOR " + " ".join(recs))
    L.append("OBSERVABLE_INCLUDE(0) " + " ".join(f"rec[-{9 - q}]" for q in (0, 1, 2)))
    return "\n".join(L)


def large_k(n=34, layers=40, t_count=26, seed=42):
    return random_clifford_t(n=n, layers=layers, t_count=t_count, seed=seed, final_measure=False, mr_every=10,
                             mr_count=4, t_distinct_qubits=True)


def batch_programs(count=4096, seed=42, shots=1000):
    """(circuit text, shots) for `count` random_clifford_t programs with
    n ~ U[16, 40] and T-count ~ U[4, 16]; program i uses seed `seed ^ i`."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(16, 41))
        t = int(rng.integers(4, 17))
        out.append((random_clifford_t(n=n, layers=40, t_count=t, seed=seed ^ i), shots))
    return out


def rz_exp_val(n=24, layers=30, n_params=16, n_terms=32, seed=42, nonvanishing=True):
    """Returns (circuit_text_builder(thetas) -> text, coeffs, pauli_terms).

    The circuit has n_params R_Z slots whose angles are filled in by the
    builder (theta in radians -> R_Z(theta / pi) in half-turns); the probes
    are `n_terms` random Pauli strings (each qubit I/X/Y/Z uniformly), with
    coefficients ~ N(0, 1) for the observable sum.

    nonvanishing: draw the strings uniformly among those whose expectation
    is not identically zero in theta.  After the Clifford layers the state
    is a product of the n_params rotated axes (each (|0> + z|1>)/sqrt 2 in
    the compiled frame) and a stabilizer state on the dormant axes, so a
    string is kept when its compiled image has X or Y (not Z alone) on the
    active axes it touches, at least one of them, and no X support on a
    dormant axis.  An unconditioned draw has <P> = 0 exactly for every probe
    of the seed-42 draw at n = 24, which would make the workload vacuous.
    The circuit-frame string of a chosen image is found by inverting the
    compiled images of X_q and Z_q over GF(2)."""
    rng = np.random.default_rng(seed)
    param_layers = sorted(rng.choice(layers, size=n_params, replace=n_params > layers))
    param_qubits = [int(rng.integers(0, n)) for _ in range(n_params)]
    layer_ops = []
    for layer in range(layers):
        ones = rng.random(n) < 0.5
        perm = rng.permutation(n)
        layer_ops.append((ones, perm))

    def body(thetas):
        thetas = np.asarray(thetas, dtype=np.float64)
        lines = []
        pi_idx = 0
        for layer, (ones, perm) in enumerate(layer_ops):
            lines.append("H " + " ".join(str(q) for q in range(n) if ones[q]))
            lines.append("S " + " ".join(str(q) for q in range(n) if not ones[q]))
            while pi_idx < n_params and param_layers[pi_idx] == layer:
                lines.append(f"R_Z({float(thetas[pi_idx] / np.pi):.17g}) {param_qubits[pi_idx]}")
                pi_idx += 1
            lines.append("CX " + " ".join(f"{a} {b}" for a, b in zip(perm[::2], perm[1::2])))
        return [l for l in lines if l.split()[1:]]

    def draw(count):
        out = []
        while len(out) < count:
            letters = rng.integers(0, 4, size=n)
            t = "*".join(f"{'IXYZ'[l]}{q}" for q, l in enumerate(letters) if l)
            if t:
                out.append(t)
        return out

    if nonvanishing:
        terms = _nonvanishing_terms(body(np.full(n_params, 0.25)), n, n_params, n_terms, rng)
    else:
        terms = draw(n_terms)
    coeffs = rng.normal(size=n_terms)

    def build(thetas):
        return "\n".join(body(thetas) + ["EXP_VAL " + " ".join(terms)])

    return build, coeffs, terms


def _gf2_solve(M, t):
    """x with M x = t over GF(2) (M square, invertible; uint8 0/1)."""
    n = M.shape[0]
    A = np.concatenate([M.copy(), t.reshape(-1, 1)], axis=1) & 1
    for c in range(n):
        piv = c + int(np.flatnonzero(A[c:, c])[0])
        A[[c, piv]] = A[[piv, c]]
        rows = np.flatnonzero(A[:, c])
        rows = rows[rows != c]
        A[rows] ^= A[c]
    return A[:, n]


def _nonvanishing_terms(lines, n, k, count, rng):
    """`count` circuit-frame Pauli strings whose compiled images act as X/Y
    (or I) on the k active axes, touching at least one, and as Z/I on the
    dormant ones."""
    from .compiler import compile as ccompile
    basis = [f"{c}{q}" for q in range(n) for c in "XZ"]
    p = ccompile("\n".join(lines + ["EXP_VAL " + " ".join(basis)]), keep_rotations=True)
    w = p.words
    ex = np.asarray(p.ev_x).reshape(-1, w)
    ez = np.asarray(p.ev_z).reshape(-1, w)
    order = [int(ins["idx0"]) for ins in p.bytecode if int(ins["opcode"]) == _op_exp_val()]
    if len(order) != 2 * n:
        raise RuntimeError("basis probes changed in compilation")

    def bits(words):
        v = sum(int(x) << (64 * i) for i, x in enumerate(words))
        return [(v >> a) & 1 for a in range(n)]

    M = np.zeros((2 * n, 2 * n), np.uint8)
    for col, mi in enumerate(order):
        M[:n, col] = bits(ex[mi])
        M[n:, col] = bits(ez[mi])
    terms = []
    while len(terms) < count:
        act = rng.integers(0, 3, size=k)  # 0 = I, 1 = X, 2 = Y on the active axes
        if not act.any():
            continue
        tx = np.zeros(n, np.uint8)
        tz = np.zeros(n, np.uint8)
        tx[:k] = act > 0
        tz[:k] = act == 2
        tz[k:] = rng.integers(0, 2, size=n - k)
        sol = _gf2_solve(M, np.concatenate([tx, tz]))
        letters = ["IXZY"[int(sol[2 * q]) + 2 * int(sol[2 * q + 1])] for q in range(n)]
        terms.append("*".join(f"{c}{q}" for q, c in enumerate(letters) if c != "I"))
    return terms


def _op_exp_val():
    from .program import Op
    return Op.EXP_VAL


def _active_x_only(lines, terms):
    """For each Pauli string, whether its compiled EXP_VAL mask has X support
    only on axes active at the probe (static rank from the EXPAND count)."""
    from .compiler import compile as ccompile
    from .program import Op
    p = ccompile("\n".join(lines + ["EXP_VAL " + " ".join(terms)]), keep_rotations=True)
    bc = p.bytecode
    expands = (Op.EXPAND, Op.EXPAND_T, Op.EXPAND_T_DAG, Op.EXPAND_ROT)
    k, ok = 0, []
    for ins in bc:
        op = int(ins["opcode"])
        if op in expands:
            k += 1
        elif op == Op.EXP_VAL:
            words = np.asarray(p.ev_x).reshape(-1, p.words)[int(ins["idx0"])]
            x = sum(int(v) << (64 * i) for i, v in enumerate(words))
            ok.append((x >> k) == 0)
        elif op in (Op.MEAS_ACTIVE_DIAGONAL, Op.MEAS_ACTIVE_INTERFERE, Op.SWAP_MEAS_INTERFERE):
            k -= 1
    if len(ok) != len(terms):
        raise RuntimeError("probe count changed in compilation")
    return ok
