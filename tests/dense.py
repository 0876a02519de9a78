"""Dense 2^n statevector model of the VM's factored state (test helper).

The VM state (src/clifft/svm/svm.h:94-104) is |psi> = gamma * P * (phi (x)
|0..0>_dormant) on virtual axes, with P = prod_q X_q^{x_q} Z_q^{z_q} (Z
applied first).  Virtual axis q is bit q of the basis index.  These helpers
expand that state and apply textbook gate matrices, so opcode semantics can
be checked against linear algebra rather than against a restatement of the
opcode itself.
"""

from __future__ import annotations

import numpy as np

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1, -1]).astype(complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
S = np.diag([1, 1j])
SDG = np.diag([1, -1j])
T = np.diag([1, np.exp(1j * np.pi / 4)])
TDG = np.diag([1, np.exp(-1j * np.pi / 4)])


def apply_1q(psi, U, q, n):
    t = psi.reshape([2] * n)  # axis (n-1-q) is qubit q (C order, bit q of index)
    ax = n - 1 - q
    t = np.moveaxis(np.tensordot(U, t, axes=([1], [ax])), 0, ax)
    return t.reshape(-1)


def apply_2q(psi, U4, lo, hi, n):
    """U4 in the basis |b_hi b_lo> with index (b_hi << 1) | b_lo."""
    t = psi.reshape([2] * n)
    a_lo, a_hi = n - 1 - lo, n - 1 - hi
    U = U4.reshape(2, 2, 2, 2)  # [hi_out, lo_out, hi_in, lo_in]
    t = np.tensordot(U, t, axes=([2, 3], [a_hi, a_lo]))
    t = np.moveaxis(t, [0, 1], [a_hi, a_lo])
    return t.reshape(-1)


def cnot(psi, c, t, n):
    idx = np.arange(1 << n)
    src = np.where((idx >> c) & 1, idx ^ (1 << t), idx)
    return psi[src]


def cz(psi, a, b, n):
    idx = np.arange(1 << n)
    sgn = np.where(((idx >> a) & 1) & ((idx >> b) & 1), -1.0, 1.0)
    return psi * sgn


def swap(psi, a, b, n):
    idx = np.arange(1 << n)
    ba, bb = (idx >> a) & 1, (idx >> b) & 1
    src = idx ^ ((ba ^ bb) << a) ^ ((ba ^ bb) << b)
    return psi[src]


def pauli(psi, xbits, zbits, n):
    """Apply X^x Z^z (Z first)."""
    idx = np.arange(1 << n)
    zpar = np.array([bin(i & zbits).count("1") & 1 for i in range(1 << n)])
    psi = psi * np.where(zpar, -1.0, 1.0)
    return psi[idx ^ xbits]


def hermitian_pauli(psi, xbits, zbits, sign, n):
    """(-1)^sign * i^{|x & z|} X^x Z^z  (Y = iXZ convention)."""
    ph = (1j) ** (bin(xbits & zbits).count("1") % 4) * (-1 if sign else 1)
    return ph * pauli(psi, xbits, zbits, n)


def expand(phi, k, px, pz, gamma, n):
    """Dense |psi> of the factored VM state (active rank k)."""
    psi = np.zeros(1 << n, dtype=complex)
    psi[: 1 << k] = phi[: 1 << k]
    return gamma * pauli(psi, int(px), int(pz), n)


def state_of(st, n):
    """Dense vector of an oracle/device State object."""
    return expand(np.asarray(st.v), st.active_k, int(st.p_x[0]), int(st.p_z[0]), st.gamma, n)


def project(psi, q, basis, m, n):
    """Projector of qubit q onto outcome m of Z (basis='z') or X (basis='x')."""
    if basis == "z":
        idx = np.arange(1 << n)
        keep = ((idx >> q) & 1) == m
        return np.where(keep, psi, 0)
    plus = apply_1q(psi, H, q, n)
    plus = project(plus, q, "z", m, n)
    return apply_1q(plus, H, q, n)


def random_state(rng, k):
    v = rng.normal(size=1 << k) + 1j * rng.normal(size=1 << k)
    return v / np.linalg.norm(v)
