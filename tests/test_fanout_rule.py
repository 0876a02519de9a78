"""The merged fan-out opcode's phase rule, checked by brute force on the CPU.

A kArrayFanout that covers CNOTs from control c to the targets T and the
ARRAY_MULTI_CZ from c over the mask m that follows them is one deferred
permutation/phase opcode (clifft_amd/csrc/svm_interp.h, MONO_FANOUT_CZ):
the sign applies before the permutation x -> x ^ T (where x_c = 1), on
x_c = 1 and parity(x & qm) = 1 with qm = m, toggled at bit c when
parity(T & m) = 1.  On random small states this must equal the CNOTs then
the MULTI_CZ applied one at a time; the rule without the toggle must not.
"""

from __future__ import annotations

import numpy as np


def _par(v: int) -> int:
    return bin(v).count("1") & 1


def _sequential(v, c, T, m):
    K = len(v)
    a = np.zeros(K, complex)
    for x in range(K):
        a[x ^ (T if (x >> c) & 1 else 0)] = v[x]
    for y in range(K):
        if (y >> c) & 1 and _par(y & m):
            a[y] = -a[y]
    return a


def _merged(v, c, T, m, toggle=True):
    qm = m ^ ((1 << c) if toggle and _par(T & m) else 0)
    b = np.zeros(len(v), complex)
    for x in range(len(v)):
        val = -v[x] if (x >> c) & 1 and _par(x & qm) else v[x]
        b[x ^ (T if (x >> c) & 1 else 0)] = val
    return b


def _cases(seed, n):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        k = int(rng.integers(2, 8))
        c = int(rng.integers(0, k))
        masks = []
        for _ in range(2):
            w = 0
            for _ in range(int(rng.integers(0, 4))):
                t = int(rng.integers(0, k))
                if t != c:
                    w ^= 1 << t
            masks.append(w)
        v = rng.normal(size=1 << k) + 1j * rng.normal(size=1 << k)
        yield v, c, masks[0], masks[1]


def test_merged_fanout_cz_equals_the_ops():
    for v, c, T, m in _cases(1, 1500):
        assert np.array_equal(_sequential(v, c, T, m), _merged(v, c, T, m)), (c, T, m)


def test_rule_without_the_toggle_fails():
    bad = sum(not np.array_equal(_sequential(v, c, T, m), _merged(v, c, T, m, toggle=False))
              for v, c, T, m in _cases(2, 400))
    assert bad > 0
