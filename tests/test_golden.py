"""Golden fixtures (tests/golden, written by make_golden.py): the compiler
reproduces the committed Program blobs from the committed circuit text, and
the oracle reproduces the committed records, expectation values and
parameter-shift gradients."""

from __future__ import annotations

import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["clifford_t20", "surface_d3", "rz24"]


def load(name):
    from clifft_amd.program import Program
    with open(os.path.join(GOLDEN, name + ".stim")) as f:
        text = f.read()
    with open(os.path.join(GOLDEN, name + ".prog"), "rb") as f:
        prog = Program.from_bytes(f.read())
    return text, prog, np.load(os.path.join(GOLDEN, name + ".npz"))


@pytest.mark.parametrize("name", NAMES)
def test_compiler_reproduces_blob(name):
    from clifft_amd.compiler import compile as cc
    text, prog, _ = load(name)
    kw = {"keep_rotations": True} if name == "rz24" else {}
    assert cc(text, **kw).to_bytes() == prog.to_bytes()


@pytest.mark.parametrize("name", ["clifford_t20", "surface_d3"])
def test_oracle_reproduces_records(orc, name):
    _, prog, g = load(name)
    shots = len(g["measurements"])
    m, d, o, _, _ = orc.sample(prog, shots, seed=int(g["seed"]), per_shot=True)
    assert np.array_equal(np.packbits(m, axis=1, bitorder="little"), g["measurements"])
    assert np.array_equal(np.packbits(d, axis=1, bitorder="little"), g["detectors"])
    assert np.array_equal(np.packbits(o, axis=1, bitorder="little"), g["observables"])
    f = orc.sample(prog, 1, seed=int(g["seed"]), per_shot=False)
    assert np.array_equal(np.concatenate([f[0][0], f[1][0], f[2][0]]), g["single_stream_first"])


def test_oracle_reproduces_gradients(orc):
    from gradprog import oracle_exp_vals, oracle_param_shift
    _, prog, g = load("rz24")
    for b in range(len(g["thetas"])):
        np.testing.assert_allclose(oracle_exp_vals(orc, prog, g["thetas"][b]), g["exp_vals"][b], atol=1e-13, rtol=0)
        np.testing.assert_allclose(oracle_param_shift(orc, prog, g["thetas"][b], g["coeffs"]), g["grads"][b],
                                   atol=1e-12, rtol=0)
    assert np.all(np.abs(g["exp_vals"]) > 1e-9) and np.abs(g["grads"]).max() > 1e-2
