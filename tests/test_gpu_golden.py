"""Device executor vs the committed golden fixtures (tests/golden)."""

from __future__ import annotations

import numpy as np
import pytest

from test_golden import load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["clifford_t20", "surface_d3"])
def test_device_reproduces_records(name):
    from clifft_amd import svm
    _, prog, g = load(name)
    print("CASE golden", name, flush=True)
    shots = len(g["measurements"])
    res = svm.sample(prog, shots, seed=int(g["seed"]), packed=True)
    assert np.array_equal(res.measurements, g["measurements"])
    assert np.array_equal(res.detectors, g["detectors"])
    assert np.array_equal(res.observables, g["observables"])


def test_device_adjoint_reproduces_gradients():
    from clifft_amd import grad
    _, prog, g = load("rz24")
    print("CASE golden rz24", flush=True)
    adj = grad.Adjoint(prog)
    ev, gr = adj.run(g["thetas"], g["coeffs"])
    np.testing.assert_allclose(ev, g["exp_vals"], atol=1e-12, rtol=0)
    np.testing.assert_allclose(gr, g["grads"], atol=1e-10, rtol=0)
