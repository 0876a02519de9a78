"""Device adjoint gradients vs the oracle's parameter-shift gradients and
expectation values, central finite differences of the device forward pass,
and the torch.autograd.Function around the kernel.

Tolerances: expectation values 1e-12 absolute (same algorithm, rounding
only); adjoint vs parameter-shift 1e-10 (both exact up to rounding, the
shift takes 2P+1 separate evaluations); vs central differences with
h = 1e-6, 1e-7 (truncation h^2 plus cancellation eps/h)."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from gradprog import oracle_exp_vals, oracle_param_shift, rz_circuit

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    from clifft_amd import _lib, grad
    from clifft_amd.compiler import compile as cc
    assert _lib.backend() == "hip:gfx950", _lib.backend()
    return grad, cc


CASES = [  # (seed, n, layers, n_params, repeat_qubits, fusion)
    (0, 8, 4, 6, True, True), (1, 8, 4, 6, True, False), (2, 8, 4, 6, True, True), (4, 10, 5, 8, True, True),
    (6, 9, 4, 9, False, True), (7, 12, 3, 12, False, True), (8, 14, 3, 14, False, True)]


@pytest.mark.parametrize("case", CASES)
def test_adjoint_matches_oracle(mods, orc, case):
    grad, cc = mods
    seed, n, layers, npar, rep, fusion = case
    announce("adjoint", case)
    build, coeffs = rz_circuit(n=n, layers=layers, n_params=npar, seed=seed, repeat_qubits=rep)
    rng = np.random.default_rng(seed + 100)
    thetas = rng.uniform(0, 2 * np.pi, (3, npar))
    p = cc(build(thetas[0]), keep_rotations=True, fusion=fusion)
    adj = grad.Adjoint(p)
    ev, g = adj.run(thetas, coeffs)
    for b in range(len(thetas)):
        np.testing.assert_allclose(ev[b], oracle_exp_vals(orc, p, thetas[b]), atol=1e-12, rtol=0)
        np.testing.assert_allclose(g[b], oracle_param_shift(orc, p, thetas[b], coeffs), atol=1e-10, rtol=0)
    assert np.abs(g).max() > 1e-3


def test_adjoint_vs_central_fd(mods):
    grad, cc = mods
    announce("adjoint-fd")
    build, coeffs = rz_circuit(n=8, layers=4, n_params=6, seed=2)
    th = np.random.default_rng(9).uniform(0, 2 * np.pi, 6)
    p = cc(build(th), keep_rotations=True)
    adj = grad.Adjoint(p)
    _, g = adj.run(th[None], coeffs)
    h = 1e-6
    shifted = np.repeat(th[None], 2 * len(th), axis=0)
    for i in range(len(th)):
        shifted[2 * i, i] += h
        shifted[2 * i + 1, i] -= h
    ev, _ = adj.run(shifted, grads=False)
    F = ev @ coeffs
    fd = (F[0::2] - F[1::2]) / (2 * h)
    np.testing.assert_allclose(g[0], fd, atol=1e-7, rtol=0)


def test_autograd_function(mods):
    import torch
    grad, cc = mods
    announce("autograd")
    build, coeffs = rz_circuit(n=8, layers=4, n_params=6, seed=1)
    rng = np.random.default_rng(5)
    th_np = rng.uniform(0, 2 * np.pi, (4, 6))
    p = cc(build(th_np[0]), keep_rotations=True)
    adj = grad.Adjoint(p)
    ev_ref, g_ref = adj.run(th_np, coeffs)
    th = torch.tensor(th_np, dtype=torch.float64, device="cuda:0", requires_grad=True)
    c = torch.tensor(coeffs, dtype=torch.float64, device="cuda:0")
    F = grad.expectation(adj, th, c)
    w = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64, device="cuda:0")
    (F * w).sum().backward()
    np.testing.assert_allclose(F.detach().cpu().numpy(), ev_ref @ coeffs, atol=1e-13, rtol=0)
    np.testing.assert_allclose(th.grad.cpu().numpy(), g_ref * w.cpu().numpy()[:, None], atol=1e-13, rtol=0)
    th2 = torch.tensor(th_np[:1], dtype=torch.float64, device="cuda:0", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: grad.expectation(adj, t, c), (th2,), eps=1e-6, atol=1e-6)
    # dF/dcoeffs = exp_vals, alone and together with dF/dthetas
    c2 = torch.tensor(coeffs, dtype=torch.float64, device="cuda:0", requires_grad=True)
    th3 = torch.tensor(th_np, dtype=torch.float64, device="cuda:0", requires_grad=True)
    (grad.expectation(adj, th3, c2) * w).sum().backward()
    np.testing.assert_allclose(c2.grad.cpu().numpy(), w.cpu().numpy() @ ev_ref, atol=1e-13, rtol=0)
    np.testing.assert_allclose(th3.grad.cpu().numpy(), g_ref * w.cpu().numpy()[:, None], atol=1e-13, rtol=0)
    c3 = torch.tensor(coeffs, dtype=torch.float64, device="cuda:0", requires_grad=True)
    assert torch.autograd.gradcheck(lambda cc_: grad.expectation(adj, th2.detach(), cc_), (c3,), eps=1e-6,
                                    atol=1e-8)
    # shapes the kernel would read past are refused before any launch
    with pytest.raises(ValueError):
        grad.expectation(adj, th[0].detach(), c)
    with pytest.raises(ValueError):
        grad.expectation(adj, th[:, :5].detach(), c)
    with pytest.raises(ValueError):
        grad.expectation(adj, th.detach(), c[:-1])


def test_workload_rz24_matches_oracle(mods, orc):
    """The 24-qubit, 16-parameter, 32-probe workload (peak rank 16) at two
    theta vectors against the oracle's parameter-shift gradient."""
    grad, cc = mods
    from clifft_amd import workloads as W
    announce("rz24")
    build, coeffs, _ = W.rz_exp_val()
    th = np.random.default_rng(42).uniform(0, 2 * np.pi, (2, 16))
    p = cc(build(th[0]), keep_rotations=True)
    adj = grad.Adjoint(p)
    ev, g = adj.run(th, coeffs)
    for b in range(2):
        np.testing.assert_allclose(ev[b], oracle_exp_vals(orc, p, th[b]), atol=1e-12, rtol=0)
        np.testing.assert_allclose(g[b], oracle_param_shift(orc, p, th[b], coeffs), atol=1e-10, rtol=0)
    assert np.all(np.abs(ev) > 1e-9) and np.abs(g).max() > 1e-2  # the probes are not identically zero


def test_adjoint_rejects_measurements(mods):
    grad, cc = mods
    from clifft_amd._lib import SvmError
    p = cc("H 0\nR_Z(0.3) 0\nM 0\nEXP_VAL X0", keep_rotations=True)
    with pytest.raises(SvmError):
        grad.Adjoint(p)
