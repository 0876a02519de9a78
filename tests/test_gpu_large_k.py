"""The large-rank workload on the device against the oracle, shot by shot.

34 qubits of random Clifford layers with 26 T gates on distinct qubits and
MR on 4 random qubits every 10 layers compile to peak rank 25 (2^25
complex128 amplitudes, 512 MiB per shot in flight): 26 EXPAND_T run as
fused expansion passes in HBM.  The variant with a final M on every qubit
also runs 25 active measurements from rank 25 down, so the probability
reductions and the compaction sweeps run at rank >= 20.  Records are
bit-exact against the oracle's per-shot streams.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def svm():
    import clifft_amd.svm as svm
    from clifft_amd import _lib
    assert _lib.backend() == "hip:gfx950", _lib.backend()
    return svm


def large_k_program(final_measure):
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile
    return ccompile(W.random_clifford_t(n=34, layers=40, t_count=26, seed=42, final_measure=final_measure,
                                        mr_every=10, mr_count=4, t_distinct_qubits=True))


@pytest.mark.parametrize("final_measure,shots", [(False, 4), (True, 2)])
def test_large_k34_bit_exact(svm, orc, final_measure, shots):
    announce("large-k34", final_measure, shots)
    p = large_k_program(final_measure)
    h = p.active_k_history()
    assert p.peak_rank == 25 and h.max() == 25
    if final_measure:  # active measurements at rank >= 20
        from clifft_amd.program import Op
        act = (Op.MEAS_ACTIVE_DIAGONAL, Op.MEAS_ACTIVE_INTERFERE, Op.SWAP_MEAS_INTERFERE)
        assert sum(1 for i, op in enumerate(p.bytecode["opcode"]) if op in act and h[i] >= 20) >= 5
    res = svm.sample(p, shots, seed=42)
    m, d, o, e, _ = orc.sample(p, shots, seed=42, per_shot=True)
    assert m.shape == (shots, p.num_measurements)
    assert np.array_equal(res.measurements, m)
    assert np.array_equal(res.detectors, d)
    assert np.array_equal(res.observables, o)


def test_large_k34_graph_replay(svm, orc):
    """Per-program HIP graph replay of the large-rank workload: a
    one-program batch captured once and replayed with new seeds, with eager
    work between replays; every replay equals an eager launch with the same
    seed, and the seed-42 launch the oracle's per-shot records."""
    import torch
    from clifft_amd import batch
    announce("large-k34-graph")
    p = large_k_program(True)
    shots = 2
    m = torch.zeros((shots, (p.num_measurements + 7) // 8), dtype=torch.uint8, device="cuda:0")
    b = batch.Batch([p], [shots], [{"measurements": m.data_ptr()}], packed=True)
    ref = np.packbits(orc.sample(p, shots, seed=42, per_shot=True)[0], axis=1, bitorder="little")
    for seed in (42, 7, 9):
        b.run([seed])
        torch.cuda.synchronize()
        eager = m.cpu().numpy().copy()
        m.fill_(255)
        torch.cuda.synchronize()
        b.run_graph([seed])
        torch.cuda.synchronize()
        assert np.array_equal(m.cpu().numpy(), eager), seed
        if seed == 42:
            assert np.array_equal(eager, ref)
        _ = torch.ones(1000, device="cuda:0").sum().item()  # unrelated eager work between replays
    b.close()
