"""Batched execution vs per-program launches and the oracle.

Every program of a batch must produce the records `svm.sample(program,
shots, seed)` produces (the per-shot stream of (seed, shot) is the same; the
team shape, and so the reduction order, can differ between the two paths,
which moves expectation values by rounding only).  The batch runs one
persistent launch per launch group (programs sharing a kernel shape), each
draining its own device queue.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from randprog import random_program

pytestmark = pytest.mark.gpu

ATOL = 1e-12


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    import clifft_amd.batch as batch
    import clifft_amd.svm as svm
    from clifft_amd import _lib
    assert _lib.backend() == "hip:gfx950", _lib.backend()
    return svm, batch


# peaks spanning every tier: lanes, wavefront teams, LDS workgroup teams, global
PEAKS = [0, 2, 3, 4, 6, 7, 8, 9, 11, 12, 13, 14, 15]


def mixed_programs(seed, count, wide=False):
    rng = np.random.default_rng(seed)
    progs, shots, seeds = [], [], []
    for i in range(count):
        peak = PEAKS[i % len(PEAKS)]
        nq = int(rng.integers(peak + 2, peak + 14)) if not wide or i % 3 else 70 + i
        progs.append(random_program(seed * 1000 + i, num_qubits=nq, peak=peak, length=int(rng.integers(60, 220))))
        shots.append(int(rng.integers(1, 90)) if peak >= 12 else int(rng.integers(1, 400)))
        seeds.append(int(rng.integers(0, 2**63)))
    return progs, shots, seeds


def assert_same(a, b, what):
    assert np.array_equal(a.measurements, b.measurements), what
    assert np.array_equal(a.detectors, b.detectors), what
    assert np.array_equal(a.observables, b.observables), what
    np.testing.assert_allclose(a.exp_vals, b.exp_vals, atol=ATOL, rtol=0, err_msg=what)


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("packed", [False, True])
def test_batch_matches_plain(mods, wide, packed):
    svm, batch = mods
    announce("batch-vs-plain", wide, packed)
    progs, shots, seeds = mixed_programs(11 + wide, 22, wide=wide)
    got = batch.batch_sample(progs, shots, seeds, packed=packed)
    plain = batch.batch_sample(progs, shots, seeds, packed=packed, per_program=True)
    for i, p in enumerate(progs):
        ref = svm.sample(p, shots[i], seed=seeds[i], packed=packed)
        assert_same(got[i], ref, f"program {i} (peak {p.peak_rank}) persistent")
        assert_same(plain[i], ref, f"program {i} (peak {p.peak_rank}) per-program")


def test_batch_matches_oracle(mods, orc):
    svm, batch = mods
    announce("batch-vs-oracle")
    progs, shots, seeds = mixed_programs(5, 11)
    got = batch.batch_sample(progs, shots, seeds)
    for i, p in enumerate(progs):
        m, d, o, e, _ = orc.sample(p, shots[i], seed=seeds[i], per_shot=True)
        assert np.array_equal(got[i].measurements, m), i
        assert np.array_equal(got[i].detectors, d), i
        assert np.array_equal(got[i].observables, o), i
        np.testing.assert_allclose(got[i].exp_vals, e, atol=ATOL, rtol=0)


def _device_batch(batch, progs, shots, packed=False):
    import torch
    dev = torch.device("cuda:0")
    cols = (lambda n: (n + 7) // 8) if packed else (lambda n: n)
    bufs, outs = [], []
    for p, s in zip(progs, shots):
        m = torch.zeros((s, cols(p.num_measurements)), dtype=torch.uint8, device=dev)
        d = torch.zeros((s, cols(p.num_detectors)), dtype=torch.uint8, device=dev)
        o = torch.zeros((s, cols(p.num_observables)), dtype=torch.uint8, device=dev)
        e = torch.zeros((s, p.num_exp_vals), dtype=torch.float64, device=dev)
        bufs.append((m, d, o, e))
        outs.append({"measurements": m.data_ptr() if m.numel() else None,
                     "detectors": d.data_ptr() if d.numel() else None,
                     "observables": o.data_ptr() if o.numel() else None,
                     "exp_vals": e.data_ptr() if e.numel() else None})
    return batch.Batch(progs, shots, outs, packed=packed), bufs


def _snapshot(bufs):
    return [tuple(t.cpu().numpy().copy() for t in b) for b in bufs]


TUNINGS = [dict(wave_rank_max=7, lds_rank_max=10, gwave_rank_max=14, wave_chunk=16, block_chunk=4, lane_chunk=256),
           dict(gwave_rank_max=0), dict(concurrent=0), dict(global_threads=512), dict(global_threads=1024),
           dict(lds_rank_max=8),
           dict(lds_rank_max=13), dict(wave_rank_max=5), dict(wave_rank_max=8, lds_rank_max=8),
           dict(wave_chunk=1, block_chunk=1), dict(wave_chunk=128, block_chunk=64),
           dict(global_threads=512, wave_rank_max=6, lds_rank_max=12, wave_chunk=128, block_chunk=32),
           dict(lane_rank_max=-1), dict(lane_rank_max=5, lane_chunk=100), dict(lane_rank_max=1, lane_chunk=64),
           dict(lane_rank_max=6, lane_chunk=1024), dict(concurrent=0, lane_rank_max=4, lane_chunk=3000),
           dict(gwave_rank_max=13), dict(gwave_rank_max=16, wave_chunk=3), dict(gwave_rank_max=30, lds_rank_max=8),
           dict(wave_rank_max=10, lds_rank_max=10), dict(wave_rank_max=6, lds_rank_max=6, gwave_rank_max=16),
           dict(lane_rank_max=7), dict(lane_rank_max=7, lane_chunk=64, concurrent=0)]


def test_batch_tunings_agree(mods):
    import torch
    svm, batch = mods
    announce("batch-tunings")
    progs, shots, seeds = mixed_programs(21, 22)
    b, bufs = _device_batch(batch, progs, shots)
    b.run(seeds)
    torch.cuda.synchronize()
    ref = _snapshot(bufs)
    for t in TUNINGS:
        announce("tuning", t)
        for buf in bufs:
            for x in buf:
                x.zero_()
        b.set_tuning(**t)
        b.run(seeds)
        torch.cuda.synchronize()
        got = _snapshot(bufs)
        for i in range(len(progs)):
            for j in range(3):
                assert np.array_equal(got[i][j], ref[i][j]), (t, i, j)
            np.testing.assert_allclose(got[i][3], ref[i][3], atol=ATOL, rtol=0)
    b.close()


def test_batch_lane_rank7_matches_oracle(mods, orc):
    """Peak-7 programs one shot per lane (128 KiB of LDS columns, one
    wavefront per workgroup, a launch group apart from the smaller lane
    programs), 32-bit and 64-bit frame words, against the oracle."""
    svm, batch = mods
    announce("batch-lane-rank7")
    rng = np.random.default_rng(77)
    progs, shots, seeds = [], [], []
    for i in range(8):
        peak = 7 if i % 4 else 5
        nq = int(rng.integers(10, 30)) if i % 2 else int(rng.integers(34, 64))
        progs.append(random_program(7700 + i, num_qubits=nq, peak=peak, length=int(rng.integers(120, 260))))
        shots.append(int(rng.integers(60, 300)))
        seeds.append(int(rng.integers(0, 2**63)))
    assert any(p.peak_rank == 7 and p.num_qubits > 32 for p in progs)
    assert any(p.peak_rank == 7 and p.num_qubits <= 32 for p in progs)
    import torch
    b, bufs = _device_batch(batch, progs, shots)
    b.set_tuning(lane_rank_max=7)
    kinds = [g["kind"] for g in b.groups()]
    assert kinds.count("lane") >= 3, kinds
    b.run(seeds)
    torch.cuda.synchronize()
    got = _snapshot(bufs)
    b.close()
    for i, p in enumerate(progs):
        m, d, o, e, _ = orc.sample(p, shots[i], seed=seeds[i], per_shot=True)
        assert np.array_equal(got[i][0], m), i
        assert np.array_equal(got[i][1], d), i
        assert np.array_equal(got[i][2], o), i
        np.testing.assert_allclose(got[i][3], e, atol=ATOL, rtol=0)


def test_batch_autotune(mods):
    import torch
    svm, batch = mods
    announce("batch-autotune")
    progs, shots, seeds = mixed_programs(31, 33)
    b, bufs = _device_batch(batch, progs, shots, packed=True)
    chosen = b.tune(sample_shots=16)
    assert chosen["global_threads"] in (256, 512, 1024) and chosen["concurrent"] in (0, 1)
    assert chosen["wave_rank_max"] <= chosen["lds_rank_max"] <= 13
    assert b.tuning == chosen
    b.run(seeds)
    torch.cuda.synchronize()
    got = _snapshot(bufs)
    for i, p in enumerate(progs):
        ref = svm.sample(p, shots[i], seed=seeds[i], packed=True)
        assert np.array_equal(got[i][0], ref.measurements), i
        assert np.array_equal(got[i][1], ref.detectors), i
        assert np.array_equal(got[i][2], ref.observables), i
    b.close()


@pytest.mark.parametrize("concurrent", [1, 0])
def test_batch_graph_replays(mods, concurrent):
    """Captured graph replays (launch groups on parallel streams, or in
    sequence), with eager work and synchronizations between them, each match
    an eager run with the same seeds."""
    import torch
    svm, batch = mods
    announce("batch-graph", concurrent)
    progs, shots, seeds = mixed_programs(41, 15)
    b, bufs = _device_batch(batch, progs, shots)
    b.set_tuning(concurrent=concurrent)
    rng = np.random.default_rng(3)
    for rep in range(4):
        sd = [int(x) for x in rng.integers(0, 2**63, size=len(progs))]
        b.run(sd)
        torch.cuda.synchronize()
        eager = _snapshot(bufs)
        for buf in bufs:
            for x in buf:
                x.fill_(7)
        torch.cuda.synchronize()
        b.run_graph(sd)
        torch.cuda.synchronize()
        graph = _snapshot(bufs)
        for i in range(len(progs)):
            for j in range(3):
                assert np.array_equal(graph[i][j], eager[i][j]), (rep, i, j)
            np.testing.assert_allclose(graph[i][3], eager[i][3], atol=0, rtol=0)
        _ = torch.ones(1000, device="cuda:0").sum().item()  # unrelated eager work between replays
    b.close()


def test_batch_groups_cover_every_shape(mods):
    """The mixed batch plans one launch group per kernel shape: lanes (by
    frame width), wavefront teams, LDS workgroup teams by size, global
    wavefront teams (peak <= gwave_rank_max) and global workgroup teams;
    every program sits in exactly one group."""
    svm, batch = mods
    progs, shots, seeds = mixed_programs(61, 26, wide=True)
    b, bufs = _device_batch(batch, progs, shots)
    for gwave, kinds in ((14, {"lane", "wave", "lds", "gwave", "global"}), (0, {"lane", "wave", "lds", "global"})):
        b.set_tuning(gwave_rank_max=gwave, wave_rank_max=7, lds_rank_max=10)
        groups = b.groups()
        assert {g["kind"] for g in groups} == kinds, gwave
        assert sum(g["programs"] for g in groups) == len(progs)
        for g in groups:
            assert 0 < g["lds_bytes"] <= 160 * 1024 - 256 and g["grid"] > 0 and g["items"] > 0
            if g["kind"] == "global":
                assert g["block"] in (256, 512, 1024) and g["amp_slices"] in (1, 2, 3)
    b.close()


def test_batch_c3_programs_match_oracle(mods, orc):
    """64 programs of the 4096-program workload generator (16-40 qubits,
    T-count = peak rank 4-16, 40 layers, final M on every qubit) in one
    batch, 24 shots each: records bit-exact against the oracle's per-shot
    streams."""
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile
    svm, batch = mods
    announce("batch-c3")
    progs = [ccompile(text) for text, _ in W.batch_programs(count=64, seed=42, shots=24)]
    assert min(p.peak_rank for p in progs) <= 6 and max(p.peak_rank for p in progs) >= 14
    seeds = [42 ^ i for i in range(len(progs))]
    ref = [orc.sample(p, 24, seed=seeds[i], per_shot=True)[0] for i, p in enumerate(progs)]
    got = batch.batch_sample(progs, 24, seeds, packed=True)
    for i, p in enumerate(progs):
        assert np.array_equal(got[i].measurements, np.packbits(ref[i], axis=1, bitorder="little")), (i, p.peak_rank)
    # the same batch with the global-amplitude programs on wavefront teams
    import torch
    b, bufs = _device_batch(batch, progs, [24] * len(progs), packed=True)
    b.set_tuning(gwave_rank_max=16)
    assert "gwave" in {g["kind"] for g in b.groups()}
    b.run(seeds)
    torch.cuda.synchronize()
    snap = _snapshot(bufs)
    for i, p in enumerate(progs):
        assert np.array_equal(snap[i][0], np.packbits(ref[i], axis=1, bitorder="little")), (i, p.peak_rank)
    b.close()


def test_batch_rejects_bad_tuning(mods):
    svm, batch = mods
    from clifft_amd._lib import SvmError
    progs, shots, seeds = mixed_programs(51, 3)
    b, bufs = _device_batch(batch, progs, shots)
    with pytest.raises(SvmError):
        b.set_tuning(global_threads=300)
    with pytest.raises(SvmError):
        b.set_tuning(concurrent=2)
    with pytest.raises(SvmError):
        b.set_tuning(wave_rank_max=11, lds_rank_max=12)
    with pytest.raises(SvmError):
        b.set_tuning(lds_rank_max=14)
    with pytest.raises(SvmError):
        b.set_tuning(lane_rank_max=8)
    b.run(seeds)  # still usable with the previous tuning
    b.close()
