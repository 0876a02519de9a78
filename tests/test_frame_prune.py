"""The frame opcodes left out of the packed bytecode copies, checked on the CPU.

`identity_frame_ops` (clifft_amd/csrc/svm_host.cpp) marks the FRAME_*
instructions that are the identity on every shot that starts from the zero
frame; the device's sampling, forced and k-fault copies skip them.  Its text
is compiled for the host with g++ (tests/cpp/frame_prune_check.cpp), and the
claim is checked where it matters: the oracle, which runs every opcode,
samples each random program and the same program with the marked
instructions deleted, with the same seed, and the two runs' records and
expectation values must be identical.  The programs cover every opcode the
random-program generator emits (noise sites, APPLY_PAULI resets, U2/U4
nodes, dormant and active measurements, post-selection, expectation values)
and compiled Clifford+T circuits, where most frame opcodes are marked.  A
planted mistake in the analysis (a CNOT that ignores its z input, a
measurement that clears its outcome bit) must make some pair differ.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import ctypes

import numpy as np
import pytest

from randprog import random_program

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECKER = os.path.join(ROOT, "tests", "cpp", "frame_prune_check.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")


def _analysis() -> str:
    with open(os.path.join(ROOT, "clifft_amd", "csrc", "svm_host.cpp")) as f:
        src = f.read()
    return src[src.index("// CLIFFT_FRAME_PRUNE=0 keeps"):src.index("// Runs of frame opcodes on axes below 64")]


def _build(tmp_path, text: str):
    (tmp_path / "frame_prune.inc").write_text(text)
    so = tmp_path / "libframe_prune.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", str(tmp_path), "-I",
                    os.path.join(ROOT, "include"), CHECKER, "-o", str(so)], check=True, timeout=120)
    lib = ctypes.CDLL(str(so))
    lib.idle_mask.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.idle_mask.restype = None
    return lib


def _mask(lib, p):
    v = p.view()
    out = np.zeros(len(p.bytecode), np.uint8)
    lib.idle_mask(ctypes.addressof(v), out.ctypes.data)
    return out.astype(bool)


def _programs():
    from clifft_amd import workloads as W
    from clifft_amd.compiler import compile as ccompile
    progs = []
    for seed in range(24):
        nq = (8, 12, 20, 40, 64)[seed % 5]
        progs.append(random_program(500 + seed, num_qubits=nq, peak=min(6, nq), length=220,
                                    noise=seed % 3 != 0, postselect=seed % 4 == 1))
    for seed in range(6):
        progs.append(ccompile(W.random_clifford_t(n=12 + 2 * seed, layers=16, t_count=4 + seed, seed=seed,
                                                  final_measure=True, mr_every=5 if seed % 2 else 0,
                                                  mr_count=2)))
    return progs


def _pruned(p, idle):
    q = p.copy()
    q.bytecode = p.bytecode[~idle].copy()
    return q


def _same_records(orc, p, q, shots=64, seed=5):
    a = orc.sample(p, shots, seed=seed, per_shot=True)
    b = orc.sample(q, shots, seed=seed, per_shot=True)
    return all(np.array_equal(x, y) for x, y in zip(a[:4], b[:4]))


def test_marked_frame_ops_are_identities(tmp_path, orc):
    lib = _build(tmp_path, _analysis())
    marked = 0
    compiled_marked = 0
    for i, p in enumerate(_programs()):
        idle = _mask(lib, p)
        assert np.all(p.bytecode["opcode"][idle] <= 5), i  # FRAME_* only
        marked += int(idle.sum())
        if p.param_instr is not None and len(p.param_instr):
            continue
        if i >= 24:
            compiled_marked += int(idle.sum())
        assert _same_records(orc, p, _pruned(p, idle)), i
    assert marked > 0
    assert compiled_marked > 200  # most frame opcodes of the Clifford+T circuits


def test_words_above_one_mark_nothing(tmp_path):
    lib = _build(tmp_path, _analysis())
    p = random_program(7, num_qubits=100, peak=5, length=150)
    assert p.words == 2
    assert not _mask(lib, p).any()


@pytest.mark.parametrize("good,bad", [
    ("const bool xa = has(mx, a), zb = has(mz, b);", "const bool xa = has(mx, a), zb = false;"),
    ("                mx |= bit(a);\n                mz &= ~bit(a);", "                mx &= ~bit(a);\n                mz &= ~bit(a);"),
])
def test_planted_mistake_is_caught(tmp_path, orc, good, bad):
    text = _analysis()
    assert good in text
    lib = _build(tmp_path, text.replace(good, bad, 1))
    differ = 0
    for p in _programs():
        idle = _mask(lib, p)
        if idle.any() and not _same_records(orc, p, _pruned(p, idle)):
            differ += 1
    assert differ > 0
