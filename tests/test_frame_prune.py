"""The frame opcodes left out of the packed bytecode copies, and the
superinstructions that pack the rest, checked on the CPU.

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

`pack_frame_runs` is compiled the same way (tests/cpp/frame_pack_check.cpp)
and its output, read with the superinstructions' documented meaning (packed
runs; fans of CZs or CNOTs that share a qubit), must move a frame and the
phase of gamma exactly as the original opcodes applied one at a time, from
random frames, on random frame-opcode sequences rich in fans, with random
instructions marked as identities (and left out of both sides).
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


# ---------------------------------------------------------------- packing
PACK_CHECKER = os.path.join(ROOT, "tests", "cpp", "frame_pack_check.cpp")


def _pack_text() -> str:
    import re
    with open(os.path.join(ROOT, "clifft_amd", "csrc", "svm_device.h")) as f:
        dev = f.read()
    consts = [m.group(0) for m in re.finditer(r"constexpr uint32_t kLaneFrame\w+ = [^;]+;", dev)]
    consts.append(re.search(r"enum : uint32_t \{ kFanCz[^}]*\};", dev).group(0))
    with open(os.path.join(ROOT, "clifft_amd", "csrc", "svm_host.cpp")) as f:
        src = f.read()
    body = src[src.index("// Runs of frame opcodes on axes below 64 become superinstructions"):
               src.index("// CLIFFT_ARRAY_FANOUT=0 keeps")]
    return "\n".join(consts) + "\n" + body


def _build_pack(tmp_path, text):
    (tmp_path / "frame_pack.inc").write_text(text)
    so = tmp_path / "libframe_pack.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", str(tmp_path), "-I",
                    os.path.join(ROOT, "include"), PACK_CHECKER, "-o", str(so)], check=True, timeout=120)
    lib = ctypes.CDLL(str(so))
    lib.pack.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    lib.pack.restype = None
    return lib


def _apply(op, a, b, st):
    """One frame opcode on st = [x, z, e] (gamma's phase as a power e of i)."""
    x, z, e = st
    bit = lambda w, q: (w >> q) & 1
    if op == 0:  # CNOT
        xa, zb = bit(x, a), bit(z, b)
        x ^= xa << b
        z ^= zb << a
    elif op == 1:  # CZ
        xa, xb = bit(x, a), bit(x, b)
        e += 2 * (xa & xb)
        z ^= (xa << b) ^ (xb << a)
    elif op == 2:  # H
        xa, za = bit(x, a), bit(z, a)
        e += 2 * (xa & za)
        if xa != za:
            x ^= 1 << a
            z ^= 1 << a
    elif op in (3, 4):  # S, S_DAG
        xa = bit(x, a)
        e += (1 if op == 3 else 3) * xa
        z ^= xa << a
    else:  # SWAP
        dx, dz = bit(x, a) ^ bit(x, b), bit(z, a) ^ bit(z, b)
        x ^= (dx << a) | (dx << b)
        z ^= (dz << a) | (dz << b)
    st[:] = [x, z, e % 4]


def _run_packed(code, st):
    bit = lambda w, q: (w >> q) & 1
    par = lambda w: bin(w).count("1") & 1
    pc, n = 0, len(code)
    while pc < n:
        ins = code[pc]
        op, a, b = int(ins["opcode"]), int(ins["axis_1"]), int(ins["axis_2"])
        raw = ins.tobytes()[8:32]
        skip = 0
        if op == 0xF0:
            words = np.frombuffer(raw[:24], dtype=np.uint64)
            for j in range(a):
                f = (int(words[j // 4]) >> (16 * (j % 4))) & 0xFFFF
                _apply(f & 7, (f >> 3) & 63, (f >> 9) & 63, st)
            skip = b
        elif op == 0xF3:
            kind, p = int(ins["reserved"]), a
            m = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
            x, z, e = st
            if kind == 0:
                xp, q = bit(x, p), par(x & m)
                e += 2 * (xp & q)
                z ^= (m if xp else 0) ^ (q << p)
            elif kind == 1:
                xp, q = bit(x, p), par(z & m)
                x ^= m if xp else 0
                z ^= q << p
            else:
                zp, q = bit(z, p), par(x & m)
                x ^= q << p
                z ^= m if zp else 0
            st[:] = [x, z, e % 4]
            skip = b
        else:
            assert op <= 5, op
            _apply(op, a, b, st)
        pc += 1 + skip


def _random_frame_code(rng, n, nq):
    from clifft_amd.program import (make_frame_cnot, make_frame_cz, make_frame_h, make_frame_s, make_frame_s_dag,
                                    make_frame_swap)
    ops = []
    while len(ops) < n:
        r = rng.random()
        p = int(rng.integers(0, nq))
        others = [int(q) for q in rng.choice([q for q in range(nq) if q != p], size=int(rng.integers(1, 9)))]
        if r < 0.3:  # a CZ fan around p, each op in either orientation
            ops += [make_frame_cz(p, q) if rng.random() < 0.5 else make_frame_cz(q, p) for q in others]
        elif r < 0.5:
            ops += [make_frame_cnot(p, q) for q in others]
        elif r < 0.65:
            ops += [make_frame_cnot(q, p) for q in others]
        else:
            a, b = (int(v) for v in rng.choice(nq, size=2, replace=False))
            ops.append([make_frame_cnot, make_frame_cz, make_frame_swap][int(rng.integers(0, 3))](a, b)
                       if rng.random() < 0.6 else
                       [make_frame_h, make_frame_s, make_frame_s_dag][int(rng.integers(0, 3))](a))
    return ops[:n]


def test_packed_frame_units_equal_the_ops(tmp_path):
    from clifft_amd.program import INSTR_DTYPE
    lib = _build_pack(tmp_path, _pack_text())
    rng = np.random.default_rng(2024)
    fans = units = 0
    for case in range(300):
        nq = int(rng.integers(3, 65))
        ops = _random_frame_code(rng, int(rng.integers(1, 120)), nq)
        code = np.array(ops, dtype=INSTR_DTYPE)
        idle = (rng.random(len(code)) < (0.0 if case % 3 == 0 else 0.3)).astype(np.uint8)
        packed = code.copy()
        lib.pack(packed.ctypes.data, len(packed), idle.ctypes.data)
        fans += int((packed["opcode"] == 0xF3).sum())
        units += int((packed["opcode"] >= 0xF0).sum())
        for _ in range(4):
            x, z = (int(rng.integers(0, 2**63)) & ((1 << nq) - 1) for _ in range(2))
            want = [x, z, 0]
            for i, ins in enumerate(code):
                if not idle[i]:
                    _apply(int(ins["opcode"]), int(ins["axis_1"]), int(ins["axis_2"]), want)
            got = [x, z, 0]
            _run_packed(packed, got)
            assert got == want, (case, nq)
    assert fans > 300 and units > fans


def test_planted_fan_mistake_is_caught(tmp_path):
    from clifft_amd.program import INSTR_DTYPE
    text = _pack_text()
    good = "(kind == kFanControl ? f.axis_2 : f.axis_1);"
    assert good in text
    lib = _build_pack(tmp_path, text.replace(good, "(kind == kFanControl ? f.axis_1 : f.axis_1);"))
    rng = np.random.default_rng(7)
    bad = 0
    for case in range(60):
        nq = int(rng.integers(3, 40))
        code = np.array(_random_frame_code(rng, 60, nq), dtype=INSTR_DTYPE)
        idle = np.zeros(len(code), np.uint8)
        packed = code.copy()
        lib.pack(packed.ctypes.data, len(packed), idle.ctypes.data)
        x, z = (int(rng.integers(1, 2**40)) & ((1 << nq) - 1) for _ in range(2))
        want, got = [x, z, 0], [x, z, 0]
        for ins in code:
            _apply(int(ins["opcode"]), int(ins["axis_1"]), int(ins["axis_2"]), want)
        _run_packed(packed, got)
        bad += got != want
    assert bad > 0
