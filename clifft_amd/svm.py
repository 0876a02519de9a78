"""Python mirror of the reference sampling API on the HIP executor.

    sample(program, shots, seed)            -> SampleResult
        (src/python/bindings.cc:885-918, src/clifft/svm/svm.cc:191-255)
    sample_survivors(program, shots, seed, keep_records)
        (src/python/bindings.cc:1020-1038, svm.cc:266-345)
    record_probabilities(program, records)  -> log-probabilities
        (src/clifft/api/record_probabilities.cc:96-191)
    State / execute(program, state)         -> single shot from a preset state
        (src/python/bindings.cc:1040-1080, svm.cc:158-164)
    sample_k / sample_k_survivors / noise_site_probabilities
        -> exactly-k-fault importance sampling (svm.h:366-378, svm.cc:347-711)

All execution goes through libclifft_svm.so; nothing here falls back to the
CPU.  Seeding: shot s draws from its own xoshiro256++ stream (SplitMix64
outputs 4s..4s+3 of `seed`), so shot 0 matches the reference's first shot.
"""

from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

from . import _lib
from .program import Program
from .sample_result import SampleResult


class DeviceProgram:
    """A Program uploaded to the GPU (bytecode + constant pool + plan).

    Created directly, it keeps its Program alive.  The copy that
    device_program() caches rides on the Program object and refers back to
    it weakly, so dropping the Program frees its device state."""

    def __init__(self, program: Program, _weak: bool = False):
        self._program = weakref.ref(program) if _weak else (lambda p=program: p)
        L = _lib.lib()
        h = ctypes.c_void_p()
        view = program.view()
        _lib.check(L.clifft_svm_program_create(ctypes.byref(view), ctypes.byref(h)), "program_create")
        self.handle = h
        self._fin = weakref.finalize(self, L.clifft_svm_program_destroy, h)

    @property
    def program(self) -> Program:
        p = self._program()
        if p is None:
            raise RuntimeError("the Program of this DeviceProgram no longer exists")
        return p

    def info(self) -> dict:
        info = _lib.ProgramInfo()
        _lib.check(_lib.lib().clifft_svm_program_info_get(self.handle, ctypes.byref(info)), "program_info")
        return {f: getattr(info, f) for f, _ in info._fields_}

    def close(self):
        self._fin()


def device_program(program) -> DeviceProgram:
    """The program's device upload, made on first use and kept on the
    Program object (freed with it)."""
    if isinstance(program, DeviceProgram):
        return program
    dp = program.__dict__.get("_device_program")
    if dp is None:
        dp = DeviceProgram(program, _weak=True)
        program._device_program = dp
    return dp


def _seed(seed):
    if seed is None:
        seed = int.from_bytes(os.urandom(8), "little")
    return int(seed) & (2**64 - 1)


def _ptr(a):
    return None if a is None or a.size == 0 else a.ctypes.data


def sample(program, shots: int, seed=None, packed: bool = False) -> SampleResult:
    """Run `shots` shots; raises ValueError on post-selected programs
    (bindings.cc:888-893)."""
    dp = device_program(program)
    prog = dp.program
    if prog.has_postselection:
        raise ValueError("sample() cannot be used with post-selected programs because it returns a fixed "
                         "number of rows and cannot discard shots. Use sample_survivors(program, shots, "
                         "keep_records=True) instead.")
    shots = int(shots)

    def cols(n):
        return (n + 7) // 8 if packed else n

    m = np.zeros((shots, cols(prog.num_measurements)), np.uint8)
    d = np.zeros((shots, cols(prog.num_detectors)), np.uint8)
    o = np.zeros((shots, cols(prog.num_observables)), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    rec = _lib.Records(_ptr(m), _ptr(d), _ptr(o), _ptr(e), None)
    flags = _lib.REC_PACKED if packed else 0
    _lib.check(_lib.lib().clifft_svm_sample(dp.handle, shots, _seed(seed), flags, ctypes.byref(rec), None), "sample")
    return SampleResult(m, d, o, exp_vals=e)


def sample_device(program, shots: int, seed, out: dict, packed: bool = False, stream=None) -> None:
    """Sample into caller-owned device buffers (e.g. torch tensors' data_ptr):
    out = {"measurements": ptr, "detectors": ptr, "observables": ptr,
    "exp_vals": ptr} (missing keys are skipped)."""
    dp = device_program(program)
    rec = _lib.Records(out.get("measurements"), out.get("detectors"), out.get("observables"),
                       out.get("exp_vals"), out.get("discarded"))
    flags = _lib.REC_DEVICE | (_lib.REC_PACKED if packed else 0)
    _lib.check(_lib.lib().clifft_svm_sample(dp.handle, int(shots), _seed(seed), flags, ctypes.byref(rec), stream),
               "sample")


def sample_survivors(program, shots: int, seed=None, keep_records: bool = False) -> SampleResult:
    dp = device_program(program)
    prog = dp.program
    shots = int(shots)
    ones = np.zeros(max(prog.num_observables, 1), np.uint64)
    summ = _lib.Survivors(shots, 0, 0, ones.ctypes.data)
    if keep_records:
        m = np.zeros((shots, prog.num_measurements), np.uint8)
        d = np.zeros((shots, prog.num_detectors), np.uint8)
        o = np.zeros((shots, prog.num_observables), np.uint8)
        e = np.zeros((shots, prog.num_exp_vals), np.float64)
        rec = _lib.Records(_ptr(m), _ptr(d), _ptr(o), _ptr(e), None)
        recp = ctypes.byref(rec)
    else:
        recp = None
    _lib.check(_lib.lib().clifft_svm_sample_survivors(dp.handle, shots, _seed(seed), 0, ctypes.byref(summ), recp,
                                                      None), "sample_survivors")
    n = int(summ.passed_shots)
    if keep_records:
        m, d, o, e = m[:n], d[:n], o[:n], e[:n]
    else:
        m = np.zeros((0, prog.num_measurements), np.uint8)
        d = np.zeros((0, prog.num_detectors), np.uint8)
        o = np.zeros((0, prog.num_observables), np.uint8)
        e = np.zeros((0, prog.num_exp_vals), np.float64)
    return SampleResult(m, d, o, total_shots=shots, passed_shots=n, logical_errors=int(summ.logical_errors),
                        observable_ones=ones[:prog.num_observables].copy(), exp_vals=e)


UNREACHABLE_LOG_PROB = np.finfo(np.float64).min


def record_probabilities(program, records) -> np.ndarray:
    """Natural-log probability of each measurement record (forced
    execution); unreachable records give UNREACHABLE_LOG_PROB.  Noise,
    detectors, observables and post-selection are rejected
    (record_probabilities.cc:25-91)."""
    dp = device_program(program)
    prog = dp.program
    ops = prog.bytecode["opcode"]
    from .program import Op
    for bad, what in ((Op.NOISE, "noise"), (Op.NOISE_BLOCK, "noise"), (Op.READOUT_NOISE, "readout noise"),
                      (Op.DETECTOR, "detectors"), (Op.POSTSELECT, "post-selection"),
                      (Op.OBSERVABLE, "observables")):
        if np.any(ops == bad):
            raise ValueError(f"record_probabilities: program contains {what}")
    if prog.num_measurements == 0:
        raise ValueError("record_probabilities: program has no measurements")
    if prog.total_meas_slots != prog.num_measurements:
        raise ValueError("record_probabilities: program has hidden measurement slots")
    rec = np.ascontiguousarray(records, dtype=np.uint8).reshape(-1, prog.num_measurements)
    if np.any(rec > 1):
        raise ValueError("record_probabilities: records must be 0/1")
    out = np.zeros(len(rec), np.float64)
    _lib.check(_lib.lib().clifft_svm_record_log_probs(dp.handle, _ptr(rec), len(rec), _ptr(out), None),
               "record_log_probs")
    return out


class State:
    """Host-side VM state for single-shot `execute` (mirrors the reference
    `State`, bindings.cc:1040-1071)."""

    def __init__(self, peak_rank, num_measurements=0, num_qubits=0, num_detectors=0, num_observables=0,
                 num_exp_vals=0, seed=0):
        nq = num_qubits if num_qubits > 0 else max(peak_rank, 1)
        self.words = (nq + 63) // 64
        self.peak_rank = peak_rank
        self.v = np.zeros(1 << peak_rank, np.complex128)
        self.v[0] = 1.0
        self.p_x = np.zeros(self.words, np.uint64)
        self.p_z = np.zeros(self.words, np.uint64)
        self.meas_record = np.zeros(num_measurements, np.uint8)
        self.det_record = np.zeros(num_detectors, np.uint8)
        self.obs_record = np.zeros(num_observables, np.uint8)
        self.exp_vals = np.zeros(num_exp_vals, np.float64)
        self.gamma = 1.0 + 0.0j
        self.active_k = 0
        self.discarded = False
        self.next_noise_idx = 0
        self.forced_reachable = True
        self.forced_log_probability = 0.0
        self.forced_record = None
        self.dust_clamps = 0
        self.rng = np.zeros(4, np.uint64)
        self.reseed(seed)

    @classmethod
    def for_program(cls, prog, seed=0):
        return cls(prog.peak_rank, prog.total_meas_slots, prog.num_qubits, prog.num_detectors,
                   prog.num_observables, prog.num_exp_vals, seed)

    def reseed(self, seed):
        """xoshiro256++ state from SplitMix64(seed) (svm.h:46-52)."""
        z = int(seed) & (2**64 - 1)
        out = []
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & (2**64 - 1)
            x = z
            x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
            x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
            out.append(x ^ (x >> 31))
        self.rng[:] = np.array(out, dtype=np.uint64)


def execute(program, state: State) -> None:
    dp = device_program(program)
    fr = None
    if state.forced_record is not None:
        fr = np.ascontiguousarray(state.forced_record, dtype=np.uint8)
    s = _lib.State()
    s.v = state.v.ctypes.data
    s.px, s.pz = state.p_x.ctypes.data, state.p_z.ctypes.data
    s.gamma_re, s.gamma_im = complex(state.gamma).real, complex(state.gamma).imag
    s.active_k = state.active_k
    s.discarded = int(state.discarded)
    s.meas_record, s.det_record, s.obs_record = _ptr(state.meas_record), _ptr(state.det_record), _ptr(state.obs_record)
    s.exp_vals = _ptr(state.exp_vals)
    for i in range(4):
        s.rng[i] = int(state.rng[i])
    s.next_noise_idx = state.next_noise_idx
    s.forced_reachable = int(state.forced_reachable)
    s.forced_log_probability = state.forced_log_probability
    s.forced_record = None if fr is None else fr.ctypes.data
    s.dust_clamps = state.dust_clamps
    _lib.check(_lib.lib().clifft_svm_execute(dp.handle, ctypes.byref(s)), "execute")
    state.gamma = complex(s.gamma_re, s.gamma_im)
    state.active_k = s.active_k
    state.discarded = bool(s.discarded)
    state.rng[:] = np.array([s.rng[i] for i in range(4)], dtype=np.uint64)
    state.next_noise_idx = s.next_noise_idx
    state.forced_reachable = bool(s.forced_reachable)
    state.forced_log_probability = s.forced_log_probability
    state.dust_clamps = s.dust_clamps


def _check_k(rc: int, what: str) -> None:
    if rc == _lib.E_INVALID:
        raise ValueError(_lib.lib().clifft_svm_last_error().decode(errors="replace"))
    _lib.check(rc, what)


def sample_k(program, shots: int, k: int, seed=None, packed: bool = False) -> SampleResult:
    """Every shot has exactly k faults (noise sites + readout entries); raises
    ValueError for an infeasible k, like the reference (svm.cc:525-546)."""
    dp = device_program(program)
    prog = dp.program
    shots = int(shots)

    def cols(n):
        return (n + 7) // 8 if packed else n

    m = np.zeros((shots, cols(prog.num_measurements)), np.uint8)
    d = np.zeros((shots, cols(prog.num_detectors)), np.uint8)
    o = np.zeros((shots, cols(prog.num_observables)), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    rec = _lib.Records(_ptr(m), _ptr(d), _ptr(o), _ptr(e), None)
    flags = _lib.REC_PACKED if packed else 0
    _check_k(_lib.lib().clifft_svm_sample_k(dp.handle, shots, int(k), _seed(seed), flags, ctypes.byref(rec), None),
             "sample_k")
    return SampleResult(m, d, o, exp_vals=e)


def sample_k_survivors(program, shots: int, k: int, seed=None, keep_records: bool = False) -> SampleResult:
    dp = device_program(program)
    prog = dp.program
    shots = int(shots)
    ones = np.zeros(max(prog.num_observables, 1), np.uint64)
    summ = _lib.Survivors(shots, 0, 0, ones.ctypes.data)
    m = np.zeros((shots, prog.num_measurements), np.uint8)
    d = np.zeros((shots, prog.num_detectors), np.uint8)
    o = np.zeros((shots, prog.num_observables), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    rec = _lib.Records(_ptr(m), _ptr(d), _ptr(o), _ptr(e), None)
    _check_k(_lib.lib().clifft_svm_sample_k_survivors(dp.handle, shots, int(k), _seed(seed), 0, ctypes.byref(summ),
                                                      ctypes.byref(rec) if keep_records else None, None),
             "sample_k_survivors")
    n = int(summ.passed_shots)
    if keep_records:
        m, d, o, e = m[:n], d[:n], o[:n], e[:n]
    else:
        m, d, o, e = m[:0], d[:0], o[:0], e[:0]
    return SampleResult(m, d, o, total_shots=shots, passed_shots=n, logical_errors=int(summ.logical_errors),
                        observable_ones=ones[:prog.num_observables].copy(), exp_vals=e)


def noise_site_probabilities(program) -> np.ndarray:
    dp = device_program(program)
    n = ctypes.c_uint32(0)
    _lib.check(_lib.lib().clifft_svm_noise_site_probabilities(dp.handle, None, ctypes.byref(n)), "noise_sites")
    out = np.zeros(n.value, np.float64)
    _lib.check(_lib.lib().clifft_svm_noise_site_probabilities(dp.handle, _ptr(out), ctypes.byref(n)), "noise_sites")
    return out
