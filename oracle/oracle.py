"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/libsvm_oracle.so (built by oracle/Makefile).  Used by tests/,
the smoke check and the CPU-baseline leg of bench.py as the checker; the
product (clifft_amd) never imports this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(fast: bool = False):
    name = "libsvm_oracle_fast.so" if fast else "libsvm_oracle.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, u32, u64, dbl = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
        sigs = {
            "orc_state_new": (vp, [u32, u32, u32, u32, u32, u32, u64]),
            "orc_state_free": (None, [vp]),
            "orc_state_reset": (None, [vp]),
            "orc_state_v": (ctypes.POINTER(ctypes.c_double), [vp]),
            "orc_state_px": (ctypes.POINTER(ctypes.c_uint64), [vp]),
            "orc_state_pz": (ctypes.POINTER(ctypes.c_uint64), [vp]),
            "orc_state_meas": (ctypes.POINTER(ctypes.c_uint8), [vp]),
            "orc_state_det": (ctypes.POINTER(ctypes.c_uint8), [vp]),
            "orc_state_obs": (ctypes.POINTER(ctypes.c_uint8), [vp]),
            "orc_state_ev": (ctypes.POINTER(ctypes.c_double), [vp]),
            "orc_state_rng": (ctypes.POINTER(ctypes.c_uint64), [vp]),
            "orc_state_get_gamma": (None, [vp, vp]),
            "orc_state_set_gamma": (None, [vp, dbl, dbl]),
            "orc_state_get_k": (u32, [vp]),
            "orc_state_set_k": (None, [vp, u32]),
            "orc_state_discarded": (u32, [vp]),
            "orc_state_dust_clamps": (u64, [vp]),
            "orc_state_next_noise": (u32, [vp]),
            "orc_state_set_next_noise": (None, [vp, u32]),
            "orc_state_reseed": (None, [vp, u64]),
            "orc_state_seed_shot": (None, [vp, u64, u64]),
            "orc_state_random_double": (dbl, [vp]),
            "orc_state_forced_log_prob": (dbl, [vp]),
            "orc_state_forced_reachable": (u32, [vp]),
            "orc_state_set_forced_record": (None, [vp, vp]),
            "orc_execute": (None, [vp, vp]),
            "orc_draw_next_noise": (None, [vp, vp]),
            "orc_sample": (ctypes.c_int, [vp, u64, u64, ctypes.c_int, u64, vp, vp, vp, vp, vp]),
            "orc_sample_parallel": (ctypes.c_int, [vp, u64, u64, ctypes.c_int, vp, vp, vp, vp]),
            "orc_sample_survivors": (ctypes.c_int, [vp, u64, u64, ctypes.c_int, vp, vp, vp, vp, vp, vp]),
            "orc_record_log_probs": (ctypes.c_int, [vp, vp, u64, vp]),
            "orc_execute_forced_faults": (ctypes.c_int, [vp, vp, vp, u32, vp, u32]),
            "orc_max_threads": (ctypes.c_int, []),
            "orc_sample_k": (ctypes.c_int, [vp, u64, u32, u64, ctypes.c_int, u64, vp, vp, vp, vp, vp]),
            "orc_noise_site_probabilities": (u32, [vp, vp]),
        }
        for fn, (res, args) in sigs.items():
            f = getattr(L, fn)
            f.restype = res
            f.argtypes = args
        _LIBS[name] = L
    return _LIBS[name]


def _p(a):
    return None if a is None else a.ctypes.data


class State:
    """Oracle SchrodingerState (src/clifft/svm/svm.h:115-289)."""

    def __init__(self, peak_rank, num_measurements=0, num_qubits=0, num_detectors=0,
                 num_observables=0, num_exp_vals=0, seed=0):
        self.L = lib()
        self.peak_rank = peak_rank
        self.n_meas, self.n_det, self.n_obs, self.n_ev = num_measurements, num_detectors, num_observables, num_exp_vals
        nq = num_qubits if num_qubits > 0 else max(peak_rank, 1)
        self.words = (nq + 63) // 64
        self.h = self.L.orc_state_new(peak_rank, num_measurements, num_qubits, num_detectors,
                                      num_observables, num_exp_vals, seed & (2**64 - 1))
        if not self.h:
            raise ValueError("peak_rank >= 63")

    @classmethod
    def for_program(cls, prog, seed=0):
        return cls(prog.peak_rank, prog.total_meas_slots, prog.num_qubits, prog.num_detectors,
                   prog.num_observables, prog.num_exp_vals, seed)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_state_free(self.h)
            self.h = None

    # views into C memory
    def _arr(self, fn, n, dtype):
        ptr = fn(self.h)
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype) if n > 0 else np.zeros(0, dtype)

    @property
    def v(self):
        raw = self._arr(self.L.orc_state_v, 2 << self.peak_rank, np.float64)
        return raw.view(np.complex128)

    @property
    def p_x(self):
        return self._arr(self.L.orc_state_px, self.words, np.uint64)

    @property
    def p_z(self):
        return self._arr(self.L.orc_state_pz, self.words, np.uint64)

    @property
    def meas_record(self):
        return self._arr(self.L.orc_state_meas, self.n_meas, np.uint8)

    @property
    def det_record(self):
        return self._arr(self.L.orc_state_det, self.n_det, np.uint8)

    @property
    def obs_record(self):
        return self._arr(self.L.orc_state_obs, self.n_obs, np.uint8)

    @property
    def exp_vals(self):
        return self._arr(self.L.orc_state_ev, self.n_ev, np.float64)

    @property
    def rng(self):
        return self._arr(self.L.orc_state_rng, 4, np.uint64)

    @property
    def gamma(self) -> complex:
        out = (ctypes.c_double * 2)()
        self.L.orc_state_get_gamma(self.h, out)
        return complex(out[0], out[1])

    @gamma.setter
    def gamma(self, g):
        g = complex(g)
        self.L.orc_state_set_gamma(self.h, g.real, g.imag)

    @property
    def active_k(self) -> int:
        return self.L.orc_state_get_k(self.h)

    @active_k.setter
    def active_k(self, k):
        self.L.orc_state_set_k(self.h, k)

    @property
    def discarded(self) -> bool:
        return bool(self.L.orc_state_discarded(self.h))

    @property
    def dust_clamps(self) -> int:
        return self.L.orc_state_dust_clamps(self.h)

    @property
    def next_noise_idx(self) -> int:
        return self.L.orc_state_next_noise(self.h)

    @next_noise_idx.setter
    def next_noise_idx(self, v):
        self.L.orc_state_set_next_noise(self.h, v)

    @property
    def forced_log_probability(self) -> float:
        return self.L.orc_state_forced_log_prob(self.h)

    @property
    def forced_reachable(self) -> bool:
        return bool(self.L.orc_state_forced_reachable(self.h))

    def set_forced_record(self, rec: np.ndarray):
        self._forced = np.ascontiguousarray(rec, dtype=np.uint8)
        self.L.orc_state_set_forced_record(self.h, self._forced.ctypes.data)

    def reset(self):
        self.L.orc_state_reset(self.h)

    def reseed(self, seed):
        self.L.orc_state_reseed(self.h, seed & (2**64 - 1))

    def seed_shot(self, seed, shot):
        self.L.orc_state_seed_shot(self.h, seed & (2**64 - 1), shot)

    def random_double(self) -> float:
        return self.L.orc_state_random_double(self.h)


def execute(prog, state: State):
    state._keep = prog.view()
    state.L.orc_execute(ctypes.byref(state._keep), state.h)


def draw_next_noise(prog, state: State):
    v = prog.view()
    state.L.orc_draw_next_noise(ctypes.byref(v), state.h)


def execute_forced_faults(prog, state: State, noise_idx, readout_idx):
    v = prog.view()
    n = np.ascontiguousarray(noise_idx, dtype=np.uint32)
    r = np.ascontiguousarray(readout_idx, dtype=np.uint32)
    state.L.orc_execute_forced_faults(ctypes.byref(v), state.h, _p(n), len(n), _p(r), len(r))


def sample(prog, shots, seed=0, per_shot=True, first_shot=0, fast=False):
    """Oracle sample().  per_shot=False reproduces the reference single
    stream (src/clifft/svm/svm.cc:191-255); per_shot=True is the device
    executor's per-shot stream convention.  fast=True uses the -O3
    -ffast-math build (CPU-baseline timing only)."""
    L = lib(fast)
    v = prog.view()
    m = np.zeros((shots, prog.num_measurements), np.uint8)
    d = np.zeros((shots, prog.num_detectors), np.uint8)
    o = np.zeros((shots, prog.num_observables), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    disc = np.zeros(shots, np.uint8)
    rc = L.orc_sample(ctypes.byref(v), shots, seed & (2**64 - 1), 1 if per_shot else 0, first_shot,
                      _p(m), _p(d), _p(o), _p(e), _p(disc))
    assert rc == 0
    return m, d, o, e, disc


def sample_k(prog, shots, k, seed=0, per_shot=True, first_shot=0):
    """Oracle sample_k (src/clifft/svm/svm.cc:549-622): exactly k forced
    faults per shot.  Raises ValueError where the reference raises
    std::invalid_argument."""
    L = lib()
    v = prog.view()
    m = np.zeros((shots, prog.num_measurements), np.uint8)
    d = np.zeros((shots, prog.num_detectors), np.uint8)
    o = np.zeros((shots, prog.num_observables), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    disc = np.zeros(shots, np.uint8)
    rc = L.orc_sample_k(ctypes.byref(v), shots, k, seed & (2**64 - 1), 1 if per_shot else 0, first_shot,
                        _p(m), _p(d), _p(o), _p(e), _p(disc))
    if rc == -2:
        raise ValueError(f"k ({k}) exceeds total fault sites")
    if rc == -3:
        raise ValueError(f"k-fault stratum k={k} has zero probability mass")
    assert rc == 0
    return m, d, o, e, disc


def noise_site_probabilities(prog):
    L = lib()
    v = prog.view()
    n = L.orc_noise_site_probabilities(ctypes.byref(v), None)
    out = np.zeros(n, np.float64)
    L.orc_noise_site_probabilities(ctypes.byref(v), _p(out))
    return out


def sample_parallel(prog, shots, seed=0, threads=0, fast=True):
    L = lib(fast)
    v = prog.view()
    m = np.zeros((shots, prog.num_measurements), np.uint8)
    d = np.zeros((shots, prog.num_detectors), np.uint8)
    o = np.zeros((shots, prog.num_observables), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    rc = L.orc_sample_parallel(ctypes.byref(v), shots, seed & (2**64 - 1), threads, _p(m), _p(d), _p(o), _p(e))
    assert rc == 0
    return m, d, o, e


def sample_survivors(prog, shots, seed=0, per_shot=True):
    L = lib()
    v = prog.view()
    counts = np.zeros(2, np.uint64)
    ones = np.zeros(max(prog.num_observables, 1), np.uint64)
    m = np.zeros((shots, prog.num_measurements), np.uint8)
    d = np.zeros((shots, prog.num_detectors), np.uint8)
    o = np.zeros((shots, prog.num_observables), np.uint8)
    e = np.zeros((shots, prog.num_exp_vals), np.float64)
    rc = L.orc_sample_survivors(ctypes.byref(v), shots, seed & (2**64 - 1), 1 if per_shot else 0,
                                _p(counts), _p(ones), _p(m), _p(d), _p(o), _p(e))
    assert rc == 0
    n = int(counts[0])
    return dict(passed=n, logical_errors=int(counts[1]), observable_ones=ones[:prog.num_observables],
                measurements=m[:n], detectors=d[:n], observables=o[:n], exp_vals=e[:n])


def record_log_probs(prog, records):
    """Forced execution per record; `prog` must already use *_FORCED opcodes."""
    L = lib()
    v = prog.view()
    rec = np.ascontiguousarray(records, dtype=np.uint8).reshape(-1, prog.num_measurements)
    out = np.zeros(len(rec), np.float64)
    rc = L.orc_record_log_probs(ctypes.byref(v), _p(rec), len(rec), _p(out))
    assert rc == 0
    return out
