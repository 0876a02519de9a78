"""Batched persistent execution of many programs (include/clifft_svm.h,
clifft_svm_batch_*).

No counterpart in the reference, whose VM runs one program per call
(src/clifft/svm/svm.h:302-306, src/python/bindings.cc:885-918).  Records of
every program are bit-identical to `svm.sample(program, shots, seed)`.

    batch_sample(programs, shots, seeds, packed=False, per_program=False)
        -> list[SampleResult]   (host outputs; per_program=True issues one
                                 plain launch per program instead)
    Batch(programs, shots, outs, packed=False)
        device-output handle: tune(), tuning, set_tuning(), groups(),
        run(seeds), run_graph(seeds)
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .sample_result import SampleResult
from .svm import _ptr, _seed, device_program

_TUNING_FIELDS = [f for f, _ in _lib.BatchTuning._fields_]


def _handles(programs):
    dps = [device_program(p) for p in programs]
    arr = (ctypes.c_void_p * len(dps))(*[d.handle.value for d in dps])
    return dps, arr


def _u64(values, n):
    a = np.ascontiguousarray(np.asarray(values, dtype=np.uint64).reshape(-1))
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def batch_sample(programs, shots, seeds, packed: bool = False, per_program: bool = False):
    """Sample every program; shots/seeds are per program (or one int for all)."""
    programs = list(programs)
    n = len(programs)
    if n == 0:
        return []
    shots = _u64(shots if np.ndim(shots) else [shots] * n, n)
    seeds = _u64([_seed(s) for s in (seeds if np.ndim(seeds) else [seeds] * n)], n)
    dps, handles = _handles(programs)
    results, recs = [], (_lib.Records * n)()

    def cols(k):
        return (k + 7) // 8 if packed else k

    for i, dp in enumerate(dps):
        p = dp.program
        if p.has_postselection:
            raise ValueError("batch_sample() cannot be used with post-selected programs")
        s = int(shots[i])
        m = np.zeros((s, cols(p.num_measurements)), np.uint8)
        d = np.zeros((s, cols(p.num_detectors)), np.uint8)
        o = np.zeros((s, cols(p.num_observables)), np.uint8)
        e = np.zeros((s, p.num_exp_vals), np.float64)
        recs[i] = _lib.Records(_ptr(m), _ptr(d), _ptr(o), _ptr(e), None)
        results.append(SampleResult(m, d, o, exp_vals=e))
    flags = (_lib.REC_PACKED if packed else 0) | (_lib.BATCH_PER_PROGRAM if per_program else 0)
    _lib.check(_lib.lib().clifft_svm_batch_sample(handles, shots.ctypes.data, seeds.ctypes.data, n, flags, recs,
                                                  None), "batch_sample")
    return results


class Batch:
    """A batch bound to caller-owned device output buffers.

    outs[i] is a dict of device pointers like svm.sample_device's `out`."""

    def __init__(self, programs, shots, outs, packed: bool = False):
        programs = list(programs)
        n = len(programs)
        self.n = n
        self.shots = _u64(shots if np.ndim(shots) else [shots] * n, n)
        self._dps, handles = _handles(programs)
        recs = (_lib.Records * n)()
        for i, o in enumerate(outs):
            recs[i] = _lib.Records(o.get("measurements"), o.get("detectors"), o.get("observables"),
                                   o.get("exp_vals"), o.get("discarded"))
        h = ctypes.c_void_p()
        flags = _lib.REC_DEVICE | (_lib.REC_PACKED if packed else 0)
        _lib.check(_lib.lib().clifft_svm_batch_create(handles, self.shots.ctypes.data, n, flags, recs,
                                                      ctypes.byref(h)), "batch_create")
        self.handle = h

    def close(self):
        if self.handle:
            _lib.lib().clifft_svm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def groups(self) -> list:
        """The current plan's launch groups (kind: wave / lds / global / lane /
        gwave, the last one shot per wavefront on global amplitudes)."""
        n = ctypes.c_uint32()
        _lib.check(_lib.lib().clifft_svm_batch_info(self.handle, 0, None, ctypes.byref(n)), "batch_info")
        arr = (_lib.BatchGroupInfo * max(n.value, 1))()
        _lib.check(_lib.lib().clifft_svm_batch_info(self.handle, n.value, arr, ctypes.byref(n)), "batch_info")
        kinds = ["wave", "lds", "global", "lane", "gwave"]
        out = []
        for g in arr[:n.value]:
            d = {f: getattr(g, f) for f, _ in g._fields_}
            d["kind"] = kinds[d["kind"]]
            out.append(d)
        return out

    def time_groups(self, enable: bool = True):
        """Bracket each group's launch of later eager runs with timing events."""
        _lib.check(_lib.lib().clifft_svm_batch_time_groups(self.handle, int(bool(enable))), "batch_time_groups")

    def group_times(self) -> list:
        """Kernel ms of each group in the last timed eager run."""
        n = ctypes.c_uint32()
        g = len(self.groups())
        ms = (ctypes.c_float * max(g, 1))()
        _lib.check(_lib.lib().clifft_svm_batch_group_times(self.handle, g, ms, ctypes.byref(n)), "batch_group_times")
        return [float(x) for x in ms[:n.value]]

    @property
    def tuning(self) -> dict:
        t = _lib.BatchTuning()
        _lib.check(_lib.lib().clifft_svm_batch_get_tuning(self.handle, ctypes.byref(t)), "batch_get_tuning")
        return {f: getattr(t, f) for f in _TUNING_FIELDS}

    def set_tuning(self, **kw):
        cur = self.tuning
        cur.update(kw)
        t = _lib.BatchTuning(*[int(cur[f]) for f in _TUNING_FIELDS])
        _lib.check(_lib.lib().clifft_svm_batch_set_tuning(self.handle, ctypes.byref(t)), "batch_set_tuning")

    def tune(self, sample_shots: int = 0, stream=None) -> dict:
        t = _lib.BatchTuning()
        _lib.check(_lib.lib().clifft_svm_batch_tune(self.handle, int(sample_shots), stream, ctypes.byref(t)),
                   "batch_tune")
        return {f: getattr(t, f) for f in _TUNING_FIELDS}

    def _seeds(self, seeds):
        return _u64(seeds if np.ndim(seeds) else [seeds] * self.n, self.n)

    def run(self, seeds, stream=None):
        s = self._seeds(seeds)
        _lib.check(_lib.lib().clifft_svm_batch_run(self.handle, s.ctypes.data, stream), "batch_run")

    def run_graph(self, seeds, stream=None):
        s = self._seeds(seeds)
        _lib.check(_lib.lib().clifft_svm_batch_run_graph(self.handle, s.ctypes.data, stream), "batch_run_graph")
