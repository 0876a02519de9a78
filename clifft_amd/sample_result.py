"""SampleResult: same attributes and tuple-unpacking as the reference's
src/python/clifft/_sample_result.py:11-86."""

from __future__ import annotations

import numpy as np


class SampleResult:
    __slots__ = ("measurements", "detectors", "observables", "exp_vals", "total_shots", "passed_shots",
                 "discards", "logical_errors", "observable_ones")

    def __init__(self, measurements, detectors, observables, total_shots=None, passed_shots=None,
                 logical_errors=None, observable_ones=None, exp_vals=None):
        self.measurements = measurements
        self.detectors = detectors
        self.observables = observables
        self.exp_vals = exp_vals if exp_vals is not None else np.empty((measurements.shape[0], 0), np.float64)
        self.total_shots = total_shots
        self.passed_shots = passed_shots
        self.discards = None if total_shots is None or passed_shots is None else total_shots - passed_shots
        self.logical_errors = logical_errors
        self.observable_ones = observable_ones

    def __iter__(self):
        yield self.measurements
        yield self.detectors
        yield self.observables

    def __repr__(self):
        parts = [f"shots={self.measurements.shape[0]}", f"measurements={self.measurements.shape[1]}"]
        if self.total_shots is not None:
            parts.append(f"total_shots={self.total_shots}")
        if self.passed_shots is not None:
            parts.append(f"passed_shots={self.passed_shots}")
        return f"SampleResult({', '.join(parts)})"
