"""MI355X (gfx950) executor for Clifft's Schrödinger VM.

Package-level names follow the reference's Python API for the sampling
path (src/python/clifft/__init__.py:57-111, 273-316, 377-425):

    compile(text)                                  -> Program
    sample(program, shots, seed=None)              -> SampleResult
    sample_survivors(program, shots, seed=None, keep_records=False)
    sample_k(program, shots, k, seed=None)
    sample_k_survivors(program, shots, k, seed=None, keep_records=False)
    record_probabilities(program, records, *, return_log=False)
    State, execute(program, state), svm_backend()

Additions with no reference counterpart: batch (many programs in one
persistent launch) and grad (adjoint gradients, torch.autograd.Function).
The native library is loaded on first use; nothing here runs on the CPU in
its place.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

from .program import Program
from .sample_result import SampleResult

__all__ = ["Program", "SampleResult", "State", "compile", "execute", "record_probabilities", "sample", "sample_k",
           "sample_k_survivors", "sample_survivors", "svm_backend", "noise_site_probabilities"]


def compile(text: str, **kw) -> Program:  # noqa: A001 - the reference's name
    from .compiler import compile as _compile
    return _compile(text, **kw)


def sample(program, shots: int, seed=None) -> SampleResult:
    from . import svm
    return svm.sample(program, shots, seed)


def sample_survivors(program, shots: int, seed=None, keep_records: bool = False) -> SampleResult:
    from . import svm
    return svm.sample_survivors(program, shots, seed, keep_records=keep_records)


def sample_k(program, shots: int, k: int, seed=None) -> SampleResult:
    from . import svm
    if program.has_postselection:
        raise ValueError("sample_k() cannot be used with post-selected programs because it returns a fixed number "
                         "of rows and cannot discard shots. Use sample_k_survivors() instead.")
    return svm.sample_k(program, shots, k, seed)


def sample_k_survivors(program, shots: int, k: int, seed=None, keep_records: bool = False) -> SampleResult:
    from . import svm
    return svm.sample_k_survivors(program, shots, k, seed, keep_records=keep_records)


def noise_site_probabilities(program) -> np.ndarray:
    from . import svm
    return svm.noise_site_probabilities(program)


def svm_backend() -> str:
    from . import _lib
    return _lib.backend()


def execute(program, state) -> None:
    from . import svm
    svm.execute(program, state)


def __getattr__(name):
    if name == "State":
        from .svm import State
        return State
    raise AttributeError(name)


def _records_array(program, records) -> np.ndarray:
    """Records as a [num_records, num_measurements] array of 0/1 bytes, from
    one '0'/'1' string, a sequence of them, or a 2-D bool/uint8 array."""
    n = program.num_measurements

    def from_strings(rows):
        out = np.zeros((len(rows), n), np.uint8)
        for i, r in enumerate(rows):
            if len(r) != n:
                raise ValueError(f"record at index {i} has length {len(r)}, expected {n}")
            bad = set(r) - {"0", "1"}
            if bad:
                raise ValueError(f"record at index {i} contains {sorted(bad)[0]!r}; expected only '0' and '1'")
            out[i] = np.frombuffer(r.encode(), np.uint8) - ord("0")
        return out

    if isinstance(records, str):
        return from_strings([records])
    if isinstance(records, np.ndarray):
        if records.ndim != 2:
            raise ValueError("records array must be 2D with shape (num_records, num_measurements)")
        if records.shape[1] != n:
            raise ValueError(f"records array has {records.shape[1]} columns, expected {n}")
        if records.dtype not in (np.dtype(bool), np.dtype(np.uint8)):
            raise TypeError("records array dtype must be bool or uint8")
        if records.dtype == np.uint8 and np.any(records > 1):
            raise ValueError("uint8 records array must contain only 0 and 1")
        return np.ascontiguousarray(records.astype(np.uint8, copy=False))
    if isinstance(records, Sequence) and all(isinstance(r, str) for r in records):
        return from_strings(list(records))
    raise TypeError("records must be strings or a 2D bool/uint8 NumPy array")


def record_probabilities(program, records, *, return_log: bool = False) -> np.ndarray:
    """Exact probability (or natural log, with return_log) that sample()
    emits each record; unreachable records give 0.0 (-inf)."""
    from . import svm
    if program.num_measurements == 0:
        raise ValueError("record_probabilities() requires a program with at least one measurement; use "
                         "clifft.basis_probabilities() for unitary circuits.")
    logp = svm.record_probabilities(program, _records_array(program, records))
    if return_log:
        return np.where(logp == np.finfo(np.float64).min, -np.inf, logp)
    return np.exp(logp)
