"""ctypes binding of libclifft_svm.so (the C ABI in include/clifft_svm.h).

There is no CPU fallback: if the library is missing or cannot find a gfx950
device, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclifft_svm.so")

REC_DEVICE = 1
REC_PACKED = 2
BATCH_PER_PROGRAM = 0x100

E_INVALID, E_UNSUPPORTED, E_HIP, E_NOMEM, E_NODEVICE = -1, -2, -3, -4, -5

_lib = None


class SvmError(RuntimeError):
    pass


class Records(ctypes.Structure):
    _fields_ = [("measurements", ctypes.c_void_p), ("detectors", ctypes.c_void_p),
                ("observables", ctypes.c_void_p), ("exp_vals", ctypes.c_void_p),
                ("discarded", ctypes.c_void_p)]


class Survivors(ctypes.Structure):
    _fields_ = [("total_shots", ctypes.c_uint64), ("passed_shots", ctypes.c_uint64),
                ("logical_errors", ctypes.c_uint64), ("observable_ones", ctypes.c_void_p)]


class ProgramInfo(ctypes.Structure):
    _fields_ = [("peak_rank", ctypes.c_uint32), ("tier", ctypes.c_uint32),
                ("team_waves", ctypes.c_uint32), ("shots_per_block", ctypes.c_uint32),
                ("num_array_ops", ctypes.c_uint64), ("bytes_per_shot", ctypes.c_double)]


class State(ctypes.Structure):
    _fields_ = [("v", ctypes.c_void_p), ("px", ctypes.c_void_p), ("pz", ctypes.c_void_p),
                ("gamma_re", ctypes.c_double), ("gamma_im", ctypes.c_double),
                ("active_k", ctypes.c_uint32), ("discarded", ctypes.c_uint32),
                ("meas_record", ctypes.c_void_p), ("det_record", ctypes.c_void_p),
                ("obs_record", ctypes.c_void_p), ("exp_vals", ctypes.c_void_p),
                ("rng", ctypes.c_uint64 * 4), ("next_noise_idx", ctypes.c_uint32),
                ("forced_reachable", ctypes.c_uint32), ("forced_log_probability", ctypes.c_double),
                ("forced_record", ctypes.c_void_p), ("dust_clamps", ctypes.c_uint64)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("global_threads", ctypes.c_uint32), ("wave_rank_max", ctypes.c_uint32),
                ("lds_rank_max", ctypes.c_uint32), ("wave_chunk", ctypes.c_uint32),
                ("block_chunk", ctypes.c_uint32), ("concurrent", ctypes.c_uint32),
                ("lane_rank_max", ctypes.c_int32), ("lane_chunk", ctypes.c_uint32),
                ("gwave_rank_max", ctypes.c_uint32)]


class BatchGroupInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("programs", ctypes.c_uint32), ("items", ctypes.c_uint32),
                ("grid", ctypes.c_uint32), ("block", ctypes.c_uint32), ("lds_bytes", ctypes.c_uint32),
                ("amp_elems", ctypes.c_uint32), ("amp_slices", ctypes.c_uint32),
                ("algorithmic_bytes", ctypes.c_double)]


# name -> (restype, argtypes); exactly the functions include/clifft_svm.h declares
SIGNATURES = {
    "clifft_svm_abi_version": (ctypes.c_uint32, []),
    "clifft_svm_backend": (ctypes.c_char_p, []),
    "clifft_svm_last_error": (ctypes.c_char_p, []),
    "clifft_svm_program_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "clifft_svm_program_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "clifft_svm_program_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ProgramInfo)]),
    "clifft_svm_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                         ctypes.POINTER(Records), ctypes.c_void_p]),
    "clifft_svm_sample_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                                 ctypes.POINTER(Records), ctypes.c_void_p]),
    "clifft_svm_sample_survivors": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                                   ctypes.c_uint32, ctypes.POINTER(Survivors),
                                                   ctypes.POINTER(Records), ctypes.c_void_p]),
    "clifft_svm_sample_k": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                           ctypes.c_uint32, ctypes.POINTER(Records), ctypes.c_void_p]),
    "clifft_svm_sample_k_survivors": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                                     ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(Survivors),
                                                     ctypes.POINTER(Records), ctypes.c_void_p]),
    "clifft_svm_noise_site_probabilities": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                           ctypes.POINTER(ctypes.c_uint32)]),
    "clifft_svm_record_log_probs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "clifft_svm_execute": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(State)]),
    "clifft_svm_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_uint32,
                                               ctypes.c_uint32, ctypes.POINTER(Records),
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "clifft_svm_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "clifft_svm_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchTuning)]),
    "clifft_svm_batch_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchTuning)]),
    "clifft_svm_batch_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(BatchGroupInfo),
                                             ctypes.POINTER(ctypes.c_uint32)]),
    "clifft_svm_batch_time_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "clifft_svm_batch_group_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                                    ctypes.POINTER(ctypes.c_uint32)]),
    "clifft_svm_batch_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                             ctypes.POINTER(BatchTuning)]),
    "clifft_svm_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "clifft_svm_batch_run_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "clifft_svm_adjoint_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                 ctypes.POINTER(ctypes.c_void_p)]),
    "clifft_svm_adjoint_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "clifft_svm_adjoint_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "clifft_svm_batch_sample": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                               ctypes.POINTER(Records), ctypes.c_void_p]),
}


def lib():
    """Load the native library (raises if it is not built).

    torch is imported first when it is installed: the HIP runtime it ships
    and the system one share the soname libamdhip64.so.7, so loading torch's
    first makes this library bind to the same runtime.  Tensors, streams and
    events then mean the same thing on both sides; loaded the other way
    round, the process would hold two HIP runtimes and torch would see no
    device."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SvmError(f"{LIB_PATH} is not built; run `python -m clifft_amd.build`")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().clifft_svm_last_error().decode(errors="replace")
        raise SvmError(f"{what or 'clifft_svm'} failed ({rc}): {msg}")


def backend() -> str:
    return lib().clifft_svm_backend().decode()
