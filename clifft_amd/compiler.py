"""Program producer: Stim-format circuit text -> Program.

Python face of include/clifft_compile.h (libclifft_compile.so).  Mirrors the
call shape of the reference `clifft.compile(text)`
(src/python/clifft/__init__.py:324-374) for the supported gate subset; the
emitted bytecode is this build's own lowering, not the reference's.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .program import INSTR_DTYPE, U2_DTYPE, U4_DTYPE, Program, ProgramView

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclifft_compile.so")
SRC = os.path.join(_HERE, "csrc", "compiler.cpp")
INCLUDE = os.path.join(os.path.dirname(_HERE), "include")

NO_FUSION, KEEP_ROTATIONS, NO_PASSES = 1, 2, 4

_lib = None


def build() -> str:
    tmp = LIB_PATH + ".tmp"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, "-o", tmp, SRC], check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m clifft_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.clifft_compile.restype = ctypes.c_int
        L.clifft_compile.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        L.clifft_compiled_view.restype = ctypes.c_int
        L.clifft_compiled_view.argtypes = [ctypes.c_void_p, ctypes.POINTER(ProgramView)]
        L.clifft_compiled_num_params.restype = ctypes.c_uint32
        L.clifft_compiled_num_params.argtypes = [ctypes.c_void_p]
        L.clifft_compiled_param_instr.restype = ctypes.POINTER(ctypes.c_uint32)
        L.clifft_compiled_param_instr.argtypes = [ctypes.c_void_p]
        L.clifft_compiled_param_sign.restype = ctypes.POINTER(ctypes.c_int8)
        L.clifft_compiled_param_sign.argtypes = [ctypes.c_void_p]
        L.clifft_compiled_global_weight.restype = None
        L.clifft_compiled_global_weight.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                                    ctypes.POINTER(ctypes.c_double)]
        L.clifft_compiled_destroy.restype = None
        L.clifft_compiled_destroy.argtypes = [ctypes.c_void_p]
        L.clifft_compile_last_error.restype = ctypes.c_char_p
        L.clifft_compile_last_error.argtypes = []
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    nbytes = n * np.dtype(dtype).itemsize
    buf = (ctypes.c_char * nbytes).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def compile(text: str, *, fusion: bool = True, keep_rotations: bool = False, passes: bool = True) -> Program:
    """Compile circuit text into a Program.

    keep_rotations: never fuse R_Z into U2 nodes, so each R_Z stays one
    ARRAY_ROT / EXPAND_ROT instruction (its index is in program.param_instr,
    its weight sign in program.param_sign) -- required for gradients."""
    L = lib()
    opts = (0 if fusion else NO_FUSION) | (KEEP_ROTATIONS if keep_rotations else 0) | (0 if passes else NO_PASSES)
    h = ctypes.c_void_p()
    rc = L.clifft_compile(text.encode(), opts, ctypes.byref(h))
    if rc != 0:
        raise ValueError(L.clifft_compile_last_error().decode())
    try:
        v = ProgramView()
        if L.clifft_compiled_view(h, ctypes.byref(v)) != 0:
            raise RuntimeError(L.clifft_compile_last_error().decode())
        W = v.words
        nsites = v.num_sites
        prog = Program(
            bytecode=_arr(v.bytecode, v.num_instructions, INSTR_DTYPE),
            num_qubits=v.num_qubits, peak_rank=v.peak_rank, num_measurements=v.num_measurements,
            total_meas_slots=v.total_meas_slots, num_detectors=v.num_detectors,
            num_observables=v.num_observables, num_exp_vals=v.num_exp_vals,
            pauli_x=_arr(v.pauli_x, v.num_pauli_masks * W, np.uint64),
            pauli_z=_arr(v.pauli_z, v.num_pauli_masks * W, np.uint64),
            pauli_sign=_arr(v.pauli_sign, v.num_pauli_masks, np.uint8),
            ev_x=_arr(v.ev_x, v.num_ev_masks * W, np.uint64), ev_z=_arr(v.ev_z, v.num_ev_masks * W, np.uint64),
            ev_sign=_arr(v.ev_sign, v.num_ev_masks, np.uint8),
            site_chan_off=_arr(v.site_chan_off, nsites + 1, np.uint32),
            chan_x=_arr(v.chan_x, v.num_channels * W, np.uint64),
            chan_z=_arr(v.chan_z, v.num_channels * W, np.uint64),
            chan_prob=_arr(v.chan_prob, v.num_channels, np.float64),
            hazards=_arr(v.hazards, nsites, np.float64),
            det_off=_arr(v.det_off, v.num_det_lists + 1, np.uint32),
            det_tgt=_arr(v.det_tgt, int(_arr(v.det_off, v.num_det_lists + 1, np.uint32)[-1]), np.uint32),
            obs_off=_arr(v.obs_off, v.num_obs_lists + 1, np.uint32),
            obs_tgt=_arr(v.obs_tgt, int(_arr(v.obs_off, v.num_obs_lists + 1, np.uint32)[-1]), np.uint32),
            u2_nodes=_arr(v.u2_nodes, v.num_u2, U2_DTYPE),
            u4_nodes=_arr(v.u4_nodes, v.num_u4, U4_DTYPE),
        )
        npar = L.clifft_compiled_num_params(h)
        prog.param_instr = _arr(ctypes.cast(L.clifft_compiled_param_instr(h), ctypes.c_void_p).value, npar, np.uint32)
        prog.param_sign = _arr(ctypes.cast(L.clifft_compiled_param_sign(h), ctypes.c_void_p).value, npar, np.int8)
        re, im = ctypes.c_double(), ctypes.c_double()
        L.clifft_compiled_global_weight(h, ctypes.byref(re), ctypes.byref(im))
        prog.global_weight = complex(re.value, im.value)
        return prog
    finally:
        L.clifft_compiled_destroy(h)
