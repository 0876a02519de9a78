"""Program container: bytecode + constant pool in the reference layout.

`Program` holds exactly what `clifft::CompiledModule` holds
(src/clifft/backend/backend.h:255-344) as numpy arrays, exposes it to the
C ABI as a `clifft_program_view` (include/clifft_svm.h), and round-trips
through a small binary blob (the reference has no on-disk Program format;
`Program.as_dict()` there omits the constant pool, src/python/bindings.cc:767-782).

The instruction factories mirror `make_*` in src/clifft/backend/backend.h:172-211
(implemented in src/clifft/backend/backend.cc:28-296).
"""

from __future__ import annotations

import ctypes
import io
import struct
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# Opcodes (src/clifft/backend/backend.h:25-86, enum order)
# --------------------------------------------------------------------------
OPCODE_NAMES = [
    "OP_FRAME_CNOT", "OP_FRAME_CZ", "OP_FRAME_H", "OP_FRAME_S", "OP_FRAME_S_DAG",
    "OP_FRAME_SWAP", "OP_ARRAY_CNOT", "OP_ARRAY_CZ", "OP_ARRAY_SWAP",
    "OP_ARRAY_MULTI_CNOT", "OP_ARRAY_MULTI_CZ", "OP_ARRAY_H", "OP_ARRAY_S",
    "OP_ARRAY_S_DAG", "OP_ARRAY_T", "OP_ARRAY_T_DAG", "OP_ARRAY_ROT", "OP_ARRAY_U2",
    "OP_ARRAY_U4", "OP_EXPAND", "OP_EXPAND_T", "OP_EXPAND_T_DAG", "OP_EXPAND_ROT",
    "OP_MEAS_DORMANT_STATIC", "OP_MEAS_DORMANT_RANDOM", "OP_MEAS_ACTIVE_DIAGONAL",
    "OP_MEAS_ACTIVE_INTERFERE", "OP_SWAP_MEAS_INTERFERE",
    "OP_MEAS_DORMANT_STATIC_FORCED", "OP_MEAS_DORMANT_RANDOM_FORCED",
    "OP_MEAS_ACTIVE_DIAGONAL_FORCED", "OP_MEAS_ACTIVE_INTERFERE_FORCED",
    "OP_SWAP_MEAS_INTERFERE_FORCED", "OP_APPLY_PAULI", "OP_NOISE", "OP_NOISE_BLOCK",
    "OP_READOUT_NOISE", "OP_DETECTOR", "OP_POSTSELECT", "OP_OBSERVABLE", "OP_EXP_VAL",
]
NUM_OPCODES = len(OPCODE_NAMES)
assert NUM_OPCODES == 41


class Op:
    pass


for _i, _n in enumerate(OPCODE_NAMES):
    setattr(Op, _n[3:], _i)

FLAG_SIGN = 1
FLAG_HIDDEN = 2
FLAG_IDENTITY = 4
FLAG_EXPECTED_ONE = 8

INV_SQRT2 = 0.70710678118654752440

# Sampling opcode -> forced variant (src/clifft/api/record_probabilities.cc:96-118)
FORCED_OF = {
    Op.MEAS_DORMANT_STATIC: Op.MEAS_DORMANT_STATIC_FORCED,
    Op.MEAS_DORMANT_RANDOM: Op.MEAS_DORMANT_RANDOM_FORCED,
    Op.MEAS_ACTIVE_DIAGONAL: Op.MEAS_ACTIVE_DIAGONAL_FORCED,
    Op.MEAS_ACTIVE_INTERFERE: Op.MEAS_ACTIVE_INTERFERE_FORCED,
    Op.SWAP_MEAS_INTERFERE: Op.SWAP_MEAS_INTERFERE_FORCED,
}

ARRAY_OPS = frozenset(
    [Op.ARRAY_CNOT, Op.ARRAY_CZ, Op.ARRAY_SWAP, Op.ARRAY_MULTI_CNOT, Op.ARRAY_MULTI_CZ,
     Op.ARRAY_H, Op.ARRAY_S, Op.ARRAY_S_DAG, Op.ARRAY_T, Op.ARRAY_T_DAG, Op.ARRAY_ROT,
     Op.ARRAY_U2, Op.ARRAY_U4, Op.EXPAND, Op.EXPAND_T, Op.EXPAND_T_DAG, Op.EXPAND_ROT,
     Op.MEAS_ACTIVE_DIAGONAL, Op.MEAS_ACTIVE_INTERFERE, Op.SWAP_MEAS_INTERFERE,
     Op.MEAS_ACTIVE_DIAGONAL_FORCED, Op.MEAS_ACTIVE_INTERFERE_FORCED,
     Op.SWAP_MEAS_INTERFERE_FORCED, Op.EXP_VAL])

# --------------------------------------------------------------------------
# 32-byte Instruction record (backend.h:95-163): overlapping payload views.
# --------------------------------------------------------------------------
INSTR_DTYPE = np.dtype({
    "names": ["opcode", "reserved", "flags", "pad", "axis_1", "axis_2",
              "weight_re", "weight_im", "idx0", "idx1", "mask", "tail"],
    "formats": ["u1", "u1", "u1", "u1", "<u2", "<u2", "<f8", "<f8", "<u4", "<u4", "<u8", "<u8"],
    "offsets": [0, 1, 2, 3, 4, 6, 8, 16, 8, 12, 8, 24],
    "itemsize": 32,
})

U2_DTYPE = np.dtype({
    "names": ["matrices", "gamma", "out_states"],
    "formats": [("<c16", (4, 4)), ("<c16", (4,)), ("u1", (4,))],
    "offsets": [0, 256, 320],
    "itemsize": 328,
})
U4_ENTRY_DTYPE = np.dtype({
    "names": ["matrix", "gamma", "out_state"],
    "formats": [("<c16", (4, 4)), "<c16", "u1"],
    "offsets": [0, 256, 272],
    "itemsize": 280,
})
U4_DTYPE = np.dtype([("entries", U4_ENTRY_DTYPE, (16,))])
assert U2_DTYPE.itemsize == 328 and U4_DTYPE.itemsize == 4480


def _instr(opcode, axis_1=0, axis_2=0, flags=0):
    r = np.zeros(1, dtype=INSTR_DTYPE)[0]
    r["opcode"] = opcode
    r["axis_1"] = axis_1
    r["axis_2"] = axis_2
    r["flags"] = flags
    return r


def make_frame_cnot(c, t): return _instr(Op.FRAME_CNOT, c, t)
def make_frame_cz(a, b): return _instr(Op.FRAME_CZ, a, b)
def make_frame_h(v): return _instr(Op.FRAME_H, v)
def make_frame_s(v): return _instr(Op.FRAME_S, v)
def make_frame_s_dag(v): return _instr(Op.FRAME_S_DAG, v)
def make_frame_swap(a, b): return _instr(Op.FRAME_SWAP, a, b)
def make_array_cnot(c, t): return _instr(Op.ARRAY_CNOT, c, t)
def make_array_cz(a, b): return _instr(Op.ARRAY_CZ, a, b)
def make_array_swap(a, b): return _instr(Op.ARRAY_SWAP, a, b)
def make_array_h(v): return _instr(Op.ARRAY_H, v)
def make_array_s(v): return _instr(Op.ARRAY_S, v)
def make_array_s_dag(v): return _instr(Op.ARRAY_S_DAG, v)
def make_array_t(v): return _instr(Op.ARRAY_T, v)
def make_array_t_dag(v): return _instr(Op.ARRAY_T_DAG, v)
def make_expand(v): return _instr(Op.EXPAND, v)
def make_expand_t(v): return _instr(Op.EXPAND_T, v)
def make_expand_t_dag(v): return _instr(Op.EXPAND_T_DAG, v)


def make_array_multi_cnot(target, ctrl_mask):
    r = _instr(Op.ARRAY_MULTI_CNOT, target)
    r["mask"] = ctrl_mask
    return r


def make_array_multi_cz(control, target_mask):
    r = _instr(Op.ARRAY_MULTI_CZ, control)
    r["mask"] = target_mask
    return r


def make_array_rot(v, re, im):
    r = _instr(Op.ARRAY_ROT, v)
    r["weight_re"], r["weight_im"] = re, im
    return r


def make_expand_rot(v, re, im):
    r = _instr(Op.EXPAND_ROT, v)
    r["weight_re"], r["weight_im"] = re, im
    return r


def make_array_u2(axis, cp_idx):
    r = _instr(Op.ARRAY_U2, axis)
    r["idx0"] = cp_idx
    return r


def make_array_u4(lo, hi, cp_idx):
    r = _instr(Op.ARRAY_U4, lo, hi)
    r["idx0"] = cp_idx
    return r


def make_meas(opcode, axis, classical_idx, sign=False, flags=0):
    r = _instr(opcode, axis, flags=flags | (FLAG_SIGN if sign else 0))
    r["idx0"] = classical_idx
    return r


def make_swap_meas_interfere(f, t, classical_idx, sign=False):
    r = _instr(Op.SWAP_MEAS_INTERFERE, f, t, FLAG_SIGN if sign else 0)
    r["idx0"] = classical_idx
    return r


def make_apply_pauli(cp_mask_idx, condition_idx):
    r = _instr(Op.APPLY_PAULI)
    r["idx0"], r["idx1"] = cp_mask_idx, condition_idx
    return r


def make_noise(site):
    r = _instr(Op.NOISE)
    r["idx0"] = site
    return r


def make_noise_block(start, count):
    r = _instr(Op.NOISE_BLOCK)
    r["idx0"], r["idx1"] = start, count
    return r


def make_readout_noise(entry):
    r = _instr(Op.READOUT_NOISE)
    r["idx0"] = entry
    return r


def make_detector(list_idx, det_idx, expected_one=False):
    r = _instr(Op.DETECTOR, flags=FLAG_EXPECTED_ONE if expected_one else 0)
    r["idx0"], r["idx1"] = list_idx, det_idx
    return r


def make_postselect(list_idx, det_idx, expected_one=False):
    r = _instr(Op.POSTSELECT, flags=FLAG_EXPECTED_ONE if expected_one else 0)
    r["idx0"], r["idx1"] = list_idx, det_idx
    return r


def make_observable(list_idx, obs_idx):
    r = _instr(Op.OBSERVABLE)
    r["idx0"], r["idx1"] = list_idx, obs_idx
    return r


def make_exp_val(cp_idx, ev_idx):
    r = _instr(Op.EXP_VAL)
    r["idx0"], r["idx1"] = cp_idx, ev_idx
    return r


# --------------------------------------------------------------------------
# CLIFFT_SVM_ABI_VERSION of include/clifft_svm.h
ABI_VERSION = 2

# ctypes mirror of clifft_program_view
# --------------------------------------------------------------------------
_P = ctypes.c_void_p


class ProgramView(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_uint32), ("num_qubits", ctypes.c_uint32),
        ("peak_rank", ctypes.c_uint32), ("num_measurements", ctypes.c_uint32),
        ("total_meas_slots", ctypes.c_uint32), ("num_detectors", ctypes.c_uint32),
        ("num_observables", ctypes.c_uint32), ("num_exp_vals", ctypes.c_uint32),
        ("has_postselection", ctypes.c_uint32), ("words", ctypes.c_uint32),
        ("bytecode", _P), ("num_instructions", ctypes.c_uint64),
        ("pauli_x", _P), ("pauli_z", _P), ("pauli_sign", _P), ("num_pauli_masks", ctypes.c_uint32),
        ("ev_x", _P), ("ev_z", _P), ("ev_sign", _P), ("num_ev_masks", ctypes.c_uint32),
        ("site_chan_off", _P), ("num_sites", ctypes.c_uint32),
        ("chan_x", _P), ("chan_z", _P), ("chan_prob", _P), ("num_channels", ctypes.c_uint32),
        ("hazards", _P),
        ("ro_meas_idx", _P), ("ro_prob", _P), ("num_readout", ctypes.c_uint32),
        ("det_off", _P), ("det_tgt", _P), ("num_det_lists", ctypes.c_uint32),
        ("obs_off", _P), ("obs_tgt", _P), ("num_obs_lists", ctypes.c_uint32),
        ("u2_nodes", _P), ("num_u2", ctypes.c_uint32),
        ("u4_nodes", _P), ("num_u4", ctypes.c_uint32),
        ("expected_observables", _P), ("num_expected_observables", ctypes.c_uint32),
    ]


def _ptr(a):
    return None if a is None or a.size == 0 else a.ctypes.data


def _csr(lists, dtype=np.uint32):
    off = np.zeros(len(lists) + 1, dtype=np.uint32)
    for i, l in enumerate(lists):
        off[i + 1] = off[i] + len(l)
    tgt = np.array([t for l in lists for t in l], dtype=dtype)
    return off, tgt


def compute_hazards(site_chan_off, chan_prob):
    """Cumulative hazard table (src/clifft/backend/backend.cc:861-868)."""
    n = len(site_chan_off) - 1
    out = np.zeros(n, dtype=np.float64)
    acc = 0.0
    for s in range(n):
        psum = 0.0
        for c in range(int(site_chan_off[s]), int(site_chan_off[s + 1])):
            psum += float(chan_prob[c])
        acc += -np.log1p(-min(psum, 1.0 - 2.0 ** -53))
        out[s] = acc
    return out


@dataclass
class Program:
    """A compiled program (bytecode + constant pool + metadata)."""

    bytecode: np.ndarray
    num_qubits: int
    peak_rank: int
    num_measurements: int = 0
    total_meas_slots: int = 0
    num_detectors: int = 0
    num_observables: int = 0
    num_exp_vals: int = 0
    has_postselection: bool = False
    # constant pool (flattened)
    pauli_x: np.ndarray = None
    pauli_z: np.ndarray = None
    pauli_sign: np.ndarray = None
    ev_x: np.ndarray = None
    ev_z: np.ndarray = None
    ev_sign: np.ndarray = None
    site_chan_off: np.ndarray = None
    chan_x: np.ndarray = None
    chan_z: np.ndarray = None
    chan_prob: np.ndarray = None
    hazards: np.ndarray = None
    ro_meas_idx: np.ndarray = None
    ro_prob: np.ndarray = None
    det_off: np.ndarray = None
    det_tgt: np.ndarray = None
    obs_off: np.ndarray = None
    obs_tgt: np.ndarray = None
    u2_nodes: np.ndarray = None
    u4_nodes: np.ndarray = None
    expected_observables: np.ndarray = None
    global_weight: complex = 1.0 + 0.0j
    # parameter slots: instruction index of each parameterised rotation and
    # the sign s of its weight z = exp(i pi s alpha)
    param_instr: np.ndarray = None
    param_sign: np.ndarray = None
    _view: ProgramView = field(default=None, repr=False, compare=False)

    def __post_init__(self):
        if not isinstance(self.bytecode, np.ndarray) or self.bytecode.dtype != INSTR_DTYPE:
            self.bytecode = np.array(list(self.bytecode), dtype=INSTR_DTYPE)
        w = self.words
        if self.total_meas_slots < self.num_measurements:
            self.total_meas_slots = self.num_measurements

        def z(name, dtype, shape=(0,)):
            if getattr(self, name) is None:
                setattr(self, name, np.zeros(shape, dtype=dtype))
            else:
                setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=dtype))

        z("pauli_x", np.uint64)
        z("pauli_z", np.uint64)
        z("pauli_sign", np.uint8)
        z("ev_x", np.uint64)
        z("ev_z", np.uint64)
        z("ev_sign", np.uint8)
        z("site_chan_off", np.uint32, (1,))
        z("chan_x", np.uint64)
        z("chan_z", np.uint64)
        z("chan_prob", np.float64)
        if self.hazards is None:
            self.hazards = compute_hazards(self.site_chan_off, self.chan_prob)
        z("hazards", np.float64)
        z("ro_meas_idx", np.uint32)
        z("ro_prob", np.float64)
        z("det_off", np.uint32, (1,))
        z("det_tgt", np.uint32)
        z("obs_off", np.uint32, (1,))
        z("obs_tgt", np.uint32)
        if self.u2_nodes is None:
            self.u2_nodes = np.zeros(0, dtype=U2_DTYPE)
        if self.u4_nodes is None:
            self.u4_nodes = np.zeros(0, dtype=U4_DTYPE)
        z("expected_observables", np.uint8)
        z("param_instr", np.uint32)
        z("param_sign", np.int8)
        self.pauli_x = self.pauli_x.reshape(-1)
        self.pauli_z = self.pauli_z.reshape(-1)
        self.ev_x = self.ev_x.reshape(-1)
        self.ev_z = self.ev_z.reshape(-1)
        self.chan_x = self.chan_x.reshape(-1)
        self.chan_z = self.chan_z.reshape(-1)
        assert self.pauli_x.size == len(self.pauli_sign) * w
        assert self.ev_x.size == len(self.ev_sign) * w
        assert self.chan_x.size == len(self.chan_prob) * w
        assert len(self.hazards) == len(self.site_chan_off) - 1

    # ---------------------------------------------------------------- meta
    @property
    def words(self) -> int:
        return max((self.num_qubits + 63) // 64, 1) if self.num_qubits > 0 else max((self.peak_rank + 63) // 64, 1)

    @property
    def num_sites(self) -> int:
        return len(self.site_chan_off) - 1

    def __len__(self):
        return len(self.bytecode)

    def view(self) -> ProgramView:
        """The C-ABI view; arrays stay owned by this object."""
        v = ProgramView()
        v.abi_version = ABI_VERSION
        v.num_qubits = self.num_qubits
        v.peak_rank = self.peak_rank
        v.num_measurements = self.num_measurements
        v.total_meas_slots = self.total_meas_slots
        v.num_detectors = self.num_detectors
        v.num_observables = self.num_observables
        v.num_exp_vals = self.num_exp_vals
        v.has_postselection = int(bool(self.has_postselection))
        v.words = self.words
        v.bytecode = _ptr(self.bytecode)
        v.num_instructions = len(self.bytecode)
        v.pauli_x, v.pauli_z, v.pauli_sign = _ptr(self.pauli_x), _ptr(self.pauli_z), _ptr(self.pauli_sign)
        v.num_pauli_masks = len(self.pauli_sign)
        v.ev_x, v.ev_z, v.ev_sign = _ptr(self.ev_x), _ptr(self.ev_z), _ptr(self.ev_sign)
        v.num_ev_masks = len(self.ev_sign)
        v.site_chan_off = _ptr(self.site_chan_off)
        v.num_sites = self.num_sites
        v.chan_x, v.chan_z, v.chan_prob = _ptr(self.chan_x), _ptr(self.chan_z), _ptr(self.chan_prob)
        v.num_channels = len(self.chan_prob)
        v.hazards = _ptr(self.hazards)
        v.ro_meas_idx, v.ro_prob = _ptr(self.ro_meas_idx), _ptr(self.ro_prob)
        v.num_readout = len(self.ro_prob)
        v.det_off, v.det_tgt = _ptr(self.det_off), _ptr(self.det_tgt)
        v.num_det_lists = len(self.det_off) - 1
        v.obs_off, v.obs_tgt = _ptr(self.obs_off), _ptr(self.obs_tgt)
        v.num_obs_lists = len(self.obs_off) - 1
        v.u2_nodes, v.num_u2 = _ptr(self.u2_nodes), len(self.u2_nodes)
        v.u4_nodes, v.num_u4 = _ptr(self.u4_nodes), len(self.u4_nodes)
        v.expected_observables = _ptr(self.expected_observables)
        v.num_expected_observables = len(self.expected_observables)
        self._view = v
        return v

    def active_k_history(self) -> np.ndarray:
        """Static active rank before each instruction (the trajectory the
        reference records in SourceMap::active_k_history, source_map.h:88-98):
        EXPAND* raise k by one, active measurements lower it by one."""
        hist = np.zeros(len(self.bytecode), dtype=np.int32)
        k = 0
        ops = self.bytecode["opcode"]
        for i, op in enumerate(ops):
            hist[i] = k
            if Op.EXPAND <= op <= Op.EXPAND_ROT:
                k += 1
            elif op in (Op.MEAS_ACTIVE_DIAGONAL, Op.MEAS_ACTIVE_INTERFERE, Op.SWAP_MEAS_INTERFERE,
                        Op.MEAS_ACTIVE_DIAGONAL_FORCED, Op.MEAS_ACTIVE_INTERFERE_FORCED,
                        Op.SWAP_MEAS_INTERFERE_FORCED):
                k -= 1
        return hist

    def forced(self) -> "Program":
        """Copy with every measurement opcode replaced by its *_FORCED variant
        (src/clifft/api/record_probabilities.cc:96-118)."""
        bc = self.bytecode.copy()
        for src, dst in FORCED_OF.items():
            bc["opcode"][bc["opcode"] == src] = dst
        p = self.copy()
        p.bytecode = bc
        return p

    def copy(self) -> "Program":
        d = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()
             if k not in ("_view", "_device_program")}
        return Program(**d)

    # ------------------------------------------------------------ blob I/O
    _MAGIC = b"CLFTPRG1"
    _ARRAYS = ["bytecode", "pauli_x", "pauli_z", "pauli_sign", "ev_x", "ev_z", "ev_sign",
               "site_chan_off", "chan_x", "chan_z", "chan_prob", "hazards", "ro_meas_idx",
               "ro_prob", "det_off", "det_tgt", "obs_off", "obs_tgt", "u2_nodes", "u4_nodes",
               "expected_observables", "param_instr", "param_sign"]
    _SCALARS = ["num_qubits", "peak_rank", "num_measurements", "total_meas_slots",
                "num_detectors", "num_observables", "num_exp_vals", "has_postselection"]

    def to_bytes(self) -> bytes:
        """Binary blob: magic, 8 u32 scalars, global weight, then each array
        as (u64 byte length, raw little-endian bytes) in a fixed order."""
        out = io.BytesIO()
        out.write(self._MAGIC)
        out.write(struct.pack("<8I", *[int(getattr(self, s)) for s in self._SCALARS]))
        out.write(struct.pack("<2d", complex(self.global_weight).real, complex(self.global_weight).imag))
        for name in self._ARRAYS:
            b = np.ascontiguousarray(getattr(self, name)).tobytes()
            out.write(struct.pack("<Q", len(b)))
            out.write(b)
        return out.getvalue()

    @classmethod
    def from_bytes(cls, blob: bytes) -> "Program":
        if blob[:8] != cls._MAGIC:
            raise ValueError("not a Program blob")
        pos = 8
        scal = struct.unpack_from("<8I", blob, pos)
        pos += 32
        gw = struct.unpack_from("<2d", blob, pos)
        pos += 16
        dtypes = {"bytecode": INSTR_DTYPE, "u2_nodes": U2_DTYPE, "u4_nodes": U4_DTYPE,
                  "pauli_x": np.uint64, "pauli_z": np.uint64, "ev_x": np.uint64, "ev_z": np.uint64,
                  "chan_x": np.uint64, "chan_z": np.uint64, "pauli_sign": np.uint8,
                  "ev_sign": np.uint8, "expected_observables": np.uint8, "chan_prob": np.float64,
                  "hazards": np.float64, "ro_prob": np.float64, "param_sign": np.int8}
        arrays = {}
        for name in cls._ARRAYS:
            (n,) = struct.unpack_from("<Q", blob, pos)
            pos += 8
            arrays[name] = np.frombuffer(blob[pos:pos + n], dtype=dtypes.get(name, np.uint32)).copy()
            pos += n
        kw = dict(zip(cls._SCALARS, scal))
        kw["has_postselection"] = bool(kw["has_postselection"])
        return cls(global_weight=complex(*gw), **kw, **arrays)


def build_program(bytecode, peak_rank, num_qubits=0, num_measurements=0, num_detectors=0,
                  num_observables=0, num_exp_vals=0, total_meas_slots=None, pauli_masks=(),
                  exp_val_masks=(), noise_sites=(), readout_noise=(), detector_targets=(),
                  observable_targets=(), u2_nodes=None, u4_nodes=None, expected_observables=(),
                  has_postselection=None):
    """Convenience constructor from Python lists, like the reference tests'
    make_program (tests/test_svm.cc:28-37).

    pauli_masks / exp_val_masks: list of (x_bits:int, z_bits:int, sign:bool)
    noise_sites: list of sites, each a list of (x_bits, z_bits, prob)
    readout_noise: list of (meas_idx, prob)
    """
    nq = num_qubits if num_qubits > 0 else max(peak_rank, 1)
    words = (nq + 63) // 64

    def split(bits):
        return [(int(bits) >> (64 * w)) & 0xFFFFFFFFFFFFFFFF for w in range(words)]

    def arena(masks):
        x = np.array([split(m[0]) for m in masks], dtype=np.uint64).reshape(-1)
        z = np.array([split(m[1]) for m in masks], dtype=np.uint64).reshape(-1)
        s = np.array([1 if m[2] else 0 for m in masks], dtype=np.uint8)
        return x, z, s

    px, pz, ps = arena(pauli_masks)
    ex, ez, es = arena(exp_val_masks)
    off = [0]
    chans = []
    for site in noise_sites:
        chans.extend(site)
        off.append(len(chans))
    cx, cz, _ = arena([(c[0], c[1], 0) for c in chans])
    cprob = np.array([c[2] for c in chans], dtype=np.float64)
    det_off, det_tgt = _csr(detector_targets)
    obs_off, obs_tgt = _csr(observable_targets)
    bc = np.array(list(bytecode), dtype=INSTR_DTYPE) if not isinstance(bytecode, np.ndarray) else bytecode
    if has_postselection is None:
        has_postselection = bool(np.any(bc["opcode"] == Op.POSTSELECT)) if len(bc) else False
    return Program(
        bytecode=bc, num_qubits=num_qubits, peak_rank=peak_rank,
        num_measurements=num_measurements,
        total_meas_slots=total_meas_slots if total_meas_slots is not None else num_measurements,
        num_detectors=num_detectors, num_observables=num_observables, num_exp_vals=num_exp_vals,
        has_postselection=has_postselection,
        pauli_x=px, pauli_z=pz, pauli_sign=ps, ev_x=ex, ev_z=ez, ev_sign=es,
        site_chan_off=np.array(off, dtype=np.uint32), chan_x=cx, chan_z=cz, chan_prob=cprob,
        ro_meas_idx=np.array([r[0] for r in readout_noise], dtype=np.uint32),
        ro_prob=np.array([r[1] for r in readout_noise], dtype=np.float64),
        det_off=det_off, det_tgt=det_tgt, obs_off=obs_off, obs_tgt=obs_tgt,
        u2_nodes=u2_nodes, u4_nodes=u4_nodes,
        expected_observables=np.array(list(expected_observables), dtype=np.uint8),
    )
