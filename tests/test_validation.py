"""Host-side program validation rejects malformed bytecode before any
device work (runs without a GPU: validation precedes device selection)."""

import ctypes

import numpy as np
import pytest

import clifft_amd._lib as L
from clifft_amd.program import (Op, build_program, make_array_cnot, make_array_h, make_array_u4, make_expand,
                                make_meas, make_noise_block)


def create(prog):
    h = ctypes.c_void_p()
    v = prog.view()
    rc = L.lib().clifft_svm_program_create(ctypes.byref(v), ctypes.byref(h))
    if rc == 0:
        L.lib().clifft_svm_program_destroy(h)
    return rc, L.lib().clifft_svm_last_error().decode()


@pytest.mark.parametrize("bc,peak,what", [
    ([make_expand(1)], 2, "EXPAND must target axis k"),
    ([make_expand(0), make_expand(1), make_expand(2)], 2, "exceeds peak_rank"),
    ([make_expand(0), make_array_h(1)], 2, "not active"),
    ([make_expand(0), make_array_cnot(0, 0)], 2, "distinct active axes"),
    ([make_expand(0), make_expand(1), make_array_u4(1, 0, 0)], 2, "U4"),
    ([make_meas(Op.MEAS_ACTIVE_DIAGONAL, 0, 0)], 2, "k-1"),
    ([make_noise_block(0, 3)], 2, "noise block"),
])
def test_rejects_malformed(bc, peak, what):
    p = build_program(bc, peak_rank=peak, num_qubits=4, num_measurements=1)
    rc, msg = create(p)
    assert rc == L.E_INVALID and what in msg, (rc, msg)


def test_rejects_bad_hazards():
    p = build_program([make_noise_block(0, 2)], peak_rank=1, num_qubits=2,
                      noise_sites=[[(1, 0, 0.1)], [(1, 0, 0.1)]])
    p.hazards = np.array([0.2, 0.1])  # decreasing
    rc, msg = create(p)
    assert rc == L.E_INVALID and "hazard" in msg


def test_rejects_bad_record_index():
    p = build_program([make_meas(Op.MEAS_DORMANT_STATIC, 0, 5)], peak_rank=1, num_qubits=2, num_measurements=1)
    rc, msg = create(p)
    assert rc == L.E_INVALID and "record index" in msg


def test_valid_program_passes_validation():
    p = build_program([make_expand(0), make_array_h(0)], peak_rank=1, num_qubits=2)
    rc, msg = create(p)
    # with no GPU the call reaches device selection and reports that instead
    assert rc in (0, L.E_NODEVICE), (rc, msg)


def test_rejects_peak_rank_beyond_32bit_indices():
    # 32-bit amplitude indices: peak 30 is the largest accepted rank
    for peak, ok in ((30, True), (31, False), (40, False)):
        p = build_program([make_expand(0), make_array_h(0)], peak_rank=peak, num_qubits=48)
        rc, msg = create(p)
        if ok:
            assert rc in (0, L.E_NODEVICE), (rc, msg)
        else:
            assert rc == L.E_UNSUPPORTED and "peak_rank <= 30" in msg, (rc, msg)


def test_device_program_freed_with_program(monkeypatch):
    """device_program() keeps its upload on the Program and refers back to
    it weakly: dropping the Program destroys the device handle."""
    import gc

    from clifft_amd import svm

    created, destroyed = [], []

    class FakeLib:
        def clifft_svm_program_create(self, view, out):
            out._obj.value = 1000 + len(created)
            created.append(out._obj.value)
            return 0

        def clifft_svm_program_destroy(self, h):
            destroyed.append(h.value)
            return 0

    fake = FakeLib()
    monkeypatch.setattr(svm._lib, "lib", lambda: fake)
    p = build_program([make_expand(0), make_array_h(0)], peak_rank=1, num_qubits=2)
    dp = svm.device_program(p)
    assert svm.device_program(p) is dp and dp.program is p and created == [1000]
    del dp
    assert destroyed == []  # still reachable through the program
    del p
    gc.collect()
    assert destroyed == [1000]
    # a DeviceProgram made directly keeps its program alive
    q = build_program([make_expand(0)], peak_rank=1, num_qubits=2)
    d2 = svm.DeviceProgram(q)
    del q
    gc.collect()
    assert d2.program.peak_rank == 1 and destroyed == [1000]
