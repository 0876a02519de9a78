"""Package-level API shapes and argument handling (CPU only: the calls that
would run on the device are not made)."""

from __future__ import annotations

import numpy as np
import pytest

import clifft_amd


def test_exports():
    for name in clifft_amd.__all__:
        assert hasattr(clifft_amd, name), name


def test_records_parsing():
    from clifft_amd import _records_array
    p = clifft_amd.compile("H 0\nM 0\nH 1\nM 1")
    assert _records_array(p, "01").tolist() == [[0, 1]]
    assert _records_array(p, ["00", "11"]).tolist() == [[0, 0], [1, 1]]
    assert _records_array(p, np.array([[True, False]])).tolist() == [[1, 0]]
    with pytest.raises(ValueError, match="length"):
        _records_array(p, "0")
    with pytest.raises(ValueError, match="expected only"):
        _records_array(p, "0x")
    with pytest.raises(TypeError):
        _records_array(p, np.zeros((1, 2), np.int64))
    with pytest.raises(ValueError, match="only 0 and 1"):
        _records_array(p, np.full((1, 2), 2, np.uint8))
    with pytest.raises(TypeError):
        _records_array(p, [1, 0])


def test_unitary_program_rejected_before_device():
    p = clifft_amd.compile("H 0\nT 0\nEXP_VAL X0")
    with pytest.raises(ValueError, match="at least one measurement"):
        clifft_amd.record_probabilities(p, np.zeros((1, 0), np.uint8))
