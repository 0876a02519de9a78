"""The C-ABI library loads and exports every entry point include/*.h
declares (no device calls: runs without a GPU)."""

import ctypes
import glob
import os
import re

import numpy as np

import clifft_amd._lib as L
from clifft_amd.program import INSTR_DTYPE, U2_DTYPE, U4_DTYPE, ProgramView

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(clifft_\w+)\s*\(", text))


def test_every_header_is_covered():
    headers = {os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))}
    assert headers == {"clifft_svm.h", "clifft_compile.h"}


def test_svm_library_exports_every_declared_symbol():
    lib = L.lib()
    decl = declared_functions("clifft_svm.h")
    assert decl, "no declarations found"
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert decl == set(L.SIGNATURES), "ctypes signature table out of sync with the header"


def test_compile_library_exports_every_declared_symbol():
    from clifft_amd import compiler as C
    lib = C.lib()
    decl = declared_functions("clifft_compile.h")
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert decl and not missing, missing


def test_abi_version_and_layouts():
    from clifft_amd.program import ABI_VERSION
    assert L.lib().clifft_svm_abi_version() == ABI_VERSION == 2
    assert ctypes.sizeof(L.BatchTuning) == 36 and ctypes.sizeof(L.BatchGroupInfo) == 40
    # reference layouts: Instruction 32 B (backend.h:163), FusedU2Node 328 B, FusedU4Node 4480 B
    assert INSTR_DTYPE.itemsize == 32 and U2_DTYPE.itemsize == 328 and U4_DTYPE.itemsize == 4480
    offs = {n: INSTR_DTYPE.fields[n][1] for n in INSTR_DTYPE.names}
    assert (offs["opcode"], offs["flags"], offs["axis_1"], offs["axis_2"], offs["weight_re"], offs["idx1"]) == \
        (0, 2, 4, 6, 8, 12)
    assert ctypes.sizeof(L.State) == 8 * 3 + 16 + 8 + 8 * 4 + 32 + 8 + 8 + 8 + 8


def test_last_error_is_empty_string_initially():
    assert isinstance(L.lib().clifft_svm_last_error(), bytes)
