"""The C-ABI library loads and exports every symbol include/echidna_hip.h
declares; host-side argument checks work without a device."""

import os
import re
import subprocess

import numpy as np
import pytest

from echidna_amd import backend
from echidna_amd.tape import GpuError, GpuTapeData, record

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "echidna_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ehip_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = backend.load_library()
    declared = header_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", backend.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ehip_\w+)", nm))
    assert set(declared) <= exported
    assert set(backend.exported_symbols()) == set(declared)


def test_library_is_built_for_gfx950():
    blob = open(backend.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_no_device_on_cpu_host():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert backend.HipContext.new(0) is None


def test_compile_rejects_invalid_and_custom():
    tape, _ = record(lambda v: v[0] * 3.0, [2.0])
    d = GpuTapeData.from_tape_f64(tape)
    backend.compile_program(d)
    d.arg0[2] = 9999
    with pytest.raises(GpuError, match="invalid GpuTapeData"):
        backend.compile_program(d)
    d = GpuTapeData.from_tape_f64(tape)
    d.opcodes[2] = 43
    with pytest.raises(GpuError) as e:
        backend.compile_program(d)
    assert e.value.code == GpuError.CUSTOM_OPS_NOT_SUPPORTED


def test_python_validate_matches_reference_messages():
    tape, _ = record(lambda v: v[0] * 3.0, [2.0])
    d = GpuTapeData.from_tape(tape)
    d.validate()
    d.opcodes[2] = 24
    d.arg1[2] = 5
    with pytest.raises(GpuError, match="should be the UNUSED sentinel"):
        d.validate()
