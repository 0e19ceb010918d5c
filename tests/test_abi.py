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


_OOM_CHILD = r"""
import ctypes, resource, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from echidna_amd import backend
lib = backend.load_library()
n = 1 << 21
# x0 * x0 * ... : a chain of 2M products (valid tape, about 60 MB of host
# program state to compile)
op = np.full(n, 4, np.uint32); op[0] = 0
a0 = np.arange(-1, n - 1, dtype=np.int64).astype(np.uint32); a0[0] = 0xFFFFFFFF
a1 = np.zeros(n, np.uint32); a1[0] = 0xFFFFFFFF
c = np.zeros(n)
out5 = (ctypes.c_uint64 * 5)()
P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
vm = [int(l.split()[1]) for l in open("/proc/self/status") if l.startswith("VmSize")][0] * 1024
resource.setrlimit(resource.RLIMIT_AS, (vm + (24 << 20), resource.getrlimit(resource.RLIMIT_AS)[1]))
rc = lib.ehip_compile_program(0, P(op, ctypes.c_uint32), P(a0, ctypes.c_uint32), P(a1, ctypes.c_uint32),
                              P(c, ctypes.c_double), n, 1, n - 1, None, 0, 1, 1, None, 0, None, 0, None, 0, out5)
resource.setrlimit(resource.RLIMIT_AS, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
print(rc, lib.ehip_last_error().decode(), flush=True)
"""


def test_host_allocation_failure_is_a_status_not_an_abort():
    """A C++ exception never crosses the C ABI: std::bad_alloc while compiling
    a program under a tight address-space limit comes back as status 3
    (GpuError::OutOfMemory, src/gpu/mod.rs:276-303), not std::terminate."""
    import sys as _sys
    r = subprocess.run([_sys.executable, "-c", _OOM_CHILD, ROOT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    rc, msg = r.stdout.strip().split(" ", 1)
    assert int(rc) == backend.GpuError.OUT_OF_MEMORY, r.stdout
    assert "out of memory" in msg
