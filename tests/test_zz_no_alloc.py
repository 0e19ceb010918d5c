"""The allocation-free fwd+grad entry point (ehip_gradient_batch_dev_noalloc)
against the allocating one (ehip_gradient_batch_dev), bit for bit, on two
calls in a row, the second with a smaller batch: with the register file in
HBM (the fixed 64 MiB scratch; the first batch needs several launches of
it) and with the tape-specialised kernel (no register file), f32 and f64."""

import numpy as np
import pytest
import torch

from echidna_amd import workloads as wl
from echidna_amd.backend import TUNE_GRADIENT, HipContext, load_library
from echidna_amd.tape import GpuError, GpuTapeData

pytestmark = pytest.mark.gpu


def _bits(t):
    return t.cpu().numpy().view(np.uint8)


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("shape", ["hbm_register_file", "specialised"])
def test_noalloc_matches_allocating_entry_point(hip, f64, shape):
    assert load_library().ehip_noalloc_scratch_bytes() == 64 << 20
    n = 32
    tape = wl.tanh_mlp(n, 64, 1)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    if shape == "specialised":
        HipContext.specialize(bufs, TUNE_GRADIENT)
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 0, fetch=2)
    else:
        # 4 KiB of LDS: almost every register of the 96 in the HBM file
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 4096, fetch=1)
    tdt = torch.float64 if f64 else torch.float32
    for seed, B in ((1, 300_000), (2, 1_001)):
        x = torch.from_numpy(wl.points_normal(seed, B, n)).to(tdt).cuda()
        ry, rg = torch.empty(B, dtype=tdt, device="cuda"), torch.empty(B, n, dtype=tdt, device="cuda")
        hip.gradient_batch_dev(bufs, x, B, ry, rg)
        y = torch.full((B,), float("nan"), dtype=tdt, device="cuda")
        g = torch.full((B, n), float("nan"), dtype=tdt, device="cuda")
        hip.gradient_batch_dev_noalloc(bufs, x, B, y, g)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(y), _bits(ry)) and np.array_equal(_bits(g), _bits(rg)), (shape, f64, B)


def test_noalloc_argument_errors(hip):
    tape = wl.tanh_mlp(4, 3, 2)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    x = torch.zeros(8, 4, device="cuda")
    y = torch.zeros(8, device="cuda")
    g = torch.zeros(8, 4, device="cuda")
    with pytest.raises(GpuError, match="batch_size must be non-zero"):
        hip.gradient_batch_dev_noalloc(bufs, x, 0, y, g)
    lib = load_library()
    assert lib.ehip_gradient_batch_dev_noalloc(hip.handle, bufs.handle, None, 8, None, None, None) == GpuError.OTHER
