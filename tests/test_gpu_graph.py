"""hipGraph capture of repeated evaluations: every replay equals the eager
launch bit for bit, with eager work and synchronisations between replays."""

import numpy as np
import pytest
import torch

from echidna_amd import philox
from echidna_amd import workloads as wl
from echidna_amd.backend import TUNE_GRADIENT, TUNE_HVP, TUNE_TAYLOR
from echidna_amd.tape import GpuError, GpuTapeData

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("f64", [False, True])
def test_graph_replays_match_eager(hip, f64):
    tape = wl.tanh_mlp2(14, 20, 12, 61)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    tdt = torch.float64 if f64 else torch.float32
    B, n = 5000, 14
    x = torch.from_numpy(wl.points_normal(1, B, n)).to(tdt).cuda()
    v = torch.from_numpy(philox.normal(2, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n)).to(tdt).cuda()
    y = torch.empty(B, dtype=tdt, device="cuda")
    g = torch.empty(B, n, dtype=tdt, device="cuda")
    h = torch.empty(B, n, dtype=tdt, device="cuda")
    c = torch.empty(B, 1, 3, dtype=tdt, device="cuda")
    hip.gradient_batch_dev(bufs, x, B, y, g)
    torch.cuda.synchronize()
    y0, g0 = y.clone(), g.clone()
    hip.hvp_batch_dev(bufs, x, n, v, B, g, h)
    torch.cuda.synchronize()
    hg0, h0 = g.clone(), h.clone()
    hip.taylor_kth_batch_dev(bufs, x, n, v, B, 3, c)
    torch.cuda.synchronize()
    c0 = c.clone()
    graphs = [
        (hip.graph_capture(bufs, TUNE_GRADIENT, B, x, d_y=y, d_g=g, repeats=3), lambda: (y, g), (y0, g0)),
        (hip.graph_capture(bufs, TUNE_HVP, B, x, d_dirs=v, d_g=g, d_h=h, repeats=2), lambda: (g, h), (hg0, h0)),
        (hip.graph_capture(bufs, TUNE_TAYLOR, B, x, d_y=c, d_dirs=v, order=3), lambda: (c,), (c0,)),
    ]
    for rep in range(3):
        for graph, now, ref in graphs:
            for t in now():
                t.fill_(float("nan"))
            torch.cuda.synchronize()
            graph.launch()
            torch.cuda.synchronize()
            for a, b in zip(now(), ref):
                assert np.array_equal(a.cpu().numpy().view(np.uint8), b.cpu().numpy().view(np.uint8)), rep
            # eager work between replays
            hip.gradient_batch_dev(bufs, x, B, torch.empty_like(y), torch.empty_like(g))
            torch.cuda.synchronize()
    for graph, _, _ in graphs:
        graph.close()


def test_graph_capture_errors(hip):
    bufs = hip.upload_tape(GpuTapeData.from_tape(wl.tanh_mlp(3, 4, 62)))
    x = torch.zeros(4, 3, device="cuda")
    with pytest.raises(GpuError, match="repeats must be non-zero"):
        hip.graph_capture(bufs, TUNE_GRADIENT, 4, x, d_y=torch.zeros(4, device="cuda"), d_g=x, repeats=0)
    with pytest.raises(GpuError, match="batch_size must be non-zero"):
        hip.graph_capture(bufs, TUNE_GRADIENT, 0, x, d_y=torch.zeros(4, device="cuda"), d_g=x)
