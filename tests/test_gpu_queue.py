"""Multi-tape persistent queue: many tapes of uneven length and point count in
one launch.  Every job's results equal the per-tape launch bit for bit (same
per-lane arithmetic, only the schedule differs) and match the f64 oracle to
the composite-tape bar."""

import functools

import numpy as np
import pytest
import torch

import pyoracle as po
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuError, GpuTapeData
from test_gpu_parity import check_grad

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def suite(scale, count, f64, seed=3):
    specs = wl.random_dag_suite(seed, count, scale)
    tapes, xs = [], []
    for k, (length, n, B) in enumerate(specs):
        tapes.append(wl.random_dag(seed, k, length, n))
        xs.append(wl.points_uniform(seed * 100 + k, B, n, -1.0, 1.0).astype(np.float64 if f64 else np.float32))
    return tapes, xs


def upload(hip, tapes, f64):
    return [hip.upload_tape(GpuTapeData.from_tape_f64(t) if f64 else GpuTapeData.from_tape(t)) for t in tapes]


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("wg,lds,fetch", [(256, 0, 0), (64, 4096, 0), (128, 65536, 0), (256, 0, 1), (128, 16384, 1)])
def test_queue_matches_per_tape_launches(hip, f64, wg, lds, fetch):
    tapes, xs = suite(0.04, 24, f64)
    bufs = upload(hip, tapes, f64)
    tdt = torch.float64 if f64 else torch.float32
    jobs, outs = [], []
    for k, (b, x) in enumerate(zip(bufs, xs)):
        B, n = x.shape
        dx = torch.from_numpy(x).cuda()
        dy = torch.full((B,), float("nan"), dtype=tdt, device="cuda")
        dg = torch.full((B, n), float("nan"), dtype=tdt, device="cuda") if k % 3 else None
        jobs.append((b, 1 if dg is not None else 0, dx, B, dy, dg))
        outs.append((dx, dy, dg))
    hip.set_fetch(fetch)
    try:
        hip.multi_batch_dev(jobs, wg=wg, lds_budget=lds)
    finally:
        hip.set_fetch(0)
    for (b, what, dx, B, dy, dg), x, tape in zip(jobs, xs, tapes):
        ry = torch.empty_like(dy)
        if what == 1:
            rg = torch.empty_like(dg)
            hip.gradient_batch_dev(b, dx, B, ry, rg)
            torch.cuda.synchronize()
            assert np.array_equal(dg.cpu().numpy().view(np.uint8), rg.cpu().numpy().view(np.uint8))
        else:
            hip.forward_batch_dev(b, dx, B, ry)
            torch.cuda.synchronize()
        assert np.array_equal(dy.cpu().numpy().view(np.uint8), ry.cpu().numpy().view(np.uint8))


def test_queue_matches_oracle_f64(hip):
    tapes, xs = suite(0.03, 12, True, seed=5)
    bufs = upload(hip, tapes, True)
    jobs = []
    for b, x in zip(bufs, xs):
        B, n = x.shape
        jobs.append((b, 1, torch.from_numpy(x).cuda(), B, torch.empty(B, dtype=torch.float64, device="cuda"),
                     torch.empty(B, n, dtype=torch.float64, device="cuda")))
    hip.multi_batch_dev(jobs)
    for (b, _, dx, B, dy, dg), x, tape in zip(jobs, xs, tapes):
        check_grad(tape, x[:64], dy.cpu().numpy()[:64], dg.cpu().numpy()[:64])


def test_queue_single_point_and_empty(hip):
    tape = wl.tanh_mlp(5, 7, 40)
    b = hip.upload_tape(GpuTapeData.from_tape(tape))
    x = torch.from_numpy(wl.points_normal(1, 1, 5).astype(np.float32)).cuda()
    y = torch.empty(1, device="cuda")
    g = torch.empty(1, 5, device="cuda")
    hip.multi_batch_dev([(b, 1, x, 1, y, g)])
    y_ref, g_ref = po.gradient(tape, x.cpu().numpy(), dtype=np.float32)
    assert abs(float(y.cpu()[0]) - float(y_ref[0, 0])) <= 2e-5 * max(1.0, abs(float(y_ref[0, 0])))
    hip.multi_batch_dev([])  # nothing to do


def test_queue_errors(hip):
    t32 = hip.upload_tape(GpuTapeData.from_tape(wl.tanh_mlp(3, 4, 41)))
    t64 = hip.upload_tape(GpuTapeData.from_tape_f64(wl.tanh_mlp(3, 4, 41)))
    x32 = torch.zeros(4, 3, device="cuda")
    x64 = torch.zeros(4, 3, dtype=torch.float64, device="cuda")
    y = torch.zeros(4, device="cuda")
    with pytest.raises(GpuError, match="share a dtype"):
        hip.multi_batch_dev([(t32, 0, x32, 4, y, None), (t64, 0, x64, 4, y, None)])
    with pytest.raises(GpuError, match="without a gradient buffer"):
        hip.multi_batch_dev([(t32, 1, x32, 4, y, None)])
    with pytest.raises(GpuError, match="batch_size must be non-zero"):
        hip.multi_batch_dev([(t32, 0, x32, 0, y, None)])
    with pytest.raises(GpuError, match="workgroup size"):
        hip.multi_batch_dev([(t32, 0, x32, 4, y, None)], wg=32)
