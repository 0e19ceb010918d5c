"""Tapes loaded from the serde JSON wire format on the GPU: a payload
uploaded in one call (ehip_tape_upload_json) evaluates bit-identically to
the same tape uploaded from its arrays, and kink records of a GPU batch
survive the NonsmoothInfo JSON round trip (tests/serde.rs:333-350)."""

import json

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import serde as sd
from echidna_amd import sparse as sp
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuError, GpuTapeData, record

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("f64", [True, False])
@pytest.mark.parametrize("name", ["mlp", "dag"])
def test_json_upload_matches_array_upload(hip, name, f64):
    tape = wl.tanh_mlp(8, 6, 3) if name == "mlp" else wl.random_dag(5, 2, 800, 40)
    text = sd.tape_to_json(tape)
    a = hip.upload_tape_json(text, f64=f64)
    b = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    B = 1000
    X = np.random.default_rng(2).uniform(-1, 1, (B, tape.num_inputs))
    if not f64:
        X = X.astype(np.float32)
    ya, ga = hip.gradient_batch(a, X.reshape(-1), B)
    yb, gb = hip.gradient_batch(b, X.reshape(-1), B)
    assert np.array_equal(ya.view(np.uint8), yb.view(np.uint8))
    assert np.array_equal(ga.view(np.uint8), gb.view(np.uint8))
    if f64:
        y_ref, g_ref = po.gradient(tape, X)
        assert np.max(np.abs(ga.reshape(B, -1) - g_ref)) <= 1e-12 * max(1.0, np.max(np.abs(g_ref)))


def test_json_upload_rejects_corrupt_payload(hip):
    v = json.loads(sd.tape_to_json(wl.rosenbrock(3)))
    v["num_variables"] += 1
    with pytest.raises(GpuError, match=r"num_variables \(\d+\) != opcodes.len\(\)"):
        hip.upload_tape_json(json.dumps(v))


def test_kink_records_roundtrip(hip):
    tape, _ = record(lambda v: v[0].abs() + v[0].max(v[1]) + (v[1] * 2.0).floor(), [0.0, 0.0])
    bufs = hip.upload_tape_json(sd.tape_to_json(tape))
    X = np.array([[0.0, 0.0], [-0.5, 0.25], [1.5, -2.0]])
    res = sp.forward_nonsmooth(hip, bufs, X.reshape(-1), 3)
    for b in range(3):
        info = sd.point_info(res, b)
        assert len(info.kinks) == 3
        info2 = sd.NonsmoothInfo.from_json(info.to_json())
        assert info2 == info and info2.signature() == info.signature()
