"""GPU parity: the HIP backend against the CPU oracle (f64 reference sweeps).

Bars (stated here, used below):
  * per-opcode table: the reference's own ULP budgets (tests/gpu_cpu_parity.rs),
    f64 GPU vs f64 CPU and f32 GPU vs f64 CPU rounded to f32;
  * composite tapes, f64: elementwise |gpu - cpu| <= 1e-12 * max(|cpu|, M),
    M the sum of the absolute terms behind the entry (tests/parity_bar.py);
  * composite tapes, f32 vs the f64 oracle: the same form with 2e-5 (F32_TOL).
"""

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuError, GpuTapeData, record
from parity_bar import assert_f64
from parity_cases import CASES, ulp_diff_f32, ulp_diff_f64

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 2e-5


def check_grad(tape, x, y, g, tol=F64_TOL, what=""):
    """Value and gradient of every point against the oracle at `tol`."""
    x = np.asarray(x, np.float64).reshape(-1, tape.num_inputs)
    y_ref, g_ref = po.gradient(tape, x)
    mag = po.magnitudes(tape, x)
    B = x.shape[0]
    assert_f64(np.asarray(y).reshape(B, -1)[:, 0], y_ref[:, 0], mag["value"][:, 0], what + " value", tol)
    assert_f64(np.asarray(g).reshape(B, -1), g_ref, mag["grad"], what + " gradient", tol)


# ---------------------------------------------------------------------------
# per-opcode parity table
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_parity_table_f64(hip, case):
    name, build, points, _f32_ulp, f64_ulp = case
    tape = build()
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    x = np.asarray(points, dtype=np.float64)
    B = x.shape[0]
    y, g = hip.gradient_batch(bufs, x, B)
    y_ref, g_ref = po.gradient(tape, x)
    yf = hip.forward_batch(bufs, x, B)
    for b in range(B):
        assert ulp_diff_f64(y[b], y_ref[b, 0]) <= f64_ulp, (name, x[b], y[b], y_ref[b, 0])
        assert ulp_diff_f64(yf[b], y_ref[b, 0]) <= f64_ulp, (name, x[b], yf[b], y_ref[b, 0])
        for i in range(tape.num_inputs):
            assert ulp_diff_f64(g[b * tape.num_inputs + i], g_ref[b, i]) <= f64_ulp, (name, x[b], i, g, g_ref)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_parity_table_f32(hip, case):
    name, build, points, f32_ulp, _ = case
    tape = build()
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64_lossy(tape))
    x = np.asarray(points, dtype=np.float64)
    B = x.shape[0]
    y, g = hip.gradient_batch(bufs, x.astype(np.float32), B)
    y_ref, g_ref = po.gradient(tape, x)
    with np.errstate(over="ignore"):
        for b in range(B):
            assert ulp_diff_f32(y[b], np.float32(y_ref[b, 0])) <= f32_ulp, (name, x[b], y[b], y_ref[b, 0])
            for i in range(tape.num_inputs):
                assert ulp_diff_f32(g[b * tape.num_inputs + i], np.float32(g_ref[b, i])) <= f32_ulp, (name, x[b], i)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def test_rosenbrock_n16_f64(hip):
    """Extended Rosenbrock n=16, B=1024, x ~ U[-2,2]^16 from seed 0 (f64)."""
    tape = wl.rosenbrock(16)
    x = wl.points_uniform(0, 1024, 16, -2.0, 2.0)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    y, g = hip.gradient_batch(bufs, x, 1024)
    check_grad(tape, x, y, g)


def test_rosenbrock_n16_f32(hip):
    tape = wl.rosenbrock(16)
    x = wl.points_uniform(0, 1024, 16, -2.0, 2.0)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    y, g = hip.gradient_batch(bufs, x.astype(np.float32), 1024)
    check_grad(tape, x.astype(np.float32), y, g, F32_TOL)


@pytest.fixture(scope="module")
def mlp32():
    return wl.tanh_mlp(32, 64, 1)


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("wg,lds", [(64, 0), (256, 65536), (64, 4096)])
def test_mlp32_f64(hip, mlp32, fuse, wg, lds):
    """tanh MLP n=32, H=64 (seed 1); LDS-only, wide workgroup and HBM-spill register files."""
    hip.set_fusion(fuse)
    hip.set_launch(wg, lds)
    try:
        B = 3000
        x = wl.points_normal(1, B, 32)
        bufs = hip.upload_tape(GpuTapeData.from_tape_f64(mlp32))
        y, g = hip.gradient_batch(bufs, x, B)
        yf = hip.forward_batch(bufs, x, B)
    finally:
        hip.set_fusion(True)
        hip.set_launch(0, 0)
    assert np.array_equal(y, yf)
    check_grad(mlp32, x, y, g)


def test_mlp32_f32(hip, mlp32):
    B = 4096
    x = wl.points_normal(1, B, 32).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(mlp32))
    y, g = hip.gradient_batch(bufs, x, B)
    check_grad(mlp32, x, y, g, F32_TOL)
    # same-precision CPU path
    _, g32 = po.gradient(mlp32, x, dtype=np.float32)
    mag = po.magnitudes(mlp32, x.astype(np.float64))
    assert_f64(g.reshape(B, 32), g32.astype(np.float64), mag["grad"], "f32 oracle", F32_TOL)


def test_mlp32_full_batch_f32_device(hip, mlp32):
    """B = 2^20 through the device-pointer entry point; oracle on a strided sample."""
    torch = pytest.importorskip("torch")
    B = 1 << 20
    x = wl.points_normal(1, B, 32).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(mlp32))
    dx = torch.from_numpy(x).cuda()
    dy = torch.empty(B, dtype=torch.float32, device="cuda")
    dg = torch.empty(B, 32, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    hip.gradient_batch_dev(bufs, dx, B, dy, dg)
    hip_sync = torch.cuda.synchronize
    hip_sync()
    idx = np.concatenate([np.arange(0, B, 997), [B - 3, B - 2, B - 1]])  # and the last partial workgroup
    check_grad(mlp32, x[idx], dy.cpu().numpy()[idx], dg.cpu().numpy()[idx], F32_TOL)


def test_random_dags(hip):
    for k in range(4):
        tape = wl.random_dag(3, k, 400 + 300 * k, 6 + 5 * k)
        x = wl.points_uniform(10 + k, 257, tape.num_inputs, -1.0, 1.0)
        bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
        y, g = hip.gradient_batch(bufs, x, 257)
        check_grad(tape, x, y, g, what=f"dag {k}")


def test_multi_output_forward(hip):
    from echidna_amd.tape import record_multi
    tape, _ = record_multi(lambda v: [v[0] * v[1], v[0].sin() + v[1], v[1]], [1.0, 2.0])
    x = np.array([[0.5, 2.0], [1.5, -1.0], [3.0, 0.25]])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    assert hip.num_outputs(bufs) == 3
    y = hip.forward_batch(bufs, x, 3).reshape(3, 3)
    assert_f64(y, po.forward(tape, x), po.magnitudes(tape, x)["value"], "outputs")
    # gradient seeds output 0
    y0, g = hip.gradient_batch(bufs, x, 3)
    check_grad(tape, x, y0.reshape(3, 3)[:, 0], g)


# ---------------------------------------------------------------------------
# validation and error paths
# ---------------------------------------------------------------------------

def _tiny():
    return GpuTapeData(opcodes=np.array([0, 1, 4], np.uint32), arg0=np.array([0xFFFFFFFF, 0xFFFFFFFF, 0], np.uint32),
                       arg1=np.array([0xFFFFFFFF, 0xFFFFFFFF, 1], np.uint32),
                       constants=np.array([0.0, 3.0, 0.0], np.float32), num_ops=3, num_inputs=1, num_variables=3,
                       output_index=2, output_indices=np.array([2], np.uint32))


def test_upload_rejects_invalid(hip):
    hip.upload_tape(_tiny())
    d = _tiny(); d.arg0[2] = 9999
    with pytest.raises(GpuError, match="invalid GpuTapeData"):
        hip.upload_tape(d)
    d = _tiny(); d.opcodes[2] = 24; d.arg1[2] = 9999
    with pytest.raises(GpuError, match="invalid GpuTapeData"):
        hip.upload_tape(d)
    d = _tiny(); d.output_index = 0xFFFFFFFF; d.output_indices = np.zeros(0, np.uint32)
    with pytest.raises(GpuError, match="invalid GpuTapeData"):
        hip.upload_tape(d)
    d = _tiny(); d.output_indices = np.array([3], np.uint32)
    with pytest.raises(GpuError, match="invalid GpuTapeData"):
        hip.upload_tape(d)
    d = _tiny(); d.opcodes[2] = 43
    with pytest.raises(GpuError) as e:
        hip.upload_tape(d)
    assert e.value.code == GpuError.CUSTOM_OPS_NOT_SUPPORTED


def test_batch_argument_errors(hip):
    bufs = hip.upload_tape(_tiny())
    with pytest.raises(GpuError, match="batch_size must be non-zero"):
        hip.forward_batch(bufs, np.zeros(0, np.float32), 0)
    # a zero-input tape: a single Const output
    z = GpuTapeData(opcodes=np.array([1], np.uint32), arg0=np.array([0xFFFFFFFF], np.uint32),
                    arg1=np.array([0xFFFFFFFF], np.uint32), constants=np.array([2.0], np.float32), num_ops=1,
                    num_inputs=0, num_variables=1, output_index=0, output_indices=np.array([0], np.uint32))
    zb = hip.upload_tape(z)
    with pytest.raises(GpuError, match="no inputs"):
        hip.gradient_batch(zb, np.zeros(0, np.float32), 1)


def test_gradient_of_input_and_const_outputs(hip):
    # output is an input slot: gradient is e_0; output Const: gradient 0
    tape, _ = record(lambda v: v[0], [1.0, 2.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    y, g = hip.gradient_batch(bufs, np.array([[3.0, 4.0]]), 1)
    assert y[0] == 3.0 and list(g) == [1.0, 0.0]
    tape, _ = record(lambda v: v[0] * 0.0 + v[1] * 0.0 + 5.0, [1.0, 2.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    y, g = hip.gradient_batch(bufs, np.array([[3.0, 4.0]]), 1)
    y_ref, g_ref = po.gradient(tape, np.array([[3.0, 4.0]]))
    assert y[0] == y_ref[0, 0] and list(g) == list(g_ref[0])


def test_f32_tanh_pair_ulp_sweep(hip):
    """The f32 scalar tanh and its partial (from exp(-2|a|) and expm1(-2|a|))
    over 2^20 arguments in [-50, 50] and special values, against the f64
    oracle rounded to f32: within 4 ulp for both, inside the reference's
    16-ulp budget for tanh; both paths (interpreter and specialised kernel)."""
    from echidna_amd.backend import TUNE_GRADIENT, HipContext

    tape, _ = record(lambda v: v[0].tanh(), [0.5])
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    x = np.random.default_rng(80).uniform(-50, 50, 1 << 20)
    x[:12] = [0.0, -0.0, 1e-30, -1e-30, 1e-7, 0.5, 20.0, 44.0, 44.5, 60.0, np.inf, -np.inf]
    x = x.astype(np.float32)
    y_ref, g_ref = po.gradient(tape, x.reshape(-1, 1).astype(np.float64))
    yr, gr = y_ref[:, 0].astype(np.float32), g_ref[:, 0].astype(np.float32)
    worst = {}
    for spec in (False, True):
        if spec:
            HipContext.specialize(bufs, TUNE_GRADIENT)
        y, g = hip.gradient_batch(bufs, x, len(x))
        for name, got, ref in (("value", y, yr), ("derivative", g, gr)):
            gi = np.asarray(got, np.float32).view(np.int32).astype(np.int64)
            ri = ref.view(np.int32).astype(np.int64)
            gi = np.where(gi < 0, -(gi & 0x7FFFFFFF), gi)
            ri = np.where(ri < 0, -(ri & 0x7FFFFFFF), ri)
            worst[(spec, name)] = int(np.max(np.abs(gi - ri)))
    print("max ulp", worst)
    assert all(w <= 4 for w in worst.values()), worst
    # signed zeros and infinities exactly
    y, g = hip.gradient_batch(bufs, x[:12], 12)
    assert np.signbit(y[1]) and not np.signbit(y[0]) and y[10] == 1.0 and y[11] == -1.0
    assert g[0] == 1.0 and g[10] == 0.0 and g[11] == 0.0
