"""Tape wire format (no GPU): the library's serde JSON reader and writer
against the reference's own serde tests (tests/serde.rs), case by case:
round trips evaluated by the oracle, the Custom-op refusal, the malformed
payload of regression #26, and each corrupt-payload rejection.  The layout
is pinned by a hand-derived serde_json text of the rosenbrock tape, and the
number text by shortest round-trip digits: Python's repr (f64) and numpy's
dragon4 (f32) must give the same digits, and every value reads back
bit-identically.  The CBOR form (the reference's ciborium round trip) is
pinned by RFC 8949 Appendix A's encodings and by the same rejections,
re-encoded with a small independent encoder."""

import json
import re

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import serde as sd
from echidna_amd import workloads as wl
from echidna_amd.tape import UNUSED, BytecodeTape, GpuError, OpCode, record, record_multi


def rosenbrock(x):
    dx = x[0] - 1.0
    t = x[1] - x[0] * x[0]
    return dx * dx + 100.0 * t * t


def rosen_tape(x=(1.5, 2.5)):
    return record(rosenbrock, list(x))[0]


def tape_value():
    return json.loads(sd.tape_to_json(rosen_tape()))


def tape_value_with_unary():
    tape, _ = record(lambda v: (v[0] * v[0]).sin() + v[1], [0.5, 1.0])
    v = json.loads(sd.tape_to_json(tape))
    return v, v["opcodes"].index("Sin")


def last_real_op(v):
    ops = v["opcodes"]
    return max(i for i, op in enumerate(ops) if op not in ("Input", "Const"))


def assert_rejected(v, match):
    with pytest.raises(GpuError, match=match) as e:
        sd.tape_from_json(json.dumps(v))
    assert e.value.code == GpuError.OTHER


# ---- round trips (tests/serde.rs:13-66) -------------------------------------

# serde_json::to_string of the rosenbrock tape recorded at (1.5, 2.5): the
# recording order of dx*dx + 100*t*t, UNUSED = u32::MAX in empty operand
# slots, shortest decimals with a ".0" on integral values, the empty
# output_indices and custom_second_args of a single-output record().
ROSEN_JSON = (
    '{"opcodes":["Input","Input","Const","Sub","Mul","Sub","Mul","Const","Mul","Mul","Add"],'
    '"arg_indices":[[4294967295,4294967295],[4294967295,4294967295],[4294967295,4294967295],[0,2],[0,0],'
    '[1,4],[3,3],[4294967295,4294967295],[7,5],[8,5],[6,9]],'
    '"values":[1.5,2.5,1.0,0.5,2.25,0.25,0.25,100.0,25.0,6.25,6.5],'
    '"num_inputs":2,"num_variables":11,"output_index":10,"output_indices":[],"custom_second_args":{}}')


def test_rosenbrock_layout():
    assert sd.tape_to_json(rosen_tape()) == ROSEN_JSON


def same_tape(a, b):
    return (a.opcodes == b.opcodes and a.arg0 == b.arg0 and a.arg1 == b.arg1 and a.num_inputs == b.num_inputs
            and a.output_index == b.output_index and a.output_indices == b.output_indices
            and np.array_equal(np.asarray(a.values).view(np.uint64), np.asarray(b.values).view(np.uint64)))


@pytest.mark.parametrize("x", [[1.5, 2.5], [2.0, 3.0]])
def test_roundtrip_gradient(x):
    tape = rosen_tape((1.0, 2.0) if x == [2.0, 3.0] else (1.5, 2.5))
    tape2 = sd.tape_from_json(sd.tape_to_json(tape))
    assert same_tape(tape, tape2)
    y1, g1 = po.gradient(tape, np.array([x]))
    y2, g2 = po.gradient(tape2, np.array([x]))
    assert np.array_equal(g1, g2) and np.array_equal(y1, y2)
    # the reference's closed form: grad = (-2(1-x) - 400x(y-x^2), 200(y-x^2))
    a, b = x
    assert np.allclose(g2[0], [-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)], rtol=1e-12)


def test_roundtrip_supports_hessian():
    tape2 = sd.tape_from_json(sd.tape_to_json(rosen_tape()))
    x = np.array([1.5, 2.5])
    H = po.hessian(tape2, x)
    assert np.allclose(H, [[1200 * x[0] ** 2 - 400 * x[1] + 2, -400 * x[0]], [-400 * x[0], 200]], rtol=1e-12)


@pytest.mark.parametrize("name", ["mlp", "dag", "multi"])
def test_roundtrip_larger_tapes(name):
    if name == "mlp":
        tape = wl.tanh_mlp(6, 5, 9)
    elif name == "dag":
        tape = wl.random_dag(3, 1, 500, 20)
    else:
        tape, _ = record_multi(lambda v: [v[0] * v[1], v[1].sin(), v[0].powi(-3), v[1].atan2(v[0])], [0.3, 0.7])
    s = sd.tape_to_json(tape)
    tape2 = sd.tape_from_json(s)
    assert same_tape(tape, tape2)
    assert sd.tape_to_json(tape2) == s  # idempotent text
    v = json.loads(s)
    assert v["num_variables"] == len(v["opcodes"]) == len(v["arg_indices"]) == len(v["values"])
    X = np.random.default_rng(0).uniform(-1, 1, (4, tape.num_inputs))
    assert np.array_equal(po.forward(tape, X), po.forward(tape2, X))


def test_f32_payload_reads_as_rounded_f64():
    tape = wl.tanh_mlp(4, 3, 2)
    t32 = sd.tape_from_json(sd.tape_to_json(tape), f32=True)
    assert np.array_equal(np.asarray(t32.values, np.float32), np.asarray(tape.values).astype(np.float32))
    # an f32 tape's text reads back to the same f32 values
    t32b = sd.tape_from_json(sd.tape_to_json(tape, f32=True), f32=True)
    assert np.array_equal(np.asarray(t32b.values, np.float32), np.asarray(t32.values, np.float32))


def test_custom_op_tape_serialization_fails():
    t = BytecodeTape()
    t.new_input(1.0)
    t._push_entry(OpCode.Custom, 0, 0, 2.0)
    t.set_output(1)
    with pytest.raises(GpuError, match="cannot serialize a BytecodeTape containing custom ops") as e:
        sd.tape_to_json(t)
    assert e.value.code == GpuError.CUSTOM_OPS_NOT_SUPPORTED


def test_custom_opcode_in_payload_rejected():
    v = tape_value()
    v["opcodes"][last_real_op(v)] = "Custom"
    assert_rejected(v, r"opcodes\[10\] is Custom, which cannot be deserialized")


# ---- malformed and corrupt payloads (tests/serde.rs:174-330) ----------------

def test_regression_26_malformed_tape():
    text = ('{"opcodes":[0,1],"arg_indices":[[0,0]],"values":[0.0,1.0],"num_inputs":1,"num_variables":2,'
            '"output_index":1,"output_indices":[1],"custom_ops":{},"custom_second_args":{}}')
    with pytest.raises(GpuError):
        sd.tape_from_json(text)
    # the same payload with variant names still fails: 2 opcodes, 1 arg pair
    with pytest.raises(GpuError, match=r"arg_indices.len\(\) \(1\) != opcodes.len\(\) \(2\)"):
        sd.tape_from_json(text.replace("[0,1]", '["Input","Const"]'))


def test_corrupt_payload_baseline_roundtrips():
    assert same_tape(sd.tape_from_json(json.dumps(tape_value())), rosen_tape())


def test_serialized_payload_carries_num_variables():
    v = tape_value()
    assert isinstance(v["num_variables"], int) and v["num_variables"] == len(v["opcodes"])


def test_rejects_num_variables_mismatch():
    v = tape_value()
    v["num_variables"] += 1
    assert_rejected(v, r"num_variables \(12\) != opcodes.len\(\) \(11\)")


def test_rejects_args_on_input_entry():
    v = tape_value()
    v["arg_indices"][0] = [0, 0]
    assert_rejected(v, r"arg_indices\[0\] must be \[UNUSED, UNUSED\] for an Input entry, found \[0, 0\]")


def test_rejects_input_opcode_after_prefix():
    v = tape_value()
    v["opcodes"][last_real_op(v)] = "Input"
    assert_rejected(v, r"opcodes\[10\] is Input, but only the first 2 entries may be inputs")


def test_rejects_unset_output_index_sentinel():
    v = tape_value()
    v["output_index"] = 0xFFFFFFFF
    assert_rejected(v, r"output_index \(4294967295\) >= tape length \(11\)")


def test_rejects_forward_referencing_arg():
    v = tape_value()
    last = last_real_op(v)
    v["arg_indices"][last][0] = last
    assert_rejected(v, r"arg_indices\[10\]\[0\] \(10\) must reference an earlier slot")


def test_rejects_unary_with_arg1_index():
    v, sin_pos = tape_value_with_unary()
    v["arg_indices"][sin_pos][1] = 0
    assert_rejected(v, rf"arg_indices\[{sin_pos}\]\[1\] \(0\) should be UNUSED for unary op Sin")


def test_rejects_stale_custom_second_args():
    v = tape_value()
    v["custom_second_args"]["3"] = 1
    assert_rejected(v, "custom_second_args must be empty: custom ops cannot be deserialized")


def test_rejects_mismatched_lengths_without_panic():
    v = tape_value()
    v["arg_indices"].append([0, 0])
    assert_rejected(v, r"arg_indices.len\(\) \(12\) != opcodes.len\(\) \(11\)")
    v = tape_value()
    v["values"].pop()
    assert_rejected(v, r"values.len\(\) \(10\) != opcodes.len\(\) \(11\)")


# ---- schema-level behaviour of the serde derive -----------------------------

def test_defaults_unknown_fields_and_key_order():
    v = tape_value()
    del v["output_indices"], v["custom_second_args"]
    v["extra"] = {"nested": [1, 2, {"a": None}], "s": "xé\"\\"}
    text = json.dumps(dict(reversed(list(v.items()))), indent=2)  # any order, whitespace
    assert same_tape(sd.tape_from_json(text), rosen_tape())


@pytest.mark.parametrize("mutate,match", [
    (lambda v: v.pop("values"), "missing field `values`"),
    (lambda v: v.pop("num_inputs"), "missing field `num_inputs`"),
    (lambda v: v["opcodes"].__setitem__(3, "Subtract"), "unknown variant `Subtract`"),
    (lambda v: v["opcodes"].__setitem__(3, 3), "expected an OpCode variant name"),
    (lambda v: v.__setitem__("num_inputs", -1), "invalid value: integer `-1`, expected u32"),
    (lambda v: v.__setitem__("num_inputs", 2.0), "invalid type: floating point `2.0`, expected u32"),
    (lambda v: v.__setitem__("output_index", 2 ** 32), "invalid value: integer `4294967296`, expected u32"),
    (lambda v: v["arg_indices"].__setitem__(3, [0]), "invalid length 1, expected an array of length 2"),
    (lambda v: v["arg_indices"].__setitem__(3, [0, 2, 1]), "trailing characters"),
    (lambda v: v["values"].__setitem__(2, None), "invalid type: null, expected f64"),
    (lambda v: v["values"].__setitem__(2, "1.0"), "invalid type: string, expected f64"),
    (lambda v: v.__setitem__("custom_second_args", None), "expected a map"),
])
def test_schema_rejections(mutate, match):
    v = tape_value()
    mutate(v)
    assert_rejected(v, match)


@pytest.mark.parametrize("text,match", [
    ("", "expected struct TapeData"),
    ("[]", "expected struct TapeData"),
    (ROSEN_JSON[:-1], "EOF"),
    (ROSEN_JSON + "x", "trailing characters"),
    (ROSEN_JSON.replace('"num_inputs":2', '"num_inputs":2,"num_inputs":2'), "duplicate field `num_inputs`"),
    (ROSEN_JSON.replace("1.5,", "01.5,"), "invalid number"),
    (ROSEN_JSON.replace("1.5,", "1.,"), "invalid number"),
    (ROSEN_JSON.replace("1.5,", "1e999,"), "number out of range"),
    ('{"a":' * 200 + "1" + "}" * 200, "recursion limit exceeded"),
])
def test_text_rejections(text, match):
    with pytest.raises(GpuError, match=match):
        sd.tape_from_json(text)


def test_enum_map_form_and_escapes_accepted():
    text = ROSEN_JSON.replace('"Add"', '{"Add":null}').replace('"opcodes"', '"op\\u0063odes"')
    assert same_tape(sd.tape_from_json(text), rosen_tape())


def test_integer_and_exponent_values_parse():
    text = ROSEN_JSON.replace("[1.5,2.5,1.0,", "[15e-1,2.5E0,1,")
    assert same_tape(sd.tape_from_json(text), rosen_tape())


def test_non_finite_values_write_null_and_do_not_read_back():
    t = BytecodeTape()
    t.new_input(1.0)
    t.push_const(float("inf"))
    t._push_entry(OpCode.Mul, 0, 1, float("nan"))
    t.set_output(2)
    s = sd.tape_to_json(t)
    assert '"values":[1.0,null,null]' in s
    with pytest.raises(GpuError, match="invalid type: null, expected f64"):
        sd.tape_from_json(s)


# ---- number text ------------------------------------------------------------

def _digits_exp(txt):
    m = re.fullmatch(r"-?(\d+)(?:\.(\d+))?(?:e(-?\d+))?", txt)
    assert m, txt
    ip, fp, ex = m.group(1), m.group(2) or "", int(m.group(3) or 0)
    allds = ip + fp
    lead = len(allds) - len(allds.lstrip("0"))
    return allds.lstrip("0").rstrip("0"), ex + len(ip) - lead


def _const_tape(vals):
    t = BytecodeTape()
    for v in vals:
        t.push_const(float(v))
    return t


def _value_texts(s):
    return re.search(r'"values":\[(.*?)\]', s).group(1).split(",")


@pytest.mark.parametrize("v,text", [
    (1.0, "1.0"), (100.0, "100.0"), (0.1, "0.1"), (-0.0, "-0.0"), (0.0, "0.0"), (1e15, "1000000000000000.0"),
    (1e16, "1e16"), (1e20, "1e20"), (1e-5, "0.00001"), (1e-6, "1e-6"), (1.5e-7, "1.5e-7"), (5e-324, "5e-324"),
    (1.7976931348623157e308, "1.7976931348623157e308"), (-2.5, "-2.5"), (123456.789, "123456.789"),
])
def test_f64_text_layout(v, text):
    assert _value_texts(sd.tape_to_json(_const_tape([v]))) == [text]


def test_shortest_digits_and_exact_read_back():
    rng = np.random.default_rng(11)
    x = rng.integers(0, 2 ** 63, 20000, dtype=np.uint64).view(np.float64)
    x = np.concatenate([x[np.isfinite(x)], rng.standard_normal(2000), 2.0 ** np.arange(-1074, 1024, 1.0),
                        np.nextafter(2.0 ** np.arange(-1000, 1000, 7.0), 0)])
    s = sd.tape_to_json(_const_tape(x))
    for t, v in zip(_value_texts(s), x):
        if v != 0:
            assert _digits_exp(t) == _digits_exp(repr(float(abs(v))).replace("e+", "e")), (t, repr(v))
    back = np.asarray(sd.tape_from_json(s).values)
    assert np.array_equal(back.view(np.uint64), x.view(np.uint64))


def test_f32_shortest_digits_and_exact_read_back():
    rng = np.random.default_rng(12)
    xf = rng.integers(0, 2 ** 31, 20000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    xf = np.concatenate([xf[np.isfinite(xf)], (2.0 ** np.arange(-149, 128, 1.0)).astype(np.float32)])
    s = sd.tape_to_json(_const_tape(xf.astype(np.float64)), f32=True)
    for t, v in zip(_value_texts(s), xf):
        if v != 0:
            ref = np.format_float_scientific(v, unique=True).replace("e+", "e").replace(".e", "e")
            assert _digits_exp(t) == _digits_exp(ref), (t, ref)
    back = np.asarray(sd.tape_from_json(s, f32=True).values, np.float32)
    assert np.array_equal(back.view(np.uint32), xf.view(np.uint32))


# ---- derived serde forms ----------------------------------------------------

def test_pattern_roundtrip():
    from echidna_amd import sparse as sp
    from echidna_amd.tape import GpuTapeData

    tape = wl.random_dag(3, 1, 500, 20)
    d = GpuTapeData.from_tape_f64(tape)
    hp = sp.detect_sparsity(d)
    hp2 = sd.pattern_from_json(sd.pattern_to_json(hp))
    assert hp2.dim == hp.dim and np.array_equal(hp2.rows, hp.rows) and np.array_equal(hp2.cols, hp.cols)
    jp = sp.detect_jacobian_sparsity(d)
    jp2 = sd.pattern_from_json(sd.pattern_to_json(jp))
    assert (jp2.num_outputs, jp2.num_inputs) == (jp.num_outputs, jp.num_inputs)
    assert np.array_equal(jp2.rows, jp.rows) and np.array_equal(jp2.cols, jp.cols)
    assert json.loads(sd.pattern_to_json(hp)).keys() == {"dim", "rows", "cols"}


def test_nonsmooth_info_roundtrip():
    """tests/serde.rs:333-350 with the oracle's kink records of |x| + max(x, y) at (0, 0)."""
    tape, _ = record(lambda v: v[0].abs() + v[0].max(v[1]), [0.0, 0.0])
    _, sw, br = po.nonsmooth(tape, np.zeros((1, 2)))
    kidx = [i for i, op in enumerate(tape.opcodes) if op in (OpCode.Abs, OpCode.Max)]
    info = sd.NonsmoothInfo([sd.KinkEntry(i, tape.opcodes[i], float(sw[0, k]), int(br[0, k]))
                             for k, i in enumerate(kidx)])
    assert info.kinks
    info2 = sd.NonsmoothInfo.from_json(info.to_json())
    assert info2 == info and info2.signature() == info.signature()
    assert json.loads(info.to_json())["kinks"][0]["opcode"] == "Abs"


# ---- CBOR (ciborium's binary form; tests/serde.rs:156-170) -------------------

def cbor_encode(v) -> bytes:
    """Small independent RFC 8949 encoder for crafting payloads (floats always
    binary64, which every reader must accept)."""
    import struct

    def head(major, n):
        if n < 24:
            return bytes([major << 5 | n])
        for ai, w in ((24, 1), (25, 2), (26, 4), (27, 8)):
            if n < 1 << (8 * w):
                return bytes([major << 5 | ai]) + n.to_bytes(w, "big")
        raise ValueError(n)

    if v is None:
        return b"\xf6"
    if isinstance(v, bool):
        return b"\xf5" if v else b"\xf4"
    if isinstance(v, int):
        return head(0, v) if v >= 0 else head(1, -1 - v)
    if isinstance(v, float):
        return b"\xfb" + struct.pack(">d", v)
    if isinstance(v, str):
        return head(3, len(v.encode())) + v.encode()
    if isinstance(v, list):
        return head(4, len(v)) + b"".join(cbor_encode(x) for x in v)
    if isinstance(v, dict):
        return head(5, len(v)) + b"".join(cbor_encode(int(k) if isinstance(k, str) and k.isdigit() else k)
                                          + cbor_encode(x) for k, x in v.items())
    raise TypeError(type(v))


def _values_bytes(blob: bytes) -> bytes:
    """The encoded items of the `values` array (after its key and head)."""
    key = b"\x66values"
    i = blob.index(key) + len(key)
    j = blob.index(b"\x6anum_inputs")
    head = blob[i]
    skip = 1 if head & 31 < 24 else 1 + (1 << ((head & 31) - 24))
    return blob[i + skip:j]


# RFC 8949 Appendix A: preferred (shortest exact) float encodings
RFC_FLOATS = [(0.0, "f90000"), (-0.0, "f98000"), (1.0, "f93c00"), (1.1, "fb3ff199999999999a"), (1.5, "f93e00"),
              (65504.0, "f97bff"), (100000.0, "fa47c35000"), (3.4028234663852886e38, "fa7f7fffff"),
              (1.0e300, "fb7e37e43c8800759c"), (5.960464477539063e-8, "f90001"), (0.00006103515625, "f90400"),
              (-4.0, "f9c400"), (-4.1, "fbc010666666666666"), (float("inf"), "f97c00"), (float("nan"), "f97e00"),
              (float("-inf"), "f9fc00")]


@pytest.mark.parametrize("v,hexs", RFC_FLOATS)
def test_cbor_float_encoding_matches_rfc8949(v, hexs):
    assert _values_bytes(sd.tape_to_cbor(_const_tape([v]))).hex() == hexs


def test_cbor_integer_and_text_encoding_matches_rfc8949():
    t = _const_tape([1.0])
    t.output_indices = [0, 1, 10, 23, 24, 25, 100, 1000, 1000000, 4294967295]
    blob = sd.tape_to_cbor(t)
    key = b"\x6eoutput_indices"
    i = blob.index(key) + len(key)
    body = blob[i:blob.index(b"\x72custom_second_args")]
    assert body.hex() == "8a" + "00" + "01" + "0a" + "17" + "1818" + "1819" + "1864" + "1903e8" + "1a000f4240" + "1affffffff"
    assert blob[:1] == b"\xa8" and b"\x65Const" in blob  # map(8); text(5) "Const"
    assert blob.endswith(b"\x72custom_second_args\xa0")


def test_cbor_roundtrip_gradient():
    tape = rosen_tape()
    tape2 = sd.tape_from_cbor(sd.tape_to_cbor(tape))
    assert same_tape(tape, tape2)
    x = np.array([[1.5, 2.5]])
    assert np.array_equal(po.gradient(tape, x)[1], po.gradient(tape2, x)[1])


@pytest.mark.parametrize("name", ["mlp", "dag", "multi"])
def test_cbor_and_json_agree(name):
    if name == "mlp":
        tape = wl.tanh_mlp(6, 5, 9)
    elif name == "dag":
        tape = wl.random_dag(3, 1, 500, 20)
    else:
        tape, _ = record_multi(lambda v: [v[0] * v[1], v[1].sin(), v[0].powi(-3), v[1].atan2(v[0])], [0.3, 0.7])
    blob = sd.tape_to_cbor(tape)
    a, b = sd.tape_from_cbor(blob), sd.tape_from_json(sd.tape_to_json(tape))
    assert same_tape(a, tape) and same_tape(b, tape)
    assert sd.tape_to_cbor(a) == blob
    assert len(blob) < len(sd.tape_to_json(tape))
    t32 = sd.tape_from_cbor(sd.tape_to_cbor(tape, f32=True), f32=True)
    assert np.array_equal(np.asarray(t32.values, np.float32), np.asarray(tape.values).astype(np.float32))


def test_cbor_carries_non_finite_values():
    t = BytecodeTape()
    t.new_input(1.0)
    t.push_const(float("inf"))
    t._push_entry(OpCode.Mul, 0, 1, float("nan"))
    t.set_output(2)
    t2 = sd.tape_from_cbor(sd.tape_to_cbor(t))
    assert t2.values[0] == 1.0 and t2.values[1] == float("inf") and np.isnan(t2.values[2])


def test_cbor_reads_any_float_width_and_integer_values():
    v = tape_value()
    v["values"] = [float(x) for x in v["values"]]  # binary64 everywhere
    v["values"][2] = 1  # an integer where a float is expected
    assert same_tape(sd.tape_from_cbor(cbor_encode(v)), rosen_tape())


@pytest.mark.parametrize("mutate,match", [
    (lambda v: v.__setitem__("num_variables", v["num_variables"] + 1), r"num_variables \(12\) != opcodes.len\(\) \(11\)"),
    (lambda v: v["arg_indices"].__setitem__(0, [0, 0]), r"must be \[UNUSED, UNUSED\] for an Input entry"),
    (lambda v: v["opcodes"].__setitem__(10, "Input"), r"opcodes\[10\] is Input, but only the first 2"),
    (lambda v: v.__setitem__("output_index", 0xFFFFFFFF), r"output_index \(4294967295\) >= tape length"),
    (lambda v: v["arg_indices"][10].__setitem__(0, 10), r"arg_indices\[10\]\[0\] \(10\) must reference an earlier"),
    (lambda v: v["custom_second_args"].__setitem__("3", 1), "custom_second_args must be empty"),
    (lambda v: v["arg_indices"].append([0, 0]), r"arg_indices.len\(\) \(12\) != opcodes.len\(\) \(11\)"),
    (lambda v: v["opcodes"].__setitem__(10, "Custom"), r"opcodes\[10\] is Custom"),
    (lambda v: v.pop("values"), "missing field `values`"),
    (lambda v: v.__setitem__("num_inputs", -1), "negative integer, expected u32"),
    (lambda v: v["opcodes"].__setitem__(3, "Subtract"), "unknown variant `Subtract`"),
    (lambda v: v["values"].__setitem__(2, "1.0"), "invalid type: string, expected f64"),
])
def test_cbor_rejections(mutate, match):
    v = tape_value()
    mutate(v)
    with pytest.raises(GpuError, match=match):
        sd.tape_from_cbor(cbor_encode(v))


def test_cbor_truncated_and_trailing_payloads():
    blob = sd.tape_to_cbor(rosen_tape())
    for cut in (1, 10, len(blob) // 2, len(blob) - 1):
        with pytest.raises(GpuError, match="unexpected end of input"):
            sd.tape_from_cbor(blob[:cut])
    with pytest.raises(GpuError, match="trailing data"):
        sd.tape_from_cbor(blob + b"\x00")
    with pytest.raises(GpuError, match="expected struct TapeData"):
        sd.tape_from_cbor(b"\x80")
