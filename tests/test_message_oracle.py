"""The server messages: the oracle's jansson writer against hand-derived texts, the
messages' structure against the reference's key order (json.loads), and the
library's C-ABI (host code, no GPU) against the oracle byte for byte."""
import json

import numpy as np
import pytest

from oracle import message_oracle as M

# ------------------------------------------------------------------ jansson's writer


@pytest.mark.parametrize(
    "value, text",
    [
        (0.5, "0.5"),
        (2.0, "2.0"),  # no '.' or 'e' in %.17g: ".0" appended
        (-4.5, "-4.5"),
        (0.0, "0.0"),
        (-0.0, "-0.0"),
        (0.1, "0.10000000000000001"),  # 17 significant digits, not the shortest repr
        (float(np.float32(0.1)), "0.10000000149011612"),  # a float member promoted to double
        (1e22, "1e22"),  # '+' dropped from the exponent
        (1e-5, "1.0000000000000001e-5"),  # leading zero dropped from the exponent
        (1.5e300, "1.5000000000000001e300"),
        (2.0 ** -20, "9.5367431640625e-7"),  # %g drops trailing zeros
        (2.0 ** 70, "1.1805916207174113e21"),
        (1e-100, "1e-100"),
        (123456789012345678.0, "1.2345678901234568e17"),
        (100.0, "100.0"),
        (45.0, "45.0"),
    ],
)
def test_real_format(value, text):
    assert M.real(value) == text
    assert float(text) == value


def test_refused_reals_are_left_out():
    assert M.real(float("nan")) is None and M.real(float("inf")) is None
    assert M.dump([1.0, float("nan"), 2]) == "[1.0, 2]"
    assert M.dump(M.Obj([("a", float("inf")), ("b", 1)])) == '{"b": 1}'
    o = M.Obj([("a", 1.0)])
    o.set("a", float("nan"))  # json_object_set with NULL fails: the old value stays
    assert M.dump(o) == '{"a": 1.0}'


def test_strings_and_containers():
    assert M.string('a"b\\c/d') == '"a\\"b\\\\c/d"'  # '/' unescaped
    assert M.string("\b\f\n\r\t") == '"\\b\\f\\n\\r\\t"'
    assert M.string("\x01\x1f") == '"\\u0001\\u001F"'  # uppercase hex
    assert M.string("é中") == '"é中"'  # non-ASCII as is
    assert M.dump(M.Obj()) == "{}" and M.dump([]) == "[]"
    assert M.dump(M.Obj([("a", 1), ("b", [1, 2.5, True, "x"]), ("c", M.Obj([("d", False)]))])) == \
        '{"a": 1, "b": [1, 2.5, true, "x"], "c": {"d": false}}'
    o = M.Obj([("a", 1), ("b", 2)])
    o.set("a", 3)  # replaced in place
    o.set("c", 4)  # appended
    assert M.dump(o) == '{"a": 3, "b": 2, "c": 4}'


# ------------------------------------------------------------------ the messages

# key order of the register message: recreateJSON's "metadata", then
# lib/isaac.hpp:836-891 in code order
REGISTER_KEYS = ["metadata", "type", "name", "nodes", "framebuffer width", "framebuffer height", "max functors",
                 "functors", "projection count", "projection", "position", "rotation", "distance", "sources",
                 "interpolation", "step", "dimension", "width", "height", "depth", "protocol", "ao isEnabled",
                 "ao weight"]
# every section of the periodic message in code order (lib/isaac.hpp:2551-2836)
PERIOD_KEYS = ["metadata", "type", "meta nr", "projection", "position", "rotation", "distance", "transfer array",
               "transfer points", "functions", "weight", "interpolation", "render optimization", "step", "seed points",
               "advection step", "advection weight", "advection seeding period", "advection seeding duration",
               "advection on pause", "advection border", "iso threshold", "minmax", "background color", "clipping",
               "ao isEnabled", "ao weight", "render mode", "dither mode", "init"]


def _state(T=8):
    rng = np.random.default_rng(3)
    names = ["density", "vector \"field\"", "particles"]
    lut = rng.uniform(0.0, 1.0, size=(3, T, 4)).astype(np.float32)
    return M.State(
        source_names=names, feature_dims=[1, 3, 1], n_volume_field=2,
        functions=["idem", "length | mul(2)", "pow(3"], function_errors=[0, 0, -1],
        transfer_size=T, transfer_lut=lut,
        transfer_points=[[(0, (0.1, 0.2, 0.3, 0.0)), (T, (1.0, 0.5, 0.25, 1.0))], [(0, (0.0, 1.0, 0.0, 0.0))], []],
        weights=[2.0, 0.7, 1.0], iso_thresholds=[0.5, 0.0], min_max=[(0.0, 1.0), (-0.25, 3.5), (0.1, 0.9)],
        step=0.3, interpolation=True, background=(0.1, 0.2, 0.3),
        clipping=[((0.0, 0.0, 0.0), (1.0, 0.0, 0.0)), ((0.5, -0.5, 0.25), (0.0, 0.6, 0.8))],
        ao_enabled=True, ao_weight=0.5, render_mode=2, dither_mode=1,
        projection=[1.2, 0, 0, 0, 0, 2.1, 0, 0, 0, 0, -1.00004, -1, 0, 0, -0.004, 0], look_at=(0.1, -0.2, 0.3),
        rotation=(1.0, 0, 0, 0, 0.6, 0.8, 0, -0.8, 0.6), distance=-3.25, metadata='{"time step": 7}',
    )


def _pairs(text):
    return json.loads(text, object_pairs_hook=lambda kv: kv)


def test_register_message_structure():
    s = _state()
    sess = M.Session()
    text = sess.register(s, "sim", 1, (640, 480), (128, 96, 64))
    pairs = _pairs(text)
    assert [k for k, _ in pairs] == REGISTER_KEYS
    d = dict(pairs)
    assert d["type"] == "register" and d["name"] == "sim" and d["nodes"] == 1
    assert d["metadata"] == [("time step", 7)]
    assert (d["framebuffer width"], d["framebuffer height"]) == (640, 480)
    assert [dict(f)["name"] for f in d["functors"]] == ["idem", "add", "mul", "length", "pow", "sum"]
    assert [dict(f)["uses parameter"] for f in d["functors"]] == [False, True, True, False, True, False]
    assert d["projection"] == s.projection and d["rotation"] == list(s.rotation) and d["distance"] == -3.25
    assert [dict(x)["name"] for x in d["sources"]] == s.source_names
    assert (d["width"], d["height"], d["depth"], d["protocol"]) == (128, 96, 64, [1, 0])
    assert d["step"] == float(np.float32(0.3))


def test_period_message_sections_and_init_root():
    s = _state()
    sess = M.Session()
    sess.register(s, "sim", 1, (640, 480), (128, 96, 64))
    text = sess.period(s, 5, M.SEND)
    pairs = _pairs(text)
    assert [k for k, _ in pairs] == PERIOD_KEYS
    d = dict(pairs)
    assert d["meta nr"] == 5 and d["type"] == "period"
    arr = np.asarray(d["transfer array"])
    assert arr.shape == (3, 8, 4)
    assert (arr == (s.transfer_lut * np.float32(255.0)).astype(np.uint32)).all()
    assert [dict(p)["value"] for p in d["transfer points"][0]] == [0, 8]
    assert d["functions"][1] == [("source", "length | mul(2)"), ("error", 0)]
    assert len(d["iso threshold"]) == 2 and len(d["minmax"]) == 3 and len(d["clipping"]) == 2
    # the init root: the register root with updated values, new keys appended
    init = d["init"]
    keys = [k for k, _ in init]
    assert keys[: len(REGISTER_KEYS)] == REGISTER_KEYS
    assert keys[len(REGISTER_KEYS):] == ["render optimization", "seed points", "advection step", "advection weight",
                                         "advection seeding period", "advection seeding duration",
                                         "advection on pause", "advection border", "background color", "render mode",
                                         "dither mode"]
    assert dict(init)["type"] == "register"
    # no init-root section sent: no "init"
    text2 = sess.period(s, 6, ["weight", "transfer"])
    assert "init" not in dict(_pairs(text2))
    # a later step change updates the kept root in place
    s.step = 0.75
    d3 = dict(_pairs(sess.period(s, 7, ["step"])))
    assert dict(d3["init"])["step"] == 0.75
    assert [k for k, _ in d3["init"]] == keys


def test_exit_and_server_bytes():
    assert M.exit_message() == '{"type": "exit"}'
    # content[0] ('{') becomes ',': no space after it
    assert M.server_send_bytes(3, '{"a": 1, "b": 2}', False) == '{"uid": 3,"a": 1, "b": 2'
    assert M.server_send_bytes(-1, '{"payload": "x"}', True) == ',"payload": "x"} '
    assert M.server_send_bytes(-1, None, True) == "} "
    assert M.server_send_bytes(0, '{"type": "exit"}', True) == '{"uid": 0,"type": "exit"} '
    # a periodic message followed by its frame reads as one JSON object
    period = M.server_send_bytes(9, '{"type": "period"}', False)
    frame = M.server_send_bytes(-1, '{"payload": "data:image/jpeg;base64,AA"}', True)
    assert json.loads(period + frame) == {"uid": 9, "type": "period", "payload": "data:image/jpeg;base64,AA"}


# ------------------------------------------------------------------ the library against the oracle


def test_library_messages_equal_oracle(hip_lib):
    from isaac_amd import messages as L

    s = _state(T=16)
    ora, ses = M.Session(), L.Session()
    assert ses.register(s, "sim", 4, (1920, 1080), (256, 256, 256)) == ora.register(s, "sim", 4, (1920, 1080),
                                                                                  (256, 256, 256))
    plans = [list(M.SEND), [], ["weight"], ["step", "render_mode"], ["init"], ["ao", "clipping", "min_max"],
             ["projection", "distance"], ["transfer", "functions", "iso_threshold"]]
    for k, send in enumerate(plans):
        if k == 3:
            s.step, s.render_mode = 0.125, 5
        if k == 6:
            s.distance, s.projection = -2.5, [float(x) * 1.5 for x in s.projection]
        got, want = ses.period(s, k, send), ora.period(s, k, send)
        assert got == want, (k, send)
    assert L.exit_message() == M.exit_message()
    for uid, content, fin in ((3, '{"a": 1}', 0), (-1, '{"payload": "x"}', 1), (-1, None, 1), (0, "{}", 1),
                              (2, "x", 1)):
        assert L.server_send_bytes(uid, content, fin) == M.server_send_bytes(uid, content, bool(fin))
    ses.close()


def test_library_message_errors(hip_lib):
    import ctypes

    from isaac_amd import messages as L

    s = _state()
    ses = L.Session()
    st, keep = L.steering_state(s)
    n = ctypes.c_int64(0)
    buf = ctypes.create_string_buffer(64)
    # the periodic message needs the register message first
    assert hip_lib.isaac_hip_period_message(ses._s, ctypes.byref(st), 0, 0, buf, 64, ctypes.byref(n)) == -3
    fb = (ctypes.c_int32 * 2)(8, 8)
    gs = (ctypes.c_int32 * 3)(4, 4, 4)
    # too small: the length is reported, nothing is written past the capacity
    rc = hip_lib.isaac_hip_register_message(ses._s, ctypes.byref(st), b"x", 1, fb, gs, buf, 64, ctypes.byref(n))
    assert rc == -1 and n.value > 64
    assert ses.register(s, "x", 1, (8, 8), (4, 4, 4))
    ses.close()
