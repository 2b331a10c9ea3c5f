"""Test infrastructure: a Python restatement of the master rank's JSON messages to
the ISAAC server and of serverSend's framing, the checker of the library's
isaac_hip_*_message functions.  Nothing in the library imports it.

Follows
  register message   lib/isaac.hpp:834-892 (root made by recreateJSON, 2870-2875;
                     kept as jsonInitRoot by init, 1288-1291)
  periodic message   lib/isaac.hpp:2546-2838 (the send* sections in code order,
                     jsonInitRoot updates, "init" last)
  exit message       lib/isaac.hpp:2110-2118
  serverSend bytes   lib/isaac/isaac_communicator.hpp:125-190
  functor pool       lib/isaac/isaac_functors.hpp:42-219 (names, descriptions,
                     usesParameter), protocol version lib/isaac/isaac_version.hpp:7-8

jansson (the reference's JSON library, an unpinned dependency absent here:
INSTALL.md:45-67 installs the distribution package or git master) writes with
json_dumps(root, 0): objects in insertion order (jansson >= 2.8), ", " between
items, ": " after keys; integers "%lld"; reals "%.17g", ".0" appended when neither
'.' nor 'e' occurs, the exponent's '+' and leading zeros dropped; json_real refuses
NaN and infinities (the key or element is then left out); strings escape '"', '\\',
\\b \\f \\n \\r \\t and other control characters as \\u%04X.  That writer is
restated here from jansson's published source (dump.c, strconv.c) and pinned by
tests/test_message_oracle.py against hand-derived texts and json.loads."""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

FUNCTORS = [
    ("idem", "Does nothing. Keeps the feature dimension.", False),
    ("add", "Summarizes the input with a constant parameter. Keeps the feature dimension.", True),
    ("mul", "Multiplies the input with a constant parameter. Keeps the feature dimension.", True),
    ("length", "Calculates the length of an input. Reduces the feature dimension to 1.", False),
    ("pow", "Calculates the power of the input with a constant exponent. Keeps the feature dimension.", True),
    ("sum", "Calculates the sum of all components. Reduces the feature dimension to 1.", False),
]
MAX_FUNCTORS = 3
PROTOCOL = (1, 0)

# send flags in the order the periodic message tests them
SEND = [
    "projection", "look_at", "rotation", "distance", "transfer", "functions", "weight", "interpolation",
    "render_optimization", "step", "seed_points", "advection_step", "advection_weight", "advection_seeding_period",
    "advection_seeding_duration", "advection_on_pause", "advection_border", "iso_threshold", "min_max", "background",
    "clipping", "ao", "render_mode", "dither_mode", "init",
]
FLAG = {name: 1 << i for i, name in enumerate(SEND)}


# ------------------------------------------------------------------ jansson's writer
class Raw(str):
    """Already-written JSON text (e.g. the caller's metadata object)."""


def real(v: float) -> Optional[str]:
    v = float(v)
    if math.isnan(v) or math.isinf(v):
        return None
    s = "%.17g" % v
    if "e" not in s:
        return s if "." in s else s + ".0"
    mant, exp = s.split("e")
    neg = exp.startswith("-")
    digits = exp.lstrip("+-").lstrip("0") or "0"
    return mant + "e" + ("-" if neg else "") + digits


def string(s: str) -> str:
    out = ['"']
    for ch in s:
        c = ord(ch)
        if ch == "\\":
            out.append("\\\\")
        elif ch == '"':
            out.append('\\"')
        elif ch == "\b":
            out.append("\\b")
        elif ch == "\f":
            out.append("\\f")
        elif ch == "\n":
            out.append("\\n")
        elif ch == "\r":
            out.append("\\r")
        elif ch == "\t":
            out.append("\\t")
        elif c < 0x20:
            out.append("\\u%04X" % c)
        else:
            out.append(ch)
    out.append('"')
    return "".join(out)


class Real(float):
    """A JSON real (json_real); plain floats are written as reals too."""


def dump(v) -> Optional[str]:
    """json_dumps(v, 0); None for a value jansson refuses."""
    if isinstance(v, Raw):
        return str(v)
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return "%d" % v
    if isinstance(v, float):
        return real(v)
    if isinstance(v, str):
        return string(v)
    if isinstance(v, list):
        items = [t for t in (dump(x) for x in v) if t is not None]
        return "[" + ", ".join(items) + "]"
    if isinstance(v, Obj):
        items = [(k, dump(x)) for k, x in v.items]
        return "{" + ", ".join(string(k) + ": " + t for k, t in items if t is not None) + "}"
    raise TypeError(type(v))


class Obj:
    """An ordered JSON object; set() is json_object_set (replace in place or append;
    a refused value leaves the object unchanged)."""

    def __init__(self, items: Optional[List[Tuple[str, object]]] = None):
        self.items: List[Tuple[str, object]] = list(items or [])

    def set(self, key: str, value) -> None:
        if dump(value) is None:
            return
        for i, (k, _) in enumerate(self.items):
            if k == key:
                self.items[i] = (key, value)
                return
        self.items.append((key, value))

    def copy(self) -> "Obj":
        return Obj(self.items)


# ------------------------------------------------------------------ messages
@dataclass
class State:
    """The visualization state the messages report (combined source order)."""

    source_names: List[str]
    feature_dims: List[int]
    n_volume_field: int
    functions: List[str] = field(default_factory=list)
    function_errors: List[int] = field(default_factory=list)
    transfer_size: int = 0
    transfer_lut: Optional[Sequence] = None  # (n_sources, T, 4) float32
    transfer_points: List[List[Tuple[int, Tuple[float, float, float, float]]]] = field(default_factory=list)
    weights: List[float] = field(default_factory=list)
    iso_thresholds: List[float] = field(default_factory=list)
    min_max: List[Tuple[float, float]] = field(default_factory=list)
    step: float = 0.5
    interpolation: bool = False
    render_optimization: bool = True
    seed_points: int = 1000
    advection_step: float = 1.0
    advection_weight: float = 0.95
    advection_seeding_period: int = 1
    advection_seeding_duration: int = 1
    advection_on_pause: bool = False
    advection_border: bool = False
    background: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    clipping: List[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = field(default_factory=list)
    ao_enabled: bool = False
    ao_weight: float = 0.5
    render_mode: int = 0
    dither_mode: int = 1
    projection: Sequence[float] = tuple([1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0])
    look_at: Sequence[float] = (0.0, 0.0, 0.0)
    rotation: Sequence[float] = (1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0)
    distance: float = -4.5
    metadata: str = "{}"


def f32(x) -> float:
    """A float member of the reference (isaac_float) as jansson receives it."""
    import numpy as np

    return float(np.float32(x))


def reals(v, single=False) -> list:
    return [f32(x) if single else float(x) for x in v]


class Session:
    """jsonInitRoot across the master rank's messages."""

    def __init__(self):
        self.init: Optional[Obj] = None

    def register(self, s: State, name: str, nodes: int, framebuffer, global_size_scaled) -> str:
        r = Obj([("metadata", Raw(s.metadata or "{}"))])
        r.set("type", "register")
        r.set("name", name)
        r.set("nodes", int(nodes))
        r.set("framebuffer width", int(framebuffer[0]))
        r.set("framebuffer height", int(framebuffer[1]))
        r.set("max functors", MAX_FUNCTORS)
        r.set("functors", [Obj([("name", n), ("description", d), ("uses parameter", u)]) for n, d, u in FUNCTORS])
        r.set("projection count", 1)
        r.set("projection", reals(s.projection))
        r.set("position", reals(s.look_at))
        r.set("rotation", reals(s.rotation))
        r.set("distance", float(s.distance))
        r.set("sources", [Obj([("name", n), ("feature dimension", int(d))]) for n, d in zip(s.source_names, s.feature_dims)])
        r.set("interpolation", bool(s.interpolation))
        r.set("step", f32(s.step))
        r.set("dimension", 3)
        r.set("width", int(global_size_scaled[0]))
        r.set("height", int(global_size_scaled[1]))
        r.set("depth", int(global_size_scaled[2]))
        r.set("protocol", list(PROTOCOL))
        r.set("ao isEnabled", bool(s.ao_enabled))
        r.set("ao weight", f32(s.ao_weight))
        self.init = r
        return dump(r)

    def period(self, s: State, meta_nr: int, send: Sequence[str]) -> str:
        if self.init is None:
            raise RuntimeError("register first")
        send = set(send)
        init = self.init.copy()
        send_init = "init" in send
        r = Obj([("metadata", Raw(s.metadata or "{}"))])
        r.set("type", "period")
        r.set("meta nr", int(meta_nr))

        def both(key, value):
            nonlocal send_init
            r.set(key, value)
            init.set(key, value)
            send_init = True

        if "projection" in send:
            both("projection", reals(s.projection))
        if "look_at" in send:
            both("position", reals(s.look_at))
        if "rotation" in send:
            both("rotation", reals(s.rotation))
        if "distance" in send:
            both("distance", float(s.distance))
        if "transfer" in send:
            import numpy as np

            lut = np.asarray(s.transfer_lut, np.float32).reshape(len(s.source_names), s.transfer_size, 4)
            # isaac_uint(x * isaac_float(255)) per channel
            arr = [[[int(np.uint32(np.float32(c) * np.float32(255.0))) for c in px] for px in src] for src in lut]
            r.set("transfer array", arr)
            r.set("transfer points", [[Obj([("value", int(v)), ("r", f32(c[0])), ("g", f32(c[1])), ("b", f32(c[2])),
                                             ("a", f32(c[3]))]) for v, c in pts] for pts in s.transfer_points])
        if "functions" in send:
            r.set("functions", [Obj([("source", f), ("error", int(e))]) for f, e in zip(s.functions, s.function_errors)])
        if "weight" in send:
            r.set("weight", reals(s.weights, True))
        if "interpolation" in send:
            both("interpolation", bool(s.interpolation))
        if "render_optimization" in send:
            both("render optimization", bool(s.render_optimization))
        if "step" in send:
            both("step", f32(s.step))
        if "seed_points" in send:
            both("seed points", int(s.seed_points))
        if "advection_step" in send:
            both("advection step", f32(s.advection_step))
        if "advection_weight" in send:
            both("advection weight", f32(s.advection_weight))
        if "advection_seeding_period" in send:
            both("advection seeding period", int(s.advection_seeding_period))
        if "advection_seeding_duration" in send:
            both("advection seeding duration", int(s.advection_seeding_duration))
        if "advection_on_pause" in send:
            both("advection on pause", bool(s.advection_on_pause))
        if "advection_border" in send:
            both("advection border", bool(s.advection_border))
        if "iso_threshold" in send:
            r.set("iso threshold", reals(s.iso_thresholds[: s.n_volume_field], True))
        if "min_max" in send:
            r.set("minmax", [Obj([("min", f32(a)), ("max", f32(b))]) for a, b in s.min_max])
        if "background" in send:
            both("background color", reals(s.background, True))
        if "clipping" in send:
            r.set("clipping", [Obj([("position", reals(p, True)), ("normal", reals(n, True))]) for p, n in s.clipping])
        if "ao" in send:
            both("ao isEnabled", bool(s.ao_enabled))
            both("ao weight", f32(s.ao_weight))
        if "render_mode" in send:
            both("render mode", int(s.render_mode))
        if "dither_mode" in send:
            both("dither mode", int(s.dither_mode))
        if send_init:
            r.set("init", init)
        self.init = init
        return dump(r)


def exit_message() -> str:
    return dump(Obj([("type", "exit")]))


def server_send_bytes(uid: int, content: Optional[str], finishing: bool) -> str:
    """serverSend(content, starting = uid >= 0, finishing)."""
    out = ""
    if uid >= 0:
        out += '{"uid": %d' % uid
    if content is not None:
        if len(content) >= 2:
            out += "," + content[1:-1]
    if finishing:
        out += "} "
    return out
