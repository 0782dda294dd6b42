"""Read circuits written as Julia source in the reference's operation vocabulary.

The reference keeps benchmark circuits as Julia literals, e.g. the
441-qubit surface-code circuit of benchmark/surface-code-circuit.jl:4-17807
(`circ = Any[sMRZ(1, 0), ..., sCNOT(24, 23), ..., sMZ(441, 2751)]`).  This
module turns such text into this package's operation objects without
evaluating any code: a small recursive-descent reader accepts constructor
calls of known operation names whose arguments are integers, floats, tuples,
vectors, unit ranges `a:b` and `P"..."` Pauli literals, and rejects anything
else with ArgumentError.

    ops = parse_circuit(open("surface-code-circuit.jl").read())
"""
from __future__ import annotations

import re

from . import ops as _ops
from ._lib import ArgumentError

_TOKEN = re.compile(r"""
    (?P<ws>\s+|\#[^\n]*)
  | (?P<pstr>P"[^"]*")
  | (?P<num>[-+]?(?:\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+|\d+))
  | (?P<name>[A-Za-z_][A-Za-z0-9_!.]*)
  | (?P<punct>[()\[\],:=;])
""", re.VERBOSE)

# constructor name -> callable taking the parsed positional arguments
_CONSTRUCTORS = {c.__name__: c for c in _ops.SINGLE_QUBIT_GATES + _ops.TWO_QUBIT_GATES}
_CONSTRUCTORS.update({
    "sMZ": _ops.sMZ, "sMRZ": _ops.sMRZ, "sMX": _ops.sMX, "sMRX": _ops.sMRX,
    "sMY": _ops.sMY, "sMRY": _ops.sMRY,
    "PauliError": _ops.PauliError,
    "ClassicalXOR": _ops.ClassicalXOR,
    "PauliMeasurement": _ops.PauliMeasurement,
    "SingleQubitOperator": _ops.SingleQubitOperator,
    "NoiseOp": _ops.NoiseOp, "NoiseOpAll": _ops.NoiseOpAll, "NoisyGate": _ops.NoisyGate,
    "UnbiasedUncorrelatedNoise": _ops.UnbiasedUncorrelatedNoise,
    "PauliNoise": lambda *a: _ops.PauliNoise_(a[0]) if len(a) == 1 else _ops.PauliNoise(*a),
    "DepolarizationNoise": _ops.DepolarizationNoise,
})
_BOOLS = {"true": True, "false": False}


def _tokens(text: str):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if m is None:
            raise ArgumentError(f"unexpected character {text[pos]!r} at offset {pos}")
        pos = m.end()
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group(kind), m.start()))
    out.append(("end", "", len(text)))
    return out


class _Reader:
    def __init__(self, text: str):
        self.t = _tokens(text)
        self.i = 0

    def peek(self, k=0):
        return self.t[min(self.i + k, len(self.t) - 1)]

    def take(self, value=None):
        tok = self.t[self.i]
        if value is not None and tok[1] != value:
            raise ArgumentError(f"expected {value!r} at offset {tok[2]}, found {tok[1]!r}")
        self.i += 1
        return tok

    def expr(self):
        lo = self.atom()
        if self.peek()[1] == ":":  # unit range a:b, both ends included
            self.take(":")
            hi = self.atom()
            if not (isinstance(lo, int) and isinstance(hi, int)):
                raise ArgumentError("ranges need integer ends")
            return tuple(range(lo, hi + 1))
        return lo

    def seq(self, close):
        items = []
        while self.peek()[1] != close:
            items.append(self.expr())
            if self.peek()[1] == ",":
                self.take(",")
            elif self.peek()[1] != close:
                tok = self.peek()
                raise ArgumentError(f"expected ',' or {close!r} at offset {tok[2]}")
        self.take(close)
        return items

    def atom(self):
        kind, val, at = self.take()
        if kind == "num":
            return float(val) if any(c in val for c in ".eE") else int(val)
        if kind == "pstr":
            from .pauli import P
            return P(val[2:-1], device="cpu")
        if kind == "punct" and val == "(":
            items = self.seq(")")
            return items[0] if len(items) == 1 else tuple(items)
        if kind == "punct" and val == "[":
            return tuple(self.seq("]"))
        if kind == "name":
            if val in _BOOLS:
                return _BOOLS[val]
            if self.peek()[1] == "[":  # typed vector literal, e.g. Any[...]
                self.take("[")
                return tuple(self.seq("]"))
            if self.peek()[1] == "(":
                self.take("(")
                args = self.seq(")")
                ctor = _CONSTRUCTORS.get(val)
                if ctor is None:
                    raise ArgumentError(f"unknown operation {val!r} at offset {at}")
                return ctor(*args)
            raise ArgumentError(f"unexpected name {val!r} at offset {at}")
        raise ArgumentError(f"unexpected {val!r} at offset {at}")


def parse_ops(text: str) -> list:
    """Operations of one Julia vector literal or comma/newline separated
    constructor calls, in order."""
    r = _Reader(text)
    v = r.expr()
    if r.peek()[0] != "end":
        raise ArgumentError(f"trailing text at offset {r.peek()[2]}")
    return list(v) if isinstance(v, tuple) else [v]


def parse_circuit(text: str, name: str | None = None) -> list:
    """The operations of the vector assigned in a Julia file such as
    benchmark/surface-code-circuit.jl (`using ...` lines and comments are
    skipped).  With `name`, the assignment `name = ...`; otherwise the first
    assignment in the text."""
    body = re.sub(r"^\s*using\s+[^\n]*$", "", text, flags=re.MULTILINE)
    pat = (re.escape(name) if name else r"[A-Za-z_][A-Za-z0-9_]*") + r"\s*=\s*"
    m = re.search(r"^\s*" + pat, body, flags=re.MULTILINE)
    if m is None:
        raise ArgumentError("no circuit assignment found")
    r = _Reader(body[m.end():])
    v = r.expr()
    return list(v) if isinstance(v, tuple) else [v]
