"""Static reader for the pickled table files PETITE ships.

PETITE stores its tables as pickles (``sm_xsec.pkl``, ``dark_xsec.pkl``,
``dark_weights.pkl``, ``dark_drate.pkl``) and as ``.npy`` object arrays whose
payload is a pickle (``<P>_AdaptiveMaps.npy``, rows ``[params, AdaptiveMap]``;
see src/PETITE/shower.py:153-175 and utilities/generate_integrators.py:226-245
of the reference for how they are written and consumed).

This module never unpickles.  It walks the opcode stream with
``pickletools.genops`` (a disassembler) and interprets a fixed set of data
opcodes with its own stack machine.  Globals named in the stream are never
imported or called: each is matched against a short whitelist of names that
denote plain data (numpy arrays, dtypes and scalars, and the grid container of
a VEGAS adaptive map) and rebuilt here from the raw bytes/lists in the stream.
Any other global, or any opcode outside the interpreted set, raises.
"""
from __future__ import annotations

import pickletools
import struct

import numpy as np

_RECONSTRUCT = {("numpy._core.multiarray", "_reconstruct"), ("numpy.core.multiarray", "_reconstruct")}
_SCALAR = {("numpy._core.multiarray", "scalar"), ("numpy.core.multiarray", "scalar")}
_NDARRAY = {("numpy", "ndarray")}
_DTYPE = {("numpy", "dtype")}
_ADAPTIVE_MAP = {("vegas._vegas", "AdaptiveMap"), ("vegas", "AdaptiveMap")}
_ALLOWED = _RECONSTRUCT | _SCALAR | _NDARRAY | _DTYPE | _ADAPTIVE_MAP


class AdaptiveMapGrid:
    """Inert stand-in for a pickled VEGAS AdaptiveMap: only its node lists."""

    def __init__(self, grid):
        self.grid = [np.asarray(g, dtype=np.float64) for g in grid]

    def __repr__(self):
        return f"AdaptiveMapGrid(ndim={len(self.grid)}, nodes={[len(g) for g in self.grid]})"


class _Global:
    __slots__ = ("module", "name")

    def __init__(self, module, name):
        if (module, name) not in _ALLOWED:
            raise ValueError(f"static_pickle: refusing global {module}.{name}")
        self.module, self.name = module, name

    @property
    def key(self):
        return (self.module, self.name)


class _Dtype:
    def __init__(self, code):
        self.code = code
        self.order = "|"

    def to_numpy(self):
        code = self.code
        if code[0] in "<>|=":
            return np.dtype(code)
        if code.startswith("O"):
            return np.dtype(object)
        order = self.order if self.order in "<>" else ("<" if code[0] not in "bBS?" else "|")
        return np.dtype(order + code) if code[0] not in "b?" else np.dtype(code)


class _ArrayStub:
    def __init__(self):
        self.value = None


class _Mark:
    pass


_MARK = _Mark()


def _pop_mark(stack):
    items = []
    while True:
        x = stack.pop()
        if x is _MARK:
            items.reverse()
            return items
        items.append(x)


def _finish(x):
    if isinstance(x, _ArrayStub):
        return x.value
    return x


def _reduce(func, args):
    if not isinstance(func, _Global):
        raise ValueError("static_pickle: REDUCE on a non-global")
    key = func.key
    if key in _RECONSTRUCT:
        return _ArrayStub()
    if key in _DTYPE:
        return _Dtype(args[0])
    if key in _SCALAR:
        dt, raw = args
        return np.frombuffer(raw, dtype=dt.to_numpy(), count=1)[0].item()
    if key in _ADAPTIVE_MAP:
        return AdaptiveMapGrid(_finish(args[0]))
    raise ValueError(f"static_pickle: REDUCE of {key}")


def _build(obj, state):
    if isinstance(obj, _Dtype):
        obj.order = state[1]
        return obj
    if isinstance(obj, _ArrayStub):
        _, shape, dt, fortran, payload = state
        dtn = dt.to_numpy()
        if dtn == np.dtype(object):
            arr = np.empty(len(payload), dtype=object)
            for i, v in enumerate(payload):
                arr[i] = _finish(v)
            arr = arr.reshape(shape, order="F" if fortran else "C")
        else:
            arr = np.frombuffer(bytes(payload), dtype=dtn).copy()
            arr = arr.reshape(shape, order="F" if fortran else "C")
        obj.value = arr
        return obj
    raise ValueError(f"static_pickle: BUILD on {type(obj).__name__}")


def loads(data: bytes):
    stack = []
    memo = {}
    for op, arg, _pos in pickletools.genops(data):
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "STOP":
            break
        if name == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif name in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif name in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif name == "MARK":
            stack.append(_MARK)
        elif name in ("SHORT_BINUNICODE", "BINUNICODE", "BINUNICODE8", "UNICODE"):
            stack.append(arg)
        elif name in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8"):
            stack.append(bytes(arg))
        elif name in ("BININT", "BININT1", "BININT2", "LONG1", "LONG4", "INT", "LONG"):
            stack.append(int(arg))
        elif name in ("BINFLOAT", "FLOAT"):
            stack.append(float(arg))
        elif name == "NONE":
            stack.append(None)
        elif name == "NEWTRUE":
            stack.append(True)
        elif name == "NEWFALSE":
            stack.append(False)
        elif name == "EMPTY_LIST":
            stack.append([])
        elif name == "EMPTY_DICT":
            stack.append({})
        elif name == "EMPTY_TUPLE":
            stack.append(())
        elif name == "TUPLE":
            stack.append(tuple(_pop_mark(stack)))
        elif name == "TUPLE1":
            stack[-1:] = [(stack[-1],)]
        elif name == "TUPLE2":
            stack[-2:] = [tuple(stack[-2:])]
        elif name == "TUPLE3":
            stack[-3:] = [tuple(stack[-3:])]
        elif name == "LIST":
            stack.append(list(_pop_mark(stack)))
        elif name == "DICT":
            items = _pop_mark(stack)
            stack.append({_finish(items[i]): _finish(items[i + 1]) for i in range(0, len(items), 2)})
        elif name == "APPEND":
            v = stack.pop()
            stack[-1].append(_finish(v))
        elif name == "APPENDS":
            items = _pop_mark(stack)
            stack[-1].extend(_finish(v) for v in items)
        elif name == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][_finish(k)] = _finish(v)
        elif name == "SETITEMS":
            items = _pop_mark(stack)
            d = stack[-1]
            for i in range(0, len(items), 2):
                d[_finish(items[i])] = _finish(items[i + 1])
        elif name == "STACK_GLOBAL":
            nm = stack.pop()
            mod = stack.pop()
            stack.append(_Global(mod, nm))
        elif name == "GLOBAL":
            mod, nm = arg.split(" ", 1)
            stack.append(_Global(mod, nm))
        elif name == "REDUCE":
            args = stack.pop()
            func = stack.pop()
            stack.append(_reduce(func, args))
        elif name == "BUILD":
            state = stack.pop()
            stack[-1] = _build(stack[-1], state)
        else:
            raise ValueError(f"static_pickle: opcode {name} not interpreted")
    if len(stack) != 1:
        raise ValueError("static_pickle: malformed stream")
    return _finish(stack[0])


def load_pickle_file(path):
    with open(path, "rb") as fh:
        return _deep_finish(loads(fh.read()))


def load_object_npy(path):
    """Read a ``.npy`` v1/v2 file whose dtype is object (payload is a pickle)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if raw[:6] != b"\x93NUMPY":
        raise ValueError(f"{path}: not an .npy file")
    major = raw[6]
    if major == 1:
        hlen = struct.unpack("<H", raw[8:10])[0]
        start = 10 + hlen
    else:
        hlen = struct.unpack("<I", raw[8:12])[0]
        start = 12 + hlen
    return _deep_finish(loads(raw[start:]))


def _deep_finish(x):
    x = _finish(x)
    if isinstance(x, dict):
        return {k: _deep_finish(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_deep_finish(v) for v in x]
    if isinstance(x, tuple):
        return tuple(_deep_finish(v) for v in x)
    if isinstance(x, np.ndarray) and x.dtype == object:
        out = np.empty(x.shape, dtype=object)
        for idx in np.ndindex(x.shape):
            out[idx] = _deep_finish(x[idx])
        return out
    return x
