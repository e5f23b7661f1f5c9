"""Reader of wavefunction files written by the reference's
``Wavefunction.save`` (reference ``src/fqe/wavefunction.py:726-765``), with
no code execution.

The reference pickles the list

    [symmetry map, conserved, conserve_spin, conserve_number, norb,
     [key, FqeData], [key, FqeData], ...]

so reading it with ``pickle.load`` would run whatever callables the file
names.  Here the pickle byte stream is only parsed: ``pickletools.genops``
yields its opcodes, and a small stack machine rebuilds plain data (None,
bools, numbers, strings, bytes, tuples, lists, dicts, sets) from them.  A
reference to a global (``GLOBAL`` / ``STACK_GLOBAL``) becomes an inert
``Global`` record: nothing is imported.  A call of a global (``REDUCE``,
``NEWOBJ``, ``INST``, ``OBJ``) becomes an inert ``Instance`` or
``Opaque`` record, except for the few pure-data constructors a numpy array
needs, which are rebuilt here from their arguments (``numpy.dtype`` from its
type string, ``ndarray`` from its shape, dtype and raw bytes, numpy scalars,
``complex``, and ``_codecs.encode(text, "latin1")``, the byte strings of
protocols 0-2).  Persistent ids, the extension registry and out-of-band
buffers are refused.  Finally the data must have the layout above, with
``fqe.fqe_data.FqeData`` sectors whose ``coeff`` arrays have the sector's
shape; anything else raises ``ValueError``.
"""

from __future__ import annotations

import math
import pickletools
from dataclasses import dataclass
from typing import Any, Dict, List, Tuple

import numpy

__all__ = ["Global", "Instance", "Opaque", "parse_pickle",
           "read_reference_wavefunction"]

# Refuse streams or arrays beyond these sizes (a corrupt or hostile length
# field must not allocate without bound).
MAX_FILE_BYTES = 1 << 36
MAX_ARRAY_BYTES = 1 << 36


@dataclass(frozen=True)
class Global:
    """A reference to module.name in the stream; never imported."""
    module: str
    name: str


@dataclass
class Instance:
    """An object of class ``cls`` created by NEWOBJ / OBJ / INST or a
    ``copyreg`` reconstructor; ``state`` is what BUILD gave it."""
    cls: Global
    args: tuple = ()
    state: Any = None


@dataclass
class Opaque:
    """A call of a global that is not a known data constructor: kept as a
    record, never performed."""
    func: Any
    args: Any
    state: Any = None


@dataclass
class _ArrayStub:
    """numpy's ``_reconstruct(ndarray, (0,), b'b')``, completed by BUILD."""
    array: Any = None


@dataclass
class _DtypeStub:
    """numpy's ``dtype(str, align, copy)``, completed by BUILD (byte
    order)."""
    spec: str
    dtype: Any = None


_MARK = object()


def _is_numpy(g: Any, *names: str) -> bool:
    return (isinstance(g, Global) and g.module.split(".")[0] == "numpy"
            and g.name in names)


def _dtype_of(obj: Any) -> numpy.dtype:
    if isinstance(obj, _DtypeStub):
        return obj.dtype if obj.dtype is not None else numpy.dtype(obj.spec)
    if isinstance(obj, numpy.dtype):
        return obj
    raise ValueError(f"not a numpy dtype in the stream: {obj!r}")


def _safe_dtype(spec: Any) -> numpy.dtype:
    """numpy dtype from a type string of a plain numeric or boolean type
    (no object, structured or void types)."""
    if not isinstance(spec, str) or len(spec) > 8:
        raise ValueError(f"unsupported dtype in the stream: {spec!r}")
    dt = numpy.dtype(spec)
    if dt.kind not in "biufc" or dt.fields is not None:
        raise ValueError(f"unsupported dtype in the stream: {spec!r}")
    return dt


def _array_from(shape: Any, dtype: numpy.dtype, fortran: Any,
                raw: Any) -> numpy.ndarray:
    if (not isinstance(shape, tuple) or
            not all(isinstance(n, int) and n >= 0 for n in shape)):
        raise ValueError("bad array shape in the stream")
    if not isinstance(raw, (bytes, bytearray)):
        raise ValueError("array data is not a byte string")
    count = math.prod(shape)
    if count * dtype.itemsize != len(raw) or len(raw) > MAX_ARRAY_BYTES:
        raise ValueError("array data does not match its shape and dtype")
    arr = numpy.frombuffer(bytes(raw), dtype=dtype, count=count)
    return arr.reshape(shape, order="F" if fortran else "C").copy()


def _call(func: Any, args: Any) -> Any:
    """The result of REDUCE(func, args): a rebuilt value for the numpy and
    builtin data constructors, an inert record otherwise."""
    if not isinstance(args, tuple):
        raise ValueError("REDUCE arguments are not a tuple")
    if _is_numpy(func, "_reconstruct"):
        return _ArrayStub()
    if _is_numpy(func, "dtype"):
        if not args:
            raise ValueError("numpy.dtype without arguments")
        return _DtypeStub(args[0], _safe_dtype(args[0]))
    if _is_numpy(func, "scalar"):
        # numpy scalar: scalar(dtype, raw bytes)
        if len(args) != 2:
            raise ValueError("bad numpy scalar in the stream")
        dt = _dtype_of(args[0])
        return _array_from((), dt, False, args[1])[()]
    if _is_numpy(func, "_frombuffer"):
        # protocol-5 form: _frombuffer(buffer, dtype, shape, order)
        if len(args) != 4:
            raise ValueError("bad numpy buffer in the stream")
        return _array_from(tuple(args[2]), _dtype_of(args[1]),
                           args[3] == "F", args[0])
    if (isinstance(func, Global) and func.module == "_codecs"
            and func.name == "encode"):
        # protocols 0-2 write bytes as _codecs.encode(text, "latin1")
        if (len(args) == 2 and isinstance(args[0], str)
                and args[1] in ("latin1", "latin-1")):
            return args[0].encode("latin-1")
        raise ValueError("bad byte string in the stream")
    if (isinstance(func, Global) and func.module in ("builtins", "__builtin__")
            and func.name == "complex"):
        if not all(isinstance(x, (int, float)) for x in args) or len(args) > 2:
            raise ValueError("bad complex number in the stream")
        return complex(*args)
    if (isinstance(func, Global) and func.module == "copyreg"
            and func.name == "_reconstructor" and args
            and isinstance(args[0], Global)):
        return Instance(args[0], tuple(args[1:]))
    return Opaque(func, args)


def _build(obj: Any, state: Any) -> Any:
    """BUILD: complete an array or dtype stub, or record the state."""
    if isinstance(obj, _ArrayStub):
        if not (isinstance(state, tuple) and len(state) == 5):
            raise ValueError("bad ndarray state in the stream")
        _, shape, dt, fortran, raw = state
        obj.array = _array_from(shape, _dtype_of(dt), fortran, raw)
        return obj
    if isinstance(obj, _DtypeStub):
        if (isinstance(state, tuple) and len(state) >= 2
                and state[1] in ("<", ">")):
            obj.dtype = obj.dtype.newbyteorder(state[1])
        return obj
    if isinstance(obj, (Instance, Opaque)):
        obj.state = state
        return obj
    raise ValueError(f"BUILD on {type(obj).__name__} in the stream")


def _finish(obj: Any, seen: Any = None) -> Any:
    """Replace array and dtype stubs by their values, recursively (each
    container once: the memo can share or nest objects)."""
    if seen is None:
        seen = set()
    if isinstance(obj, (list, dict, Instance, Opaque)):
        if id(obj) in seen:
            return obj
        seen.add(id(obj))
    if isinstance(obj, _ArrayStub):
        if obj.array is None:
            raise ValueError("ndarray without data in the stream")
        return obj.array
    if isinstance(obj, _DtypeStub):
        return obj.dtype
    if isinstance(obj, list):
        obj[:] = [_finish(x, seen) for x in obj]
        return obj
    if isinstance(obj, tuple):
        return tuple(_finish(x, seen) for x in obj)
    if isinstance(obj, dict):
        items = [(_finish(k, seen), _finish(v, seen)) for k, v in obj.items()]
        obj.clear()
        obj.update(items)
        return obj
    if isinstance(obj, (Instance, Opaque)):
        obj.state = _finish(obj.state, seen)
        obj.args = _finish(obj.args, seen)
        return obj
    return obj


def parse_pickle(data: bytes) -> Any:
    """The data of a pickle byte stream, rebuilt from its opcodes without
    importing or calling anything named in it (see the module docstring)."""
    if len(data) > MAX_FILE_BYTES:
        raise ValueError("file too large")
    stack: List[Any] = []
    memo: Dict[int, Any] = {}

    def pop_mark() -> List[Any]:
        for i in range(len(stack) - 1, -1, -1):
            if stack[i] is _MARK:
                items = stack[i + 1:]
                del stack[i:]
                return items
        raise ValueError("MARK missing in the stream")

    try:
        ops = list(pickletools.genops(data))
    except Exception as err:        # truncated or not a pickle
        raise ValueError(f"not a pickle stream: {err}") from None
    for op, arg, _ in ops:
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "STOP":
            if len(stack) != 1:
                raise ValueError("stack not empty at STOP")
            return _finish(stack[0])
        if name == "MARK":
            stack.append(_MARK)
        elif name == "POP":
            stack.pop()
        elif name == "POP_MARK":
            pop_mark()
        elif name == "DUP":
            stack.append(stack[-1])
        elif name == "NONE":
            stack.append(None)
        elif name in ("NEWTRUE", "NEWFALSE"):
            stack.append(name == "NEWTRUE")
        elif name in ("INT", "BININT", "BININT1", "BININT2", "LONG", "LONG1",
                      "LONG4", "FLOAT", "BINFLOAT"):
            stack.append(arg)
        elif name in ("STRING", "BINSTRING", "SHORT_BINSTRING", "UNICODE",
                      "BINUNICODE", "SHORT_BINUNICODE", "BINUNICODE8"):
            stack.append(arg if isinstance(arg, str) else
                         arg.decode("latin-1"))
        elif name in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8"):
            stack.append(bytes(arg))
        elif name == "BYTEARRAY8":
            stack.append(bytearray(arg))
        elif name == "EMPTY_LIST":
            stack.append([])
        elif name == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif name == "APPENDS":
            items = pop_mark()
            stack[-1].extend(items)
        elif name == "LIST":
            stack.append(pop_mark())
        elif name == "EMPTY_TUPLE":
            stack.append(())
        elif name == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif name in ("TUPLE1", "TUPLE2", "TUPLE3"):
            n = int(name[-1])
            items = tuple(stack[-n:])
            del stack[-n:]
            stack.append(items)
        elif name == "EMPTY_DICT":
            stack.append({})
        elif name == "DICT":
            items = pop_mark()
            stack.append(dict(zip(items[0::2], items[1::2])))
        elif name == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif name == "SETITEMS":
            items = pop_mark()
            for k, v in zip(items[0::2], items[1::2]):
                stack[-1][k] = v
        elif name == "EMPTY_SET":
            stack.append(set())
        elif name == "ADDITEMS":
            items = pop_mark()
            stack[-1].update(items)
        elif name == "FROZENSET":
            stack.append(frozenset(pop_mark()))
        elif name in ("GET", "BINGET", "LONG_BINGET"):
            stack.append(memo[int(arg)])
        elif name in ("PUT", "BINPUT", "LONG_BINPUT"):
            memo[int(arg)] = stack[-1]
        elif name == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif name == "GLOBAL":
            module, gname = arg.split(" ", 1)
            stack.append(Global(module, gname))
        elif name == "STACK_GLOBAL":
            gname = stack.pop()
            module = stack.pop()
            if not (isinstance(module, str) and isinstance(gname, str)):
                raise ValueError("STACK_GLOBAL of non-strings")
            stack.append(Global(module, gname))
        elif name == "REDUCE":
            args = stack.pop()
            func = stack.pop()
            stack.append(_call(func, args))
        elif name == "NEWOBJ":
            args = stack.pop()
            cls = stack.pop()
            if not isinstance(cls, Global):
                raise ValueError("NEWOBJ of a non-class")
            stack.append(Instance(cls, tuple(args)))
        elif name == "NEWOBJ_EX":
            stack.pop()                      # keyword arguments
            args = stack.pop()
            cls = stack.pop()
            if not isinstance(cls, Global):
                raise ValueError("NEWOBJ_EX of a non-class")
            stack.append(Instance(cls, tuple(args)))
        elif name == "OBJ":
            items = pop_mark()
            if not items or not isinstance(items[0], Global):
                raise ValueError("OBJ of a non-class")
            stack.append(Instance(items[0], tuple(items[1:])))
        elif name == "INST":
            module, gname = arg.split(" ", 1)
            stack.append(Instance(Global(module, gname), tuple(pop_mark())))
        elif name == "BUILD":
            state = stack.pop()
            stack[-1] = _build(stack[-1], state)
        else:
            # PERSID, BINPERSID, EXT1/2/4, NEXT_BUFFER, READONLY_BUFFER
            raise ValueError(f"pickle opcode {name} is not supported")
    raise ValueError("pickle stream without STOP")


def _state_dict(obj: Any) -> Dict[str, Any]:
    state = obj.state
    # (dict, slot dict) when the class has __slots__
    if isinstance(state, tuple) and len(state) == 2:
        state = state[0]
    if not isinstance(state, dict):
        raise ValueError("object without a state dictionary")
    return state


def read_reference_wavefunction(data: bytes) -> Dict[str, Any]:
    """The fields of a reference wavefunction file: ``symmetry_map``,
    ``conserved``, ``conserve_spin``, ``conserve_number``, ``norb`` and
    ``sectors`` ({(nele, sz): complex coefficient array}), checked against
    the layout of ``src/fqe/wavefunction.py:726-765``."""
    top = parse_pickle(data)
    if not isinstance(top, list) or len(top) < 5:
        raise ValueError("not a reference wavefunction file: expected a list "
                         "[symmetry map, conserved, conserve_spin, "
                         "conserve_number, norb, sectors...]")
    symmetry_map, conserved, conserve_spin, conserve_number, norb = top[:5]
    if not (isinstance(symmetry_map, dict) and isinstance(conserved, dict)
            and isinstance(conserve_spin, bool)
            and isinstance(conserve_number, bool)
            and isinstance(norb, int) and norb > 0):
        raise ValueError("not a reference wavefunction file: bad header")
    sectors: Dict[Tuple[int, int], numpy.ndarray] = {}
    for item in top[5:]:
        if not (isinstance(item, list) and len(item) == 2):
            raise ValueError("not a reference wavefunction file: bad sector")
        key, sector = item
        if not (isinstance(key, tuple) and len(key) == 2
                and all(isinstance(x, int) for x in key)):
            raise ValueError("not a reference wavefunction file: bad key")
        if not (isinstance(sector, Instance) and
                sector.cls == Global("fqe.fqe_data", "FqeData")):
            raise ValueError("not a reference wavefunction file: a sector "
                             "is not an fqe.fqe_data.FqeData")
        coeff = _state_dict(sector).get("coeff")
        if not isinstance(coeff, numpy.ndarray) or coeff.ndim != 2:
            raise ValueError("not a reference wavefunction file: a sector "
                             "has no coefficient matrix")
        nele, sz = key
        nalpha, nbeta = (nele + sz) // 2, (nele - sz) // 2
        want = (math.comb(norb, nalpha), math.comb(norb, nbeta))
        if ((nele + sz) % 2 or nalpha < 0 or nbeta < 0 or
                coeff.shape != want or coeff.dtype.kind != "c"):
            raise ValueError(f"not a reference wavefunction file: sector "
                             f"{key} has shape {coeff.shape}, expected {want}")
        sectors[key] = coeff
    return {"symmetry_map": symmetry_map, "conserved": conserved,
            "conserve_spin": conserve_spin,
            "conserve_number": conserve_number, "norb": norb,
            "sectors": sectors}
