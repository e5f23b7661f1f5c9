"""Torch-free ctypes binding of libfqehip.so: the module a maintainer of
the reference would add next to its own ctypes bindings
(``src/fqe/lib/__init__.py:1-6``, ``src/fqe/lib/_fqe_data.pyx:609-765``) to
run ``FqeData.apply_inplace`` for (h1, h2) tuples on the GPU.

Device memory comes from the HIP runtime through ctypes (``hipMalloc`` /
``hipMemcpy`` / ``hipFree``), the string tables from the library's own
kernels (``fqehip_string_tables``, ``fqehip_pair_table``,
``fqehip_fused_tables``), the beta-beta block of the split apply from
``fqehip_beta_block``, and the apply from ``fqehip_apply``; numpy holds the
host arrays.  Nothing here imports torch.

    import fqehip_ctypes as fh
    out = fh.apply_spatial12(coeff, h1, h2, nalpha, nbeta, norb)

Tables and prepared operators are cached per (norb, nelectrons) and per
tensor identity, as FQE keeps its FciGraph per sector.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int64, c_size_t, c_void_p
from math import comb
from typing import Dict, Optional, Tuple

import numpy

LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fqe_amd",
                       "libfqehip.so")
_H2D, _D2H = 1, 2
C128 = 0

_hip = None
_lib = None


class Sector(ctypes.Structure):
    """fqehip_sector, field for field."""
    _fields_ = [("norb", c_int), ("nalpha", c_int), ("nbeta", c_int),
                ("npair", c_int), ("lena", c_int64), ("lenb", c_int64),
                ("ptab_a", c_void_p), ("ptab_b", c_void_p),
                ("fwd_a", c_void_p), ("fwd_b", c_void_p),
                ("slot_a", c_void_p), ("nslot", c_int),
                ("hbb_hi", c_void_p), ("hbb_lo", c_void_p), ("ldh", c_int64),
                ("tile", c_int), ("hbb_f64", c_void_p), ("hbb_oz", c_void_p)]


class Operator(ctypes.Structure):
    """fqehip_operator."""
    _fields_ = [("complex_op", c_int), ("a", c_void_p), ("h1", c_void_p)]


class FqeHipError(RuntimeError):
    """A HIP or fqehip call failed."""


def _load():
    global _hip, _lib
    if _lib is None:
        hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [POINTER(c_void_p), c_size_t]
        hip.hipFree.argtypes = [c_void_p]
        hip.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        hip.hipDeviceSynchronize.argtypes = []
        hip.hipGetErrorString.restype = ctypes.c_char_p
        hip.hipGetErrorString.argtypes = [c_int]
        lib = ctypes.CDLL(LIBPATH)
        lib.fqehip_last_error.restype = ctypes.c_char_p
        lib.fqehip_string_tables.argtypes = [c_int, c_int, c_int64, c_void_p,
                                             c_void_p, c_void_p]
        lib.fqehip_pair_table.argtypes = [c_int, c_int, c_int64, c_void_p,
                                          c_int, c_void_p, c_void_p]
        lib.fqehip_fused_tables.argtypes = [c_int, c_int, c_int64, c_void_p,
                                            c_int, c_void_p, c_void_p,
                                            c_void_p, c_void_p]
        lib.fqehip_beta_block_ld.restype = c_int64
        lib.fqehip_beta_block_ld.argtypes = [c_int64]
        lib.fqehip_beta_block.argtypes = [c_int, c_int64, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p,
                                          c_int64, c_void_p]
        lib.fqehip_apply_workspace_bytes.restype = c_int64
        lib.fqehip_apply_workspace_bytes.argtypes = [c_int, POINTER(Sector)]
        lib.fqehip_apply.argtypes = [c_int, POINTER(Sector), POINTER(Operator),
                                     c_void_p, c_double, c_double, c_double,
                                     c_double, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p]
        _hip, _lib = hip, lib
    return _hip, _lib


def _check(rc: int, what: str) -> None:
    if rc:
        raise FqeHipError(f"{what}: {_lib.fqehip_last_error().decode()}")


class DeviceBuffer:
    """hipMalloc'd bytes, freed with the object."""

    def __init__(self, nbytes: int) -> None:
        hip, _ = _load()
        self.nbytes = max(int(nbytes), 16)
        self.ptr = c_void_p()
        rc = hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes)
        if rc:
            raise MemoryError(f"hipMalloc({self.nbytes}): "
                              f"{hip.hipGetErrorString(rc).decode()}")

    @classmethod
    def from_array(cls, arr: numpy.ndarray) -> "DeviceBuffer":
        arr = numpy.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: numpy.ndarray) -> None:
        arr = numpy.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            _check_hip(_hip.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes,
                                      _H2D), "hipMemcpy H2D")

    def download(self, arr: numpy.ndarray) -> numpy.ndarray:
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        if arr.nbytes:
            _check_hip(_hip.hipMemcpy(arr.ctypes.data, self.ptr, arr.nbytes,
                                      _D2H), "hipMemcpy D2H")
        return arr

    def __del__(self):
        if _hip is not None and getattr(self, "ptr", None) and self.ptr.value:
            _hip.hipFree(self.ptr)
            self.ptr = c_void_p()


def _check_hip(rc: int, what: str) -> None:
    if rc:
        raise FqeHipError(f"{what}: {_hip.hipGetErrorString(rc).decode()}")


class SpinTables:
    """Strings, de-excitation table, pair table (compressed or ordered
    basis) and the fused apply's forward / slot tables of one spin."""

    def __init__(self, norb: int, nele: int) -> None:
        _, lib = _load()
        self.norb, self.nele = norb, nele
        self.length = comb(norb, nele)
        nexc = nele * (norb - nele + 1)
        self.strings = DeviceBuffer(8 * self.length)
        self.dexc = DeviceBuffer(12 * self.length * nexc)
        _check(lib.fqehip_string_tables(norb, nele, self.length,
                                        self.strings.ptr,
                                        self.dexc.ptr if nexc else None,
                                        None), "fqehip_string_tables")
        self._tabs: Dict[bool, Tuple[DeviceBuffer, ...]] = {}

    def tables(self, compressed: bool):
        """(pair table, forward table, slot table) of the pair basis."""
        if compressed not in self._tabs:
            _, lib = _load()
            n = self.norb
            npair = n * (n + 1) // 2 if compressed else n * n
            size = 4 * npair * self.length
            pt, fwd, slot = (DeviceBuffer(size) for _ in range(3))
            dexc = self.dexc.ptr if self.nele else None
            _check(lib.fqehip_pair_table(n, self.nele, self.length, dexc,
                                         int(compressed), pt.ptr, None),
                   "fqehip_pair_table")
            if npair * self.length:
                _check(lib.fqehip_fused_tables(n, self.nele, self.length,
                                               dexc, int(compressed), pt.ptr,
                                               fwd.ptr, slot.ptr, None),
                       "fqehip_fused_tables")
            self._tabs[compressed] = (pt, fwd, slot)
        return self._tabs[compressed]


_SPIN: Dict[Tuple[int, int], SpinTables] = {}


def spin_tables(norb: int, nele: int) -> SpinTables:
    if (norb, nele) not in _SPIN:
        _SPIN[(norb, nele)] = SpinTables(norb, nele)
    return _SPIN[(norb, nele)]


class PreparedOperator:
    """The GEMM matrix of the dvec form of (h1, h2) on the device:
    V[i,j,k,l] = -h2[i,k,j,l], h1' = h1 - sum_k V[i,k,k,j]; rows are the
    output pairs plus the h1' row, columns the input pairs; the n(n+1)/2
    pair basis when V and h1' are exactly pair-symmetric; complex operators
    in the 2x2 real block form (the layout fqehip_apply documents)."""

    def __init__(self, h1: numpy.ndarray, h2: numpy.ndarray,
                 norb: int) -> None:
        v = -numpy.moveaxis(numpy.asarray(h2, dtype=numpy.complex128), 1, 2)
        h1p = (numpy.asarray(h1, dtype=numpy.complex128)
               - numpy.einsum("ikkj->ij", v))
        self.compressed = bool(
            numpy.array_equal(v, v.transpose(1, 0, 2, 3))
            and numpy.array_equal(v, v.transpose(0, 1, 3, 2))
            and numpy.array_equal(h1p, h1p.T))
        pairs = ([(i, j) for i in range(norb) for j in range(i + 1)]
                 if self.compressed else
                 [(i, j) for i in range(norb) for j in range(norb)])
        rows = [i * norb + j for i, j in pairs]
        full = numpy.vstack([
            v.reshape(norb * norb, norb * norb)[numpy.ix_(rows, rows)],
            h1p.reshape(-1)[rows][None, :]])
        self.complex_op = bool(numpy.any(full.imag))
        if self.complex_op:
            m, k = full.shape
            a = numpy.zeros((2 * m, 2 * k))
            a[0::2, 0::2], a[0::2, 1::2] = full.real, -full.imag
            a[1::2, 0::2], a[1::2, 1::2] = full.imag, full.real
        else:
            a = numpy.ascontiguousarray(full.real)
        self.npair = len(pairs)
        self.pair_symmetric = (not self.complex_op and numpy.array_equal(
            a[:-1, :], a[:-1, :].T))
        self.a = DeviceBuffer.from_array(numpy.ascontiguousarray(a))
        self._hbb: Dict[int, DeviceBuffer] = {}

    def c_struct(self) -> Operator:
        return Operator(int(self.complex_op), self.a.ptr, None)

    def beta_block(self, tb: SpinTables) -> Optional[DeviceBuffer]:
        """FP64 beta-beta block (hcol) of the complex128 split apply."""
        lenb = tb.length
        if not self.pair_symmetric or (lenb + self.npair) * 8 > 160 * 1024:
            return None
        if tb.nele not in self._hbb:
            _, lib = _load()
            ldh = lib.fqehip_beta_block_ld(lenb)
            hcol = DeviceBuffer(8 * lenb * lenb)
            hhi, hlo = DeviceBuffer(2 * ldh * ldh), DeviceBuffer(2 * ldh * ldh)
            fwd = tb.tables(self.compressed)[1]
            _check(lib.fqehip_beta_block(self.npair, lenb, fwd.ptr,
                                         self.a.ptr, hcol.ptr, hhi.ptr,
                                         hlo.ptr, ldh, None),
                   "fqehip_beta_block")
            self._hbb[tb.nele] = hcol
        return self._hbb[tb.nele]


def sector(nalpha: int, nbeta: int, norb: int,
           op: PreparedOperator) -> Tuple[Sector, tuple]:
    """Descriptor of the sector for the pair basis of ``op`` (and the
    buffers it references)."""
    ta, tb = spin_tables(norb, nalpha), spin_tables(norb, nbeta)
    pta, fwa, sla = ta.tables(op.compressed)
    ptb, fwb, _ = tb.tables(op.compressed)
    s = Sector(norb, nalpha, nbeta, op.npair, ta.length, tb.length, pta.ptr,
               ptb.ptr, None, None, None, nalpha * (norb - nalpha), None,
               None, 0, 0, None, None)
    keep = [pta, ptb]
    if not op.complex_op:
        s.fwd_a, s.fwd_b, s.slot_a = fwa.ptr, fwb.ptr, sla.ptr
        keep += [fwa, fwb, sla]
        if op.npair + 1 <= 112 and norb * norb // 4 + 1 <= 64:
            hcol = op.beta_block(tb)
            if hcol is not None:
                s.hbb_f64 = hcol.ptr
                keep.append(hcol)
    return s, tuple(keep)


def apply_spatial12(coeff: numpy.ndarray, h1: numpy.ndarray,
                    h2: numpy.ndarray, nalpha: int, nbeta: int, norb: int,
                    op: Optional[PreparedOperator] = None) -> numpy.ndarray:
    """H coeff for a complex128 [lena, lenb] coefficient block (the
    reference's ``FqeData.apply_inplace((h1, h2))`` result)."""
    _, lib = _load()
    coeff = numpy.ascontiguousarray(coeff, dtype=numpy.complex128)
    if op is None:
        op = PreparedOperator(h1, h2, norb)
    s, keep = sector(nalpha, nbeta, norb, op)
    if coeff.shape != (s.lena, s.lenb):
        raise ValueError(f"coefficients {coeff.shape} for a "
                         f"({s.lena}, {s.lenb}) sector")
    d_c = DeviceBuffer.from_array(coeff)
    d_out = DeviceBuffer(coeff.nbytes)
    ws = DeviceBuffer(lib.fqehip_apply_workspace_bytes(C128, ctypes.byref(s)))
    opc = op.c_struct()
    _check(lib.fqehip_apply(C128, ctypes.byref(s), ctypes.byref(opc), d_c.ptr,
                            1.0, 0.0, 0.0, 0.0, None, d_out.ptr, None, None,
                            None, ws.ptr, None), "fqehip_apply")
    out = d_out.download(numpy.empty_like(coeff))
    del keep
    return out
