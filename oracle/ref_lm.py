"""ctypes driver for the reference's own compiled C (``oracle/_ref``).

TEST INFRASTRUCTURE ONLY (same import rule as ``oracle/fqe_oracle.py``).

``oracle/Makefile`` compiles the reference's hot-path C sources
(``src/fqe/lib/{fqe_data,fci_graph,bitstring,binom,macros,mylapack}.c``)
unmodified into ``oracle/_ref/libfqe_ref.so``.  This module calls that
library the way the reference's bindings do:

* string tables: ``lexicographic_bitstring_generator`` ->
  ``calculate_Z_matrix`` -> ``calculate_string_address`` ->
  ``build_mapping_strings`` (count pass, fill pass) -> ``map_deexc``
  (reference ``src/fqe/fci_graph.py:36-96``, ``:204-244``, ``:301-333``;
  ``src/fqe/lib/fci_graph.py:31-118``; ``src/fqe/bitstring.py:93-118``);
* the default C apply path ``_apply_array_spatial12_lm``
  (reference ``src/fqe/fqe_data.py:685-710``) with the two
  ``lm_apply_array12_*_opt`` kernels (``src/fqe/lib/_fqe_data.pyx:609-765``),
  passing scipy's BLAS ``zaxpy`` as the reference's ``blasfunctions`` struct
  does (``src/fqe/lib/_fqe_data.pyx:25``, ``:683-686``).
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_bool, c_int, c_int64, c_void_p

import numpy
from numpy.ctypeslib import ndpointer

_HERE = os.path.dirname(os.path.abspath(__file__))
_V3 = os.path.join(_HERE, "_ref", "libfqe_ref.so")
_V4 = os.path.join(_HERE, "_ref", "libfqe_ref_v4.so")


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq",
                            "avx512vl"} <= flags
    except OSError:
        pass
    return False


# the reference builds with -march=native: the v4 (AVX-512) build where the
# host has it, else the v3 (AVX2/FMA) one
LIBPATH = _V4 if os.path.exists(_V4) and _has_avx512() else _V3
ISA = "x86-64-v4" if LIBPATH == _V4 else "x86-64-v3"

_lib = None
_blas = None


def available() -> bool:
    return os.path.exists(LIBPATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(
                f"{LIBPATH} is not built (run make -C oracle)")
        _lib = ctypes.CDLL(LIBPATH)
    return _lib


class _BlasFunctions(ctypes.Structure):
    _fields_ = [("zaxpy", c_void_p), ("zscal", c_void_p)]


def _capsule_ptr(name: str) -> int:
    import scipy.linalg.cython_blas as cblas
    cap = cblas.__pyx_capi__[name]
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get_ptr(cap, get_name(cap))


def blasfunctions() -> _BlasFunctions:
    global _blas
    if _blas is None:
        _blas = _BlasFunctions(_capsule_ptr("zaxpy"), _capsule_ptr("zscal"))
    return _blas


# ---------------------------------------------------------------------------
# string tables through the reference C
# ---------------------------------------------------------------------------

def ref_strings(norb: int, nele: int) -> numpy.ndarray:
    from math import comb
    length = comb(norb, nele)
    blist = numpy.zeros(length, dtype=numpy.uint64)
    f = lib().lexicographic_bitstring_generator
    f.argtypes = [ndpointer(dtype=numpy.uint64, flags="C"), c_int, c_int]
    f(blist, norb, nele)
    z = ref_z_matrix(norb, nele)
    out = numpy.zeros(length, dtype=numpy.uint64)
    f = lib().calculate_string_address
    f.argtypes = [ndpointer(dtype=numpy.uint64, flags="C"),
                  ndpointer(dtype=numpy.uint64, flags="C"), c_int,
                  ndpointer(dtype=numpy.int32, flags="C"), c_int]
    f(out, blist, length, z, norb)
    return out


def ref_z_matrix(norb: int, nele: int) -> numpy.ndarray:
    z = numpy.zeros((nele, norb), dtype=numpy.int32)
    if z.size:
        f = lib().calculate_Z_matrix
        f.argtypes = [ndpointer(dtype=numpy.int32, flags="C"), c_int, c_int]
        f(z, norb, nele)
    return z


def ref_maps(strings: numpy.ndarray, norb: int, nele: int):
    z = ref_z_matrix(norb, nele)
    f = lib().build_mapping_strings
    c_ptr_map = POINTER(c_int * 3)
    f.argtypes = [POINTER(c_ptr_map),
                  ndpointer(dtype=numpy.int32, flags="C"),
                  ndpointer(dtype=numpy.int32, flags="C"), c_int,
                  ndpointer(dtype=numpy.uint64, flags="C"), c_int, c_bool,
                  ndpointer(dtype=numpy.int32, flags="C"), c_int]
    lens = numpy.zeros((norb, norb), dtype=numpy.int32)
    exc = numpy.asarray([(i, j) for i in range(norb) for j in range(norb)],
                        dtype=numpy.int32).reshape(-1, 2)
    strings = numpy.ascontiguousarray(strings, dtype=numpy.uint64)
    f(None, lens, exc, len(exc), strings, len(strings), True, z, norb)
    out = {(int(i), int(j)): numpy.zeros((int(ml), 3), dtype=numpy.int32)
           for (i, j), ml in zip(exc, lens.ravel())}
    ptrs = [out[(int(i), int(j))].ctypes.data_as(c_ptr_map) for i, j in exc]
    f((c_ptr_map * len(exc))(*ptrs), lens, exc, len(exc), strings,
      len(strings), False, z, norb)
    return out


def ref_dexc(maps, nstrings: int, norb: int, nele: int) -> numpy.ndarray:
    lk = nele * (norb - nele + 1)
    dexc = numpy.zeros((nstrings, lk, 3), dtype=numpy.int32)
    index = numpy.zeros(nstrings, dtype=numpy.uint32)
    f = lib().map_deexc
    f.argtypes = [ndpointer(dtype=numpy.int32, flags="C"),
                  ndpointer(dtype=numpy.int32, flags="C"), c_int, c_int,
                  ndpointer(dtype=numpy.uint32, flags="C"), c_int]
    for (i, j), values in maps.items():
        values = numpy.ascontiguousarray(values, dtype=numpy.int32)
        f(dexc, values, lk, values.shape[0], index, i * norb + j)
    return dexc


class RefGraph:
    """String tables of one sector, built by the reference's C."""

    def __init__(self, nalpha: int, nbeta: int, norb: int) -> None:
        self.norb, self.nalpha, self.nbeta = norb, nalpha, nbeta
        self.astr = ref_strings(norb, nalpha)
        self.bstr = ref_strings(norb, nbeta)
        self.lena, self.lenb = self.astr.size, self.bstr.size
        self.alpha_map = ref_maps(self.astr, norb, nalpha)
        self.beta_map = ref_maps(self.bstr, norb, nbeta)
        self.dexca = ref_dexc(self.alpha_map, self.lena, norb, nalpha)
        self.dexcb = ref_dexc(self.beta_map, self.lenb, norb, nbeta)


# ---------------------------------------------------------------------------
# the reference's default (C, "lm") spatial 1+2-body apply
# ---------------------------------------------------------------------------

def _same_spin(coeff, h1e, h2e, dexc, len1, len2, norb, alpha, out):
    f = lib().lm_apply_array12_same_spin_opt
    f.argtypes = [ndpointer(dtype=numpy.complex128, flags="C"),
                  ndpointer(dtype=numpy.complex128, flags="C"),
                  ndpointer(dtype=numpy.int32, flags="C"), c_int, c_int,
                  c_int, ndpointer(dtype=numpy.complex128, flags="C"),
                  ndpointer(dtype=numpy.complex128, flags="C"), c_int, c_bool,
                  c_void_p]
    f(coeff, out, dexc, len1, len2, dexc.shape[1], h1e, h2e, norb, alpha,
      ctypes.addressof(blasfunctions()))
    return out


def _diff_spin(coeff, h2e, adexc, bdexc, lena, lenb, norb, out):
    f = lib().lm_apply_array12_diff_spin_opt
    f.argtypes = [ndpointer(dtype=numpy.complex128, flags="C"),
                  ndpointer(dtype=numpy.complex128, flags="C"),
                  ndpointer(dtype=numpy.int32, flags="C"),
                  ndpointer(dtype=numpy.int32, flags="C"), c_int, c_int,
                  c_int, c_int,
                  ndpointer(dtype=numpy.complex128, flags="C"), c_int,
                  c_void_p]
    f(coeff, out, adexc, bdexc, lena, lenb, adexc.shape[1], bdexc.shape[1],
      h2e, norb, ctypes.addressof(blasfunctions()))
    return out


def apply_12_lm(coeff, h1e, h2e, graph) -> numpy.ndarray:
    """The reference's ``_apply_array_spatial12_lm`` (``src/fqe/fqe_data.py:
    685-710``) driving its compiled kernels; ``graph`` supplies dexca/dexcb
    in the reference layout."""
    norb = graph.norb
    lena, lenb = graph.lena, graph.lenb
    dt = numpy.complex128
    h1e = numpy.array(h1e, dtype=dt, copy=True)
    h2e = numpy.moveaxis(numpy.array(h2e, dtype=dt, copy=True), 1, 2) * (-1.0)
    h1e -= numpy.einsum("ikkj->ij", h2e)
    h1e = numpy.ascontiguousarray(h1e)
    h2e = numpy.ascontiguousarray(h2e)
    coeff = numpy.ascontiguousarray(coeff, dtype=dt)
    dexca = numpy.ascontiguousarray(graph.dexca, dtype=numpy.int32)
    dexcb = numpy.ascontiguousarray(graph.dexcb, dtype=numpy.int32)
    out = numpy.zeros((lena, lenb), dtype=dt)
    _same_spin(coeff, h1e, h2e, dexca, lena, lenb, norb, True, out)
    coeff_t = numpy.ascontiguousarray(coeff.T)
    out_t = numpy.zeros((lenb, lena), dtype=dt)
    _same_spin(coeff_t, h1e, h2e, dexcb, lenb, lena, norb, True, out_t)
    out += out_t.T
    w = numpy.ascontiguousarray(h2e + numpy.einsum("ijkl->klij", h2e))
    _diff_spin(coeff, w, dexca, dexcb, lena, lenb, norb, out)
    return out
