"""Batched sectors of uneven size: one apply and a fixed-length Taylor
series per sector in one HIP graph launch (``fqehip_batch_*``), with the
kernel variant of every sector shape chosen by an autotuner.

The reference evolves sectors one by one from Python, each with its own
Hamiltonian object (``src/fqe/wavefunction.py:509-611``); its Taylor loop
runs until the term norm drops below ``accuracy`` (``:556-568``).  A batch
here runs exactly ``nterms`` terms for every sector, which is what a
batched propagator with a fixed step needs; ``run_sequential`` gives the
same computation sector by sector without a graph (the FP64 comparison
path).

Kernel variants (``Variant``): the split apply (pair-symmetric operator:
beta-beta GEMM + alpha row kernel; in complex128 with FP64 or Ozaki-digit
contractions, ``contraction``), the row kernel with beta scatter at 4 or 8
waves per alpha row, and the unfused gather / GEMM / scatter kernels.  ``autotune`` times one apply of each candidate per
distinct (norb, nalpha, nbeta) shape and keeps the fastest.
"""

from __future__ import annotations

import ctypes
import json
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy
import torch

from . import _lib, engine, settings
from .fci_graph import FciGraph

VARIANTS = ("split", "rows4", "rows8", "unfused")


@dataclass
class SectorSpec:
    norb: int
    nalpha: int
    nbeta: int
    h1: numpy.ndarray
    h2: numpy.ndarray
    coeff: numpy.ndarray      # [lena, lenb] complex


def _struct_for(graph: FciGraph, op: engine.PreparedOperator, prec: int,
                variant: str, contraction: Optional[str] = None):
    saved = (settings.fused, settings.split, settings.split_fp64,
             settings.contraction)
    try:
        settings.fused = variant != "unfused"
        settings.split = variant == "split"
        settings.split_fp64 = variant == "split"
        if contraction is not None:
            settings.contraction = contraction
        s, keep = engine.sector_struct(graph, op, prec)
    finally:
        (settings.fused, settings.split, settings.split_fp64,
         settings.contraction) = saved
    if variant == "rows8":
        s.tile = 8
    elif variant == "rows4":
        s.tile = 4
    return s, keep


def variants_for(op: engine.PreparedOperator, prec: int) -> List[str]:
    """Variants that apply to an operator at a precision."""
    if op.complex_op:
        return ["unfused"]
    out = ["rows4", "unfused"]
    if prec == _lib.C64:
        out = ["rows4", "rows8", "unfused"]
        if op.pair_symmetric:
            out.insert(0, "split")
    elif op.pair_symmetric:
        out.insert(1, "split")
    return out


class _Prepared:
    """Device state of one sector of a batch."""

    def __init__(self, spec: SectorSpec, dtype: torch.dtype,
                 variant: Optional[str],
                 contraction: Optional[str] = None) -> None:
        dev = _lib.device()
        self.spec = spec
        self.graph = FciGraph(spec.nalpha, spec.nbeta, spec.norb)
        self.prec = engine.precision_of(dtype)
        self.op = engine.prepare_operator((spec.h1, spec.h2), spec.norb, dev)
        self.variant = variant or variants_for(self.op, self.prec)[0]
        self.s, self.keep = _struct_for(self.graph, self.op, self.prec,
                                        self.variant, contraction)
        self.opc = self.op.c_struct()
        shape = (self.graph.lena(), self.graph.lenb())
        self.c0 = torch.from_numpy(numpy.asarray(spec.coeff).reshape(shape)
                                   ).to(dev).to(dtype)
        self.term = torch.empty_like(self.c0)
        self.apply_out = torch.empty_like(self.c0)
        self.result = torch.empty(shape, dtype=torch.complex128, device=dev)
        self.ws = engine.workspace(self.s, self.prec, dev)

    def item(self) -> _lib.BatchItem:
        return _lib.BatchItem(self.s, self.opc, _lib.ptr(self.c0),
                              _lib.ptr(self.apply_out), _lib.ptr(self.term),
                              _lib.ptr(self.result), _lib.ptr(self.ws))

    def n(self) -> int:
        return self.graph.lena() * self.graph.lenb()


def autotune(specs: Sequence[SectorSpec], dtype: torch.dtype = torch.complex64,
             reps: int = 3, cache: Optional[Dict[str, str]] = None,
             contraction: Optional[str] = None) -> Dict[str, str]:
    """Fastest variant per (norb, nalpha, nbeta) shape: one apply of the
    first sector of each shape per candidate, median of ``reps`` timed
    launches after one warm-up.  ``cache`` entries are kept as they are.
    ``contraction`` (complex128): the arithmetic of the split variant,
    "fp64" or "ozaki" (default: ``settings.contraction``)."""
    table = dict(cache or {})
    seen = {}
    for sp in specs:
        key = f"{sp.norb},{sp.nalpha},{sp.nbeta}"
        if key not in table and key not in seen:
            seen[key] = sp
    for key, sp in seen.items():
        best, best_t = None, None
        base = _Prepared(sp, dtype, None, contraction)
        for var in variants_for(base.op, base.prec):
            p = (base if var == base.variant else
                 _Prepared(sp, dtype, var, contraction))
            times = []
            for r in range(reps + 1):
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record()
                _lib.call("fqehip_apply", p.prec, ctypes.byref(p.s),
                          ctypes.byref(p.opc), _lib.ptr(p.c0), 1.0, 0.0, 0.0,
                          0.0, None, _lib.ptr(p.apply_out), None, None, None,
                          _lib.ptr(p.ws), _lib.stream_ptr())
                ev[1].record()
                torch.cuda.synchronize()
                if r:
                    times.append(ev[0].elapsed_time(ev[1]))
            t = float(numpy.median(times))
            if best_t is None or t < best_t:
                best, best_t = var, t
        table[key] = best
    return table


class BatchedTaylor:
    """One apply + ``nterms``-term Taylor series of exp(-i t H) for every
    sector, recorded into one HIP graph.  With ``group``, the sectors of
    the fused row variants without the split (complex64 rows4/rows8,
    complex128 FP64 rows) run as one group: one row launch and one
    reduction per step for all of them."""

    def __init__(self, specs: Sequence[SectorSpec], time_step: float,
                 nterms: int = 10, dtype: torch.dtype = torch.complex64,
                 tiles: Optional[Dict[str, str]] = None,
                 nstreams: int = 8, group: bool = True,
                 contraction: Optional[str] = None) -> None:
        self.time = float(time_step)
        self.nterms = int(nterms)
        self.dtype = dtype
        self.tiles = tiles or {}
        self.sectors = [
            _Prepared(sp, dtype,
                      self.tiles.get(f"{sp.norb},{sp.nalpha},{sp.nbeta}"),
                      contraction)
            for sp in specs]
        self.prec = engine.precision_of(dtype)
        self.nstreams = nstreams
        self.group = bool(group)
        self._batch = None

    def build(self) -> float:
        """Record the graph; returns the seconds taken."""
        t0 = time.perf_counter()
        items = (_lib.BatchItem * len(self.sectors))(
            *[p.item() for p in self.sectors])
        handle = ctypes.c_void_p()
        torch.cuda.synchronize()
        _lib.call("fqehip_batch_create", self.prec, len(self.sectors), items,
                  self.time, self.nterms, self.nstreams, int(self.group),
                  _lib.stream_ptr(), ctypes.byref(handle))
        self._batch = handle
        return time.perf_counter() - t0

    def nodes(self) -> int:
        return int(_lib.call("fqehip_batch_nodes", self._batch))

    def grouped(self) -> int:
        """Number of sectors that run as one group."""
        return int(_lib.call("fqehip_batch_grouped", self._batch))

    def launch(self) -> None:
        _lib.call("fqehip_batch_launch", self._batch, _lib.stream_ptr())

    def run_sequential(self) -> None:
        """The same work eagerly, one sector after another on one stream."""
        st = _lib.stream_ptr()
        for p in self.sectors:
            n = p.n()
            p.term.copy_(p.c0)
            _lib.call("fqehip_convert_to_c128", p.prec, n, _lib.ptr(p.c0), 1.0,
                      0.0, _lib.ptr(p.result), st)
            _lib.call("fqehip_apply", p.prec, ctypes.byref(p.s),
                      ctypes.byref(p.opc), _lib.ptr(p.c0), 1.0, 0.0, 0.0, 0.0,
                      None, _lib.ptr(p.apply_out), None, None, None,
                      _lib.ptr(p.ws), st)
            for k in range(1, self.nterms + 1):
                _lib.call("fqehip_apply", p.prec, ctypes.byref(p.s),
                          ctypes.byref(p.opc), _lib.ptr(p.term), 0.0,
                          -self.time / k, 0.0, 0.0, None, _lib.ptr(p.term),
                          _lib.ptr(p.result), None, None, _lib.ptr(p.ws), st)

    def results(self) -> List[Tuple[numpy.ndarray, numpy.ndarray]]:
        torch.cuda.synchronize()
        return [(p.apply_out.to(torch.complex128).cpu().numpy(),
                 p.result.cpu().numpy()) for p in self.sectors]

    def determinants(self) -> int:
        return sum(p.n() for p in self.sectors)

    def close(self) -> None:
        if self._batch is not None:
            _lib.call("fqehip_batch_destroy", self._batch)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_specs(count: int, seed: int, norb_range=(6, 12),
                 nfactors: Optional[int] = None) -> List[SectorSpec]:
    """Sectors with norb uniform in norb_range, nalpha and nbeta uniform in
    [1, norb-1], each with its own Hamiltonian (real symmetric h1, ERIs from
    Cholesky-like factors, converted to FQE's h2) and complex coefficients
    normalised per sector; draw order per sector: norb, nalpha, nbeta,
    then h1, factors, coefficients."""
    from math import comb
    rng = numpy.random.default_rng(seed)
    out = []
    for _ in range(count):
        norb = int(rng.integers(norb_range[0], norb_range[1] + 1))
        na = int(rng.integers(1, norb))
        nb = int(rng.integers(1, norb))
        k = nfactors or norb
        h1 = rng.normal(0.0, 0.1, size=(norb, norb))
        h1 = 0.5 * (h1 + h1.T)
        numpy.fill_diagonal(h1, -2.0 + rng.normal(0.0, 0.5, size=norb))
        fac = rng.normal(0.0, 0.1, size=(k, norb, norb))
        fac = 0.5 * (fac + fac.transpose(0, 2, 1))
        eri = numpy.einsum("kij,kls->ijls", fac, fac)      # (ij|ls)
        h2 = -0.5 * numpy.einsum("ikjl->ijkl", eri)        # -1/2 (ik|jl)
        lena, lenb = comb(norb, na), comb(norb, nb)
        c = rng.normal(size=(lena, lenb)) + 1j * rng.normal(size=(lena, lenb))
        c /= numpy.linalg.norm(c)
        out.append(SectorSpec(norb, na, nb, h1, h2, c))
    return out


def save_tiles(table: Dict[str, str], path: str) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load_tiles(path: str) -> Dict[str, str]:
    with open(path) as f:
        return json.load(f)
