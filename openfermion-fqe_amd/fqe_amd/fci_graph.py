"""FciGraph: determinant addressing and excitation tables, built on the GPU.

Mirrors the reference's ``fqe.fci_graph.FciGraph`` (``src/fqe/fci_graph.py:
99-419``): same constructor, accessors and table layouts (``_astr``,
``_aind``, ``_alpha_map``, ``_dexca`` ...).  The tables are produced by the
HIP kernels of ``fqehip_string_tables`` / ``fqehip_string_maps`` /
``fqehip_pair_table`` and stay resident in device memory; the host views
the reference exposes are downloaded on first access.
"""

from __future__ import annotations

import copy
from math import comb
from typing import Dict, List, Tuple

import numpy
import torch

from . import _lib

Spinmap = Dict[Tuple[int, int], numpy.ndarray]


class SpinTables:
    """Device tables of one spin (norb, nele), shared by every sector and
    every FciGraph on the same device."""

    def __init__(self, norb: int, nele: int, device: torch.device) -> None:
        self.norb = norb
        self.nele = nele
        self.length = comb(norb, nele)
        self.nexc = nele * (norb - nele + 1)
        self.device = device
        self.strings = torch.empty(self.length, dtype=torch.int64,
                                   device=device)
        self.dexc = torch.empty((self.length, self.nexc, 3),
                                dtype=torch.int32, device=device)
        _lib.call("fqehip_string_tables", norb, nele, self.length,
                  _lib.ptr(self.strings),
                  _lib.ptr(self.dexc) if self.dexc.numel() else None,
                  _lib.stream_ptr())
        self._ptab = {}
        self._maps = None
        self._map_offsets = None

    def pair_table(self, compressed: bool) -> torch.Tensor:
        """int32 [npair][length]: sign*(source+1) or 0 (see fqehip.h)."""
        key = bool(compressed)
        if key not in self._ptab:
            npair = (self.norb * (self.norb + 1) // 2 if compressed else
                     self.norb * self.norb)
            tab = torch.empty((npair, self.length), dtype=torch.int32,
                              device=self.device)
            _lib.call("fqehip_pair_table", self.norb, self.nele, self.length,
                      _lib.ptr(self.dexc) if self.dexc.numel() else None,
                      int(key), _lib.ptr(tab), _lib.stream_ptr())
            self._ptab[key] = tab
        return self._ptab[key]

    def spin_pair_table(self, first: bool) -> torch.Tensor:
        """int32 [2 norb^2][length]: the full-basis pair table of this spin
        in the first (alpha, ``first``) or second (beta) half and zeros in
        the other, so that pair p < norb^2 excites alpha strings only and
        p >= norb^2 beta strings only (spin-orbital operators on one
        sector)."""
        key = ("spin", bool(first))
        if key not in self._ptab:
            pt = self.pair_table(False)
            zero = torch.zeros_like(pt)
            self._ptab[key] = torch.cat([pt, zero] if first else [zero, pt])
        return self._ptab[key]

    def fused_tables(self, compressed: bool):
        """(forward table, slot table) of the fused apply, int32
        [npair][length] each (see fqehip_fused_tables)."""
        key = ("fused", bool(compressed))
        if key not in self._ptab:
            pt = self.pair_table(compressed)
            fwd = torch.empty_like(pt)
            slot = torch.empty_like(pt)
            if pt.numel():
                _lib.call("fqehip_fused_tables", self.norb, self.nele,
                          self.length,
                          _lib.ptr(self.dexc) if self.dexc.numel() else None,
                          int(bool(compressed)), _lib.ptr(pt), _lib.ptr(fwd),
                          _lib.ptr(slot), _lib.stream_ptr())
            self._ptab[key] = (fwd, slot)
        return self._ptab[key]

    def map_lengths(self) -> numpy.ndarray:
        n, k = self.norb, self.nele
        off = comb(n - 2, k - 1) if n >= 2 and k >= 1 else 0
        diag = comb(n - 1, k - 1) if n >= 1 and k >= 1 else 0
        lens = numpy.full((n, n), off, dtype=numpy.int64)
        numpy.fill_diagonal(lens, diag)
        return lens

    def maps(self) -> torch.Tensor:
        if self._maps is None:
            lens = self.map_lengths().reshape(-1)
            offsets = numpy.zeros(lens.size + 1, dtype=numpy.int64)
            numpy.cumsum(lens, out=offsets[1:])
            self._map_offsets = offsets
            total = int(offsets[-1])
            self._maps = torch.empty((max(total, 1), 3), dtype=torch.int32,
                                     device=self.device)
            d_off = torch.from_numpy(offsets).to(self.device)
            _lib.call("fqehip_string_maps", self.norb, self.nele, self.length,
                      _lib.ptr(self.strings), _lib.ptr(d_off),
                      _lib.ptr(self._maps), _lib.stream_ptr())
            torch.cuda.current_stream().synchronize()
        return self._maps


_TABLE_CACHE: Dict[Tuple[int, int, int], SpinTables] = {}


def spin_tables(norb: int, nele: int) -> SpinTables:
    dev = _lib.device()
    key = (norb, nele, dev.index)
    if key not in _TABLE_CACHE:
        _TABLE_CACHE[key] = SpinTables(norb, nele, dev)
    return _TABLE_CACHE[key]


def _z_matrix(norb: int, nele: int) -> numpy.ndarray:
    z = numpy.zeros((nele, norb), dtype=numpy.int32)
    for k in range(1, nele):
        for ll in range(k, norb - nele + k + 1):
            z[k - 1, ll - 1] = sum(
                comb(m, nele - k) - (comb(m - 1, nele - k - 1)
                                     if nele - k - 1 >= 0 else 0)
                for m in range(norb - ll + 1, norb - k + 1))
    for ll in range(nele, norb + 1):
        if nele:
            z[nele - 1, ll - 1] = ll - nele
    return z


class FciGraph:
    """Addressing of the determinants of one (nalpha, nbeta, norb) sector
    (reference ``src/fqe/fci_graph.py:99-154``)."""

    def __init__(self, nalpha: int, nbeta: int, norb: int) -> None:
        if norb < 0:
            raise ValueError(f'norb needs to be >= 0, passed value is {norb}')
        if nalpha < 0:
            raise ValueError(
                f'nalpha needs to be >= 0, passed value is {nalpha}')
        if nbeta < 0:
            raise ValueError(f'nbeta needs to be >= 0, passed value is {nbeta}')
        if nalpha > norb:
            raise ValueError(
                f'nalpha needs to be <= norb, passed value is {nalpha}')
        if nbeta > norb:
            raise ValueError(
                f'nbeta needs to be <= norb, passed value is {nbeta}')
        self._norb = norb
        self._nalpha = nalpha
        self._nbeta = nbeta
        self._lena = comb(norb, nalpha)
        self._lenb = comb(norb, nbeta)
        self._ta = spin_tables(norb, nalpha)
        self._tb = spin_tables(norb, nbeta)
        self._host: Dict[str, object] = {}
        self._fci_map: Dict[Tuple[int, ...], Tuple[Spinmap, Spinmap]] = {}

    # -- device tables ------------------------------------------------------
    def tables(self, alpha: bool = True) -> SpinTables:
        return self._ta if alpha else self._tb

    # -- host views (downloaded once) --------------------------------------
    def _strings_host(self, alpha: bool) -> numpy.ndarray:
        key = "astr" if alpha else "bstr"
        if key not in self._host:
            t = self.tables(alpha)
            self._host[key] = t.strings.cpu().numpy().astype(numpy.uint64)
        return self._host[key]

    def _index_host(self, alpha: bool) -> Dict[int, int]:
        key = "aind" if alpha else "bind"
        if key not in self._host:
            self._host[key] = {
                int(s): a for a, s in enumerate(self._strings_host(alpha))}
        return self._host[key]

    def _dexc_host(self, alpha: bool) -> numpy.ndarray:
        key = "dexca" if alpha else "dexcb"
        if key not in self._host:
            self._host[key] = self.tables(alpha).dexc.cpu().numpy()
        return self._host[key]

    def _maps_host(self, alpha: bool) -> Spinmap:
        key = "amap" if alpha else "bmap"
        if key not in self._host:
            t = self.tables(alpha)
            rows = t.maps().cpu().numpy()
            off = t._map_offsets
            n = self._norb
            self._host[key] = {
                (i, j): rows[off[i * n + j]:off[i * n + j + 1]].copy()
                for i in range(n) for j in range(n)}
        return self._host[key]

    _astr = property(lambda self: self._strings_host(True))
    _bstr = property(lambda self: self._strings_host(False))
    _aind = property(lambda self: self._index_host(True))
    _bind = property(lambda self: self._index_host(False))
    _dexca = property(lambda self: self._dexc_host(True))
    _dexcb = property(lambda self: self._dexc_host(False))
    _alpha_map = property(lambda self: self._maps_host(True))
    _beta_map = property(lambda self: self._maps_host(False))

    # -- reference accessors ------------------------------------------------
    def alpha_beta_transpose(self) -> 'FciGraph':
        out = copy.copy(self)
        out._nalpha, out._nbeta = self._nbeta, self._nalpha
        out._lena, out._lenb = self._lenb, self._lena
        out._ta, out._tb = self._tb, self._ta
        out._host = {}
        out._fci_map = dict(self._fci_map)
        return out

    def insert_mapping(self, dna: int, dnb: int,
                       mapping_pair: Tuple[Spinmap, Spinmap]) -> None:
        self._fci_map[(dna, dnb)] = mapping_pair

    def find_mapping(self, dna: int, dnb: int) -> Tuple[Spinmap, Spinmap]:
        return self._fci_map[(dna, dnb)]

    def alpha_map(self, iorb: int, jorb: int) -> numpy.ndarray:
        return self._alpha_map[(iorb, jorb)]

    def beta_map(self, iorb: int, jorb: int) -> numpy.ndarray:
        return self._beta_map[(iorb, jorb)]

    def lena(self) -> int:
        return self._lena

    def lenb(self) -> int:
        return self._lenb

    def nalpha(self) -> int:
        return self._nalpha

    def nbeta(self) -> int:
        return self._nbeta

    def norb(self) -> int:
        return self._norb

    def string_alpha(self, address: int) -> int:
        return self._astr[address]

    def string_beta(self, address: int) -> int:
        return self._bstr[address]

    def string_alpha_all(self) -> numpy.ndarray:
        return self._astr

    def string_beta_all(self) -> numpy.ndarray:
        return self._bstr

    def index_alpha(self, bit_string: int) -> int:
        return self._aind[int(bit_string)]

    def index_beta(self, bit_string: int) -> int:
        return self._bind[int(bit_string)]

    def index_alpha_all(self) -> Dict[int, int]:
        return self._aind

    def index_beta_all(self) -> Dict[int, int]:
        return self._bind

    def _build_string_address(self, nele: int, norb: int,
                              occupation: List[int]) -> int:
        z = _z_matrix(norb, nele)
        return int(sum(z[i, occupation[i]] for i in range(nele)))
