"""FqeDataSet: every sector of one particle number, for spin-orbital
operators (reference ``src/fqe/fqe_data_set.py:35-209``, ``:502-559``).

Mirrors the reference class for the 1- and 1+2-body spin-orbital apply and
the 1- and 2-RDMs: same constructor ``FqeDataSet(nele, norb, data)`` with
``data`` a dict {(nalpha, nbeta): FqeData}, ``apply`` / ``apply_inplace``
with (h1,) or (h1, h2) of shape (2 norb,) * 2 / (2 norb,) * 4,
``rdm1`` / ``rdm12``, ``sectors``, ``empty_copy``, ``ax_plus_y``, ``scale``
and ``fill``.  Spin-orbital index p < norb is alpha orbital p, p >= norb
beta orbital p - norb.

On the device the sectors are concatenated, ascending nbeta, into one
complex128 vector; ``fqehip_set_apply`` / ``fqehip_set_dvec`` gather over
it and contract on the FP64 matrix cores (no atomics, the same bits on
every run).  Like the reference's
``_calculate_dvec_with_coeff``, which reads the neighbouring sectors of
every sector, the set must hold every sector of its particle number;
3- and 4-body tuples raise NotImplementedError.
"""

from __future__ import annotations

import copy
from typing import Dict, Optional, Tuple

import numpy
import torch

from . import _lib
from . import engine
from .fqe_data import FqeData

Key = Tuple[int, int]


def _set_keys(nele: int, norb: int):
    return [(nele - nb, nb) for nb in range(nele + 1)
            if nele - nb <= norb and nb <= norb]


def _check_fits(nbytes: int, what: str) -> None:
    """MemoryError, before any allocation, when a workspace of the
    spin-orbital path exceeds the device memory that is free (torch's
    cached blocks included): the path keeps the (2 norb)^2 pair planes of
    every sector in HBM at once and has no chunking over pairs."""
    dev = _lib.device()
    free, _ = torch.cuda.mem_get_info(dev)
    free += torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)
    if nbytes > free:
        raise MemoryError(
            f"{what} needs {nbytes / 2**30:.1f} GiB of device memory and "
            f"{free / 2**30:.1f} GiB are free: the spin-orbital path holds "
            "all (2 norb)^2 pair planes of every sector at once (no "
            "chunking over pairs)")


class FqeDataSet:
    """reference ``src/fqe/fqe_data_set.py:35-60``"""

    def __init__(self, nele: int, norb: int,
                 data: Dict[Key, FqeData]) -> None:
        self._nele = nele
        self._norb = norb
        self._data = data

    def empty_copy(self) -> 'FqeDataSet':
        """reference ``src/fqe/fqe_data_set.py:62-74``"""
        data = {k: v.empty_copy() for k, v in self._data.items()}
        return FqeDataSet(self._nele, self._norb, data)

    def sectors(self) -> Dict[Key, FqeData]:
        return self._data

    def ax_plus_y(self, factor: complex, other: 'FqeDataSet') -> None:
        """reference ``src/fqe/fqe_data_set.py:83-96``"""
        for key in self._data:
            self._data[key].ax_plus_y(factor, other._data[key])

    def scale(self, factor: complex) -> None:
        for sector in self._data.values():
            sector.scale(factor)

    def fill(self, value: complex) -> None:
        for sector in self._data.values():
            sector.fill(value)

    # -- device layout -------------------------------------------------------
    def _keys(self):
        keys = _set_keys(self._nele, self._norb)
        if set(keys) != set(self._data.keys()):
            raise ValueError(
                f"FqeDataSet({self._nele}, {self._norb}) needs the sectors "
                f"{keys}; it holds {sorted(self._data.keys())}")
        return keys

    def _vector(self) -> torch.Tensor:
        keys = self._keys()
        parts = [self._data[k].device_coeff().reshape(-1).to(torch.complex128)
                 for k in keys]
        vec = torch.cat(parts).contiguous()
        assert vec.numel() == _lib.call("fqehip_set_length", self._norb,
                                        self._nele)
        return vec

    def _split_into(self, vec: torch.Tensor, out: 'FqeDataSet') -> None:
        off = 0
        for k in self._keys():
            sec = out._data[k]
            n = sec.lena() * sec.lenb()
            sec.set_device_coeff(vec[off:off + n].reshape(sec.lena(),
                                                          sec.lenb()))
            off += n

    def _dvec(self, vec: torch.Tensor) -> torch.Tensor:
        """[(2 norb)^2, L] complex128 D[pq] = E_pq C."""
        m = 2 * self._norb
        # planar planes, then their complex128 copy
        _check_fits(2 * m * m * vec.numel() * 16, "the spin-orbital dvec")
        d = torch.empty((m * m, 2, vec.numel()), dtype=torch.float64,
                        device=vec.device)
        _lib.call("fqehip_set_dvec", self._norb, self._nele, _lib.ptr(vec),
                  _lib.ptr(d), _lib.stream_ptr())
        return torch.complex(d[:, 0], d[:, 1]).contiguous()

    # -- apply -----------------------------------------------------------------
    def apply(self, array: Tuple[numpy.ndarray, ...]) -> 'FqeDataSet':
        """reference ``src/fqe/fqe_data_set.py:130-151``"""
        out = copy.deepcopy(self)
        out.apply_inplace(array)
        return out

    def apply_inplace(self, array: Tuple[numpy.ndarray, ...]) -> None:
        """reference ``src/fqe/fqe_data_set.py:118-128``"""
        m = 2 * self._norb
        if len(array) not in (1, 2):
            raise NotImplementedError(
                "the spin-orbital apply takes 1- and 2-body tensors")
        h1 = numpy.asarray(array[0], dtype=numpy.complex128)
        if h1.shape != (m, m):
            raise ValueError(f"h1 must have shape {(m, m)}")
        vec = self._vector()
        dev = vec.device
        out = torch.empty_like(vec)
        if len(array) == 1:
            th1 = torch.from_numpy(numpy.ascontiguousarray(h1)).to(dev)
            _lib.call("fqehip_set_apply", self._norb, self._nele,
                      _lib.ptr(th1), 0, _lib.ptr(vec), _lib.ptr(out), 0,
                      _lib.stream_ptr())
        else:
            h2 = numpy.asarray(array[1], dtype=numpy.complex128)
            if h2.shape != (m,) * 4:
                raise ValueError(f"h2 must have shape {(m,) * 4}")
            # dvec form (reference src/fqe/fqe_data_set.py:194-197)
            v = -numpy.moveaxis(h2, 1, 2)
            h1p = h1 - numpy.einsum("ikkj->ij", v)
            th1 = torch.from_numpy(numpy.ascontiguousarray(h1p)).to(dev)
            tv = torch.from_numpy(
                numpy.ascontiguousarray(v.reshape(m * m, m * m))).to(dev)
            nbytes = int(_lib.call("fqehip_set_workspace_bytes", self._norb,
                                   self._nele))
            _check_fits(nbytes, "the spin-orbital two-body apply")
            ws = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=dev)
            _lib.call("fqehip_set_apply", self._norb, self._nele,
                      _lib.ptr(th1), _lib.ptr(tv), _lib.ptr(vec),
                      _lib.ptr(out), _lib.ptr(ws), _lib.stream_ptr())
        self._split_into(out, self)

    # -- reduced density matrices -----------------------------------------
    def rdm1(self, bra: Optional['FqeDataSet'] = None) -> Tuple[numpy.ndarray]:
        """rdm1[p, q] = <bra|a+_p a_q|ket> summed over the sectors
        (reference ``src/fqe/fqe_data_set.py:502-521``)."""
        m = 2 * self._norb
        vec = self._vector()
        bvec = vec if bra is None else bra._vector()
        d2 = self._dvec(bvec)
        t = engine.block_dots(d2, vec)
        return (t.reshape(m, m).T.copy(),)

    def rdm12(self, bra: Optional['FqeDataSet'] = None
              ) -> Tuple[numpy.ndarray, numpy.ndarray]:
        """reference ``src/fqe/fqe_data_set.py:523-559``:
        rdm2[j, k, i, l] = -<E_ij bra|E_kl ket> + delta_ki rdm1[j, l]; the
        (2 norb)^2 x (2 norb)^2 dvec overlap is one FP64 Gram product."""
        m = 2 * self._norb
        vec = self._vector()
        d = self._dvec(vec)
        d2 = d if bra is None else self._dvec(bra._vector())
        out1 = engine.block_dots(d2, vec).reshape(m, m).T
        g = engine.gram(d2, d)
        out2 = -g.reshape((m,) * 4).transpose(1, 2, 0, 3).copy()
        for i in range(m):
            out2[:, i, i, :] += out1
        return out1.copy(), out2
