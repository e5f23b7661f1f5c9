"""FqeData: the coefficients of one (nalpha, nbeta) sector, in device memory.

Mirrors ``fqe.fqe_data.FqeData`` for the apply path (reference
``src/fqe/fqe_data.py:60-475``, ``:2149-2160``, ``:2620-2811``): same
constructor, ``apply`` / ``apply_inplace`` for spatial 1- and 1+2-body
tuples, ``calculate_dvec_spatial``, ``ax_plus_y``, ``norm``, ``scale``,
``set_wfn`` and the accessors.

Residency.  The coefficients live in HBM as a torch tensor (complex128, or
complex64 for the mixed-precision engine).  ``coeff`` keeps the
reference's numpy semantics: reading it moves the sector to a host numpy
array (so in-place numpy edits take effect), and the next GPU operation
moves it back.  ``device_coeff()`` gives the device tensor without a copy.
"""

from __future__ import annotations

import copy
from typing import Optional, Tuple

import numpy
import torch

from . import _lib
from . import engine
from .fci_graph import FciGraph
from .util import validate_config, rand_wfn

_TORCH_DTYPE = {numpy.dtype(numpy.complex128): torch.complex128,
                numpy.dtype(numpy.complex64): torch.complex64}


def download(tensor: torch.Tensor) -> numpy.ndarray:
    """Device -> host copy into page-locked memory from torch's caching
    host allocator: a direct DMA (3 ms for 188 MB) instead of the staged
    pageable copy into fresh memory (about 43 ms).  The array owns the
    pinned block until it is freed."""
    if tensor.device.type == "cpu":
        return tensor.numpy()
    host = torch.empty(tensor.shape, dtype=tensor.dtype, pin_memory=True)
    host.copy_(tensor)
    return host.numpy()


class FqeData:
    """One sector of a wavefunction (reference ``src/fqe/fqe_data.py:60``)."""

    def __init__(self, nalpha: int, nbeta: int, norb: int,
                 fcigraph: Optional[FciGraph] = None,
                 dtype=numpy.complex128) -> None:
        validate_config(nalpha, nbeta, norb)
        if fcigraph is not None and (nalpha != fcigraph.nalpha() or
                                     nbeta != fcigraph.nbeta() or
                                     norb != fcigraph.norb()):
            raise ValueError("FciGraph does not match other parameters")
        self._core = fcigraph if fcigraph is not None else FciGraph(
            nalpha, nbeta, norb)
        self._dtype = numpy.dtype(dtype)
        if self._dtype not in _TORCH_DTYPE:
            raise TypeError(f"unsupported dtype {dtype}: the engine holds "
                            "complex128 or complex64 coefficients")
        self._nele = nalpha + nbeta
        self._m_s = nalpha - nbeta
        self._host: Optional[numpy.ndarray] = None
        self._dev: Optional[torch.Tensor] = torch.zeros(
            (self.lena(), self.lenb()), dtype=_TORCH_DTYPE[self._dtype],
            device=_lib.device())

    # -- residency ----------------------------------------------------------
    @property
    def coeff(self) -> numpy.ndarray:
        if self._host is None:
            self._host = download(self._dev)
            self._dev = None
        return self._host

    @coeff.setter
    def coeff(self, value) -> None:
        if isinstance(value, torch.Tensor):
            self.set_device_coeff(value)
            return
        self._host = numpy.asarray(value)
        self._dev = None

    def device_coeff(self) -> torch.Tensor:
        """The coefficients as a device tensor of the sector's dtype."""
        if self._dev is None:
            host = numpy.asarray(self._host)
            if host.shape != (self.lena(), self.lenb()):
                raise ValueError(
                    f"coefficient shape {host.shape} does not match the "
                    f"sector ({self.lena()}, {self.lenb()})")
            if not numpy.iscomplexobj(host):
                host = host.astype(self._dtype)
            # upload as given and convert on the device (a host-side
            # complex128 -> complex64 pass costs more than the wider copy)
            dev = torch.from_numpy(numpy.ascontiguousarray(host)).to(
                _lib.device())
            self._dev = dev.to(_TORCH_DTYPE[self._dtype]).contiguous()
            self._host = None
        return self._dev

    def set_device_coeff(self, tensor: torch.Tensor) -> None:
        tdt = _TORCH_DTYPE[self._dtype]
        if tensor.shape != (self.lena(), self.lenb()):
            raise ValueError("coefficient shape does not match the sector")
        self._dev = tensor.to(device=_lib.device(), dtype=tdt).contiguous()
        self._host = None

    # -- reference API --------------------------------------------------------
    def __getitem__(self, key: Tuple[int, int]) -> complex:
        return self.coeff[self._core.index_alpha(key[0]),
                          self._core.index_beta(key[1])]

    def __setitem__(self, key: Tuple[int, int], value: complex) -> None:
        self.coeff[self._core.index_alpha(key[0]),
                   self._core.index_beta(key[1])] = value

    def __deepcopy__(self, memodict={}) -> 'FqeData':
        new = FqeData.__new__(FqeData)
        new._core = self._core
        new._dtype = self._dtype
        new._nele = self._nele
        new._m_s = self._m_s
        if self._dev is not None:
            new._dev = self._dev.clone()
            new._host = None
        else:
            new._host = numpy.array(self._host, copy=True)
            new._dev = None
        return new

    def __hash__(self):
        return hash((self._nele, self._m_s))

    def get_fcigraph(self) -> FciGraph:
        return self._core

    def norb(self) -> int:
        return self._core.norb()

    def lena(self) -> int:
        return self._core.lena()

    def lenb(self) -> int:
        return self._core.lenb()

    def nalpha(self) -> int:
        return self._core.nalpha()

    def nbeta(self) -> int:
        return self._core.nbeta()

    def n_electrons(self) -> int:
        return self._nele

    def alpha_map(self, iorb: int, jorb: int):
        return self._core.alpha_map(iorb, jorb)

    def beta_map(self, iorb: int, jorb: int):
        return self._core.beta_map(iorb, jorb)

    def apply(self, array: Tuple[numpy.ndarray, ...]) -> 'FqeData':
        out = copy.deepcopy(self)
        out.apply_inplace(array)
        return out

    def apply_inplace(self, array: Tuple[numpy.ndarray, ...]) -> None:
        """Apply a dense spatial 1- or 1+2-body operator in place (reference
        ``src/fqe/fqe_data.py:421-475``)."""
        self.apply_inplace_scaled(array)

    def apply_inplace_scaled(self, array, alpha: complex = 1.0,
                             beta: complex = 0.0,
                             op: Optional[engine.PreparedOperator] = None):
        """coeff <- alpha * H coeff + beta * coeff on the device."""
        len_arr = len(array)
        if len_arr < 1 or len_arr > 4:
            raise ValueError("Number of operators in tuple must be "
                             "between 1 and 4.")
        first = next(filter(lambda x: isinstance(x, numpy.ndarray), array),
                     False)
        if isinstance(first, bool):
            return
        spatial = first.shape[0] == self.norb()
        if not spatial and first.shape[0] != 2 * self.norb():
            raise ValueError("Inconsistent number of spin-orbitals in "
                             "operators and wavefunction.")
        if len_arr > 2:
            raise NotImplementedError(
                "the GPU engine applies 1- and 1+2-body operators; 3- and "
                "4-body operators are outside its path")
        if op is None:
            if spatial:
                op = engine.prepare_operator(array, self.norb(),
                                             _lib.device())
            else:
                # spin-orbital tensors on this sector (reference
                # _apply_array_spin1 / _apply_array_spin12,
                # src/fqe/fqe_data.py:533-583, :610-635)
                op = engine.prepare_spin_operator(array, self.norb(),
                                                  _lib.device())
        cin = self.device_coeff()
        out = torch.empty_like(cin)
        engine.apply_sector(self._core, cin, op, out, alpha=alpha, beta=beta,
                            cin=cin if beta != 0 else None)
        self._dev = out
        self._host = None

    def calculate_dvec_spatial(self) -> numpy.ndarray:
        """D[i, j] = E_ij C as numpy (norb, norb, lena, lenb) (reference
        ``src/fqe/fqe_data.py:2149-2160``), gathered on the GPU."""
        return download(self.calculate_dvec_spatial_device())

    def calculate_dvec_spatial_device(self,
                                      compressed: bool = False) -> torch.Tensor:
        norb = self.norb()
        npair = norb * (norb + 1) // 2 if compressed else norb * norb
        c = self.device_coeff()
        prec = engine.precision_of(c.dtype)
        real = torch.float64 if prec == _lib.C128 else torch.float32
        n = self.lena() * self.lenb()
        d = torch.empty((npair, 2, n), dtype=real, device=c.device)
        ta = self._core.tables(True).pair_table(compressed)
        tb = self._core.tables(False).pair_table(compressed)
        _lib.call("fqehip_dvec_gather", prec, npair, self.lena(), self.lenb(),
                  _lib.ptr(ta), _lib.ptr(tb), _lib.ptr(c), _lib.ptr(d),
                  _lib.stream_ptr())
        d = torch.complex(d[:, 0], d[:, 1])
        shape = ((npair,) if compressed else (norb, norb)) + (self.lena(),
                                                               self.lenb())
        return d.reshape(shape)

    # -- reduced density matrices ------------------------------------------
    def _dvec128(self) -> torch.Tensor:
        """[norb^2, lena*lenb] complex128 device dvec (ordered pairs)."""
        d = self.calculate_dvec_spatial_device(compressed=False)
        n2 = self.norb() * self.norb()
        return d.reshape(n2, -1).to(torch.complex128).contiguous()

    def _coeff128(self) -> torch.Tensor:
        return self.device_coeff().reshape(-1).to(torch.complex128).contiguous()

    def rdm1(self, bradata: Optional['FqeData'] = None) -> Tuple[numpy.ndarray]:
        """1-RDM rdm1[p, q] = <bra|E_pq|ket> (reference
        ``src/fqe/fqe_data.py:1668-1720``): one dvec gather of the bra and
        norb^2 FP64 inner products with the ket on the device."""
        bra = self if bradata is None else bradata
        norb = self.norb()
        t = engine.block_dots(bra._dvec128(), self._coeff128())
        return (t.reshape(norb, norb).T.copy(),)

    def rdm12(self, bradata: Optional['FqeData'] = None
              ) -> Tuple[numpy.ndarray, numpy.ndarray]:
        """1- and 2-RDMs in the reference's convention
        (``src/fqe/fqe_data.py:1767-1782``):
        rdm2[j, k, i, l] = -<E_ij bra|E_kl ket> + delta_ki rdm1[j, l].
        The norb^2 x norb^2 overlap matrix of the two dvecs is one FP64
        matrix-core Gram product (fqehip_gram)."""
        bra = self if bradata is None else bradata
        norb = self.norb()
        n2 = norb * norb
        d = self._dvec128()
        d2 = d if bradata is None else bra._dvec128()
        out1 = engine.block_dots(d2, self._coeff128()).reshape(norb, norb).T
        m = engine.gram(d2, d)
        out2 = -m.reshape((norb,) * 4).transpose(1, 2, 0, 3).copy()
        for i in range(norb):
            out2[:, i, i, :] += out1
        return out1.copy(), out2

    def ax_plus_y(self, sval: complex, other: 'FqeData') -> 'FqeData':
        assert hash(self) == hash(other)
        self._dev = self.device_coeff() + other.device_coeff().to(
            _TORCH_DTYPE[self._dtype]) * sval
        self._host = None
        return self

    def norm(self) -> float:
        """FP64 2-norm on the device (fqehip_block_dot)."""
        c = self.device_coeff()
        c = (c.to(torch.complex128) if c.dtype != torch.complex128
             else c).contiguous()
        return float(numpy.sqrt(max(engine.dot128(c, c).real, 0.0)))

    def scale(self, sval: complex) -> None:
        self._dev = self.device_coeff() * sval
        self._host = None

    def fill(self, value: complex) -> None:
        self.device_coeff().fill_(value)

    def conj(self) -> None:
        self._dev = torch.conj(self.device_coeff()).resolve_conj()
        self._host = None

    def empty_copy(self) -> 'FqeData':
        new = copy.deepcopy(self)
        new._dev = torch.zeros_like(self.device_coeff())
        new._host = None
        return new

    def set_wfn(self, strategy: Optional[str] = None,
                raw_data: numpy.ndarray = numpy.empty(0)) -> None:
        """reference ``src/fqe/fqe_data.py:2761-2811``"""
        strategy_args = ['ones', 'zero', 'random', 'from_data', 'hartree-fock']
        if strategy is None and raw_data.shape == (0,):
            raise ValueError('No strategy and no data passed.'
                             ' Cannot initialize')
        if strategy == 'from_data' and raw_data.shape == (0,):
            raise ValueError('No data passed to initialize from')
        if raw_data.shape != (0,) and strategy not in ['from_data', None]:
            raise ValueError('Inconsistent strategy for set_vec passed with'
                             'data')
        if strategy not in strategy_args:
            raise ValueError('Unknown Argument passed to set_vec')
        if strategy == 'from_data':
            chkdim = raw_data.shape
            if chkdim[0] != self.lena() or chkdim[1] != self.lenb():
                raise ValueError('Dim of data passed {},{} is not compatible'
                                 ' with {},{}'.format(chkdim[0], chkdim[1],
                                                      self.lena(),
                                                      self.lenb()))
        if strategy == 'ones':
            self.set_device_coeff(torch.ones((self.lena(), self.lenb()),
                                             dtype=torch.complex128))
        elif strategy == 'zero':
            self.set_device_coeff(torch.zeros((self.lena(), self.lenb()),
                                              dtype=torch.complex128))
        elif strategy == 'random':
            self.coeff = rand_wfn(self.lena(), self.lenb())
        elif strategy == 'from_data':
            if isinstance(raw_data, torch.Tensor):
                self.set_device_coeff(raw_data)
            else:
                # the device tensor is the copy (no host-side copy first)
                self.set_device_coeff(
                    torch.from_numpy(numpy.ascontiguousarray(raw_data)))
        elif strategy == 'hartree-fock':
            z = numpy.zeros((self.lena(), self.lenb()), dtype=self._dtype)
            z[0, 0] = 1.0
            self.coeff = z
