"""Device form of the operator and the per-sector apply / Taylor drivers.

The operator of a spatial (h1, h2) tuple is put in the dvec form the
reference's Python path uses (``src/fqe/fqe_data.py:646-683``):

    H C = sum_ij h1'[i,j] E_ij C + sum_ijkl V[ij,kl] E_ij E_kl C,
    V[i,j,k,l] = -h2[i,k,j,l],  h1' = h1 - einsum('ikkj', V).

Pair basis.  When V[ij,kl] = V[ji,kl] = V[ij,lk] and h1' is symmetric
(real molecular integrals) the n(n+1)/2 unordered pairs carry the same
operator exactly (reference ``:659-681`` assumes this for any real h; here
it is checked, so a real non-symmetric h takes the ordered n^2 basis and
matches the reference's C path).  GEMM matrix A: rows = output pairs plus
one row for h1', columns = input pairs; complex operators use the real
2x2-block form [[Re, -Im], [Im, Re]] on the planar (re, im) dvec.
"""

from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy
import torch

from . import _lib
from . import settings


class PreparedOperator:
    """GEMM matrix (2-body) or pair vector (1-body) on the device."""

    def __init__(self, norb: int, compressed: bool, complex_op: bool,
                 a: Optional[torch.Tensor], h1: Optional[torch.Tensor],
                 spin: bool = False):
        self.norb = norb
        self.compressed = compressed
        self.complex_op = complex_op
        self.a = a
        self.h1 = h1
        # spin: pairs of alpha orbitals, then pairs of beta orbitals
        self.spin = spin
        self.npair = (2 * norb * norb if spin else
                      norb * (norb + 1) // 2 if compressed else norb * norb)
        # A[p][q] == A[q][p] exactly (p, q < npair): the split apply applies
        self.pair_symmetric = False
        self._hbb = {}
        self._hoz = {}

    def beta_block(self, graph):
        """Beta-beta block of this operator on the beta strings of
        ``graph`` (fqehip_beta_block): (bfloat16 hi, bfloat16 lo, ldh, beta
        forward table, FP64 columns), cached per (nbeta, device); None when
        the split apply does not apply."""
        if not self.pair_symmetric or self.a is None or self.complex_op:
            return None
        lenb = graph.lenb()
        if (lenb + self.npair) * 8 > 160 * 1024:
            return None
        key = (graph.nbeta(), self.a.device.index)
        if key not in self._hbb:
            fwb, _ = graph.tables(False).fused_tables(self.compressed)
            ldh = int(_lib.call("fqehip_beta_block_ld", lenb))
            hcol = torch.empty((lenb, lenb), dtype=torch.float64,
                               device=self.a.device)
            hhi = torch.empty((ldh, ldh), dtype=torch.int16,
                              device=self.a.device)
            hlo = torch.empty_like(hhi)
            _lib.call("fqehip_beta_block", self.npair, lenb, _lib.ptr(fwb),
                      _lib.ptr(self.a), _lib.ptr(hcol), _lib.ptr(hhi),
                      _lib.ptr(hlo), ldh, _lib.stream_ptr())
            self._hbb[key] = (hhi, hlo, ldh, fwb, hcol)
        return self._hbb[key]

    def beta_block_digits(self, graph):
        """Ozaki digits of the FP64 beta-beta block (int8 digit planes,
        column scales and offset corrections; fqehip_beta_block_digits),
        cached per (nbeta, device); None when the block does not apply."""
        blk = self.beta_block(graph)
        if blk is None:
            return None
        key = (graph.nbeta(), self.a.device.index)
        if key not in self._hoz:
            lenb = graph.lenb()
            nbytes = int(_lib.call("fqehip_beta_block_digits_bytes", lenb))
            hoz = torch.empty(max(nbytes, 16), dtype=torch.uint8,
                              device=self.a.device)
            _lib.call("fqehip_beta_block_digits", lenb, _lib.ptr(blk[4]),
                      _lib.ptr(hoz), _lib.stream_ptr())
            self._hoz[key] = hoz
        return self._hoz[key]

    def c_struct(self) -> _lib.Operator:
        return _lib.Operator(int(self.complex_op), _lib.ptr(self.a),
                             _lib.ptr(self.h1))


def _pair_indices(norb: int, compressed: bool):
    """(i, j) of each pair-basis index (representative i >= j when
    compressed, index j + i(i+1)/2)."""
    if compressed:
        return [(i, j) for i in range(norb) for j in range(i + 1)]
    return [(i, j) for i in range(norb) for j in range(norb)]


def _is_pair_symmetric(h1p: numpy.ndarray, v: numpy.ndarray) -> bool:
    return (numpy.array_equal(v, v.transpose(1, 0, 2, 3))
            and numpy.array_equal(v, v.transpose(0, 1, 3, 2))
            and numpy.array_equal(h1p, h1p.T))


def prepare_operator(tensors: Sequence[numpy.ndarray], norb: int,
                     device: torch.device) -> PreparedOperator:
    """Device form of a 1- or 2-tensor spatial operator."""
    if len(tensors) == 1:
        h1 = numpy.asarray(tensors[0], dtype=numpy.complex128)
        if h1.shape != (norb, norb):
            raise ValueError("Inconsistent number of spin-orbitals in "
                             "operators and wavefunction.")
        compressed = numpy.array_equal(h1, h1.T)
        pairs = _pair_indices(norb, compressed)
        vec = numpy.array([h1[i, j] for i, j in pairs], dtype=numpy.complex128)
        d = torch.from_numpy(vec.view(numpy.float64).copy()).to(device)
        return PreparedOperator(norb, compressed, bool(numpy.any(vec.imag)),
                                None, d)
    if len(tensors) != 2:
        raise ValueError("the GPU engine applies 1- and 2-body restricted "
                         "operators")
    h1e, h2e = tensors
    if h1e is None:
        h1e = numpy.zeros((norb, norb))
    h1e = numpy.asarray(h1e)
    h2e = numpy.asarray(h2e)
    assert h1e.shape == (norb, norb)
    assert h2e.shape == (norb, norb, norb, norb)
    v = -numpy.moveaxis(h2e.astype(numpy.complex128), 1, 2)
    h1p = h1e.astype(numpy.complex128) - numpy.einsum("ikkj->ij", v)
    compressed = _is_pair_symmetric(h1p, v)
    complex_op = bool(numpy.any(v.imag) or numpy.any(h1p.imag))
    pairs = _pair_indices(norb, compressed)
    rows = [i * norb + j for i, j in pairs]
    vm = v.reshape(norb * norb, norb * norb)[numpy.ix_(rows, rows)]
    h1row = h1p.reshape(-1)[rows]
    full = numpy.vstack([vm, h1row[None, :]])  # (npair+1) x npair, complex
    a = _real_block_form(full, complex_op)
    dev_a = torch.from_numpy(numpy.ascontiguousarray(a, dtype=numpy.float64)
                             ).to(device)
    prepared = PreparedOperator(norb, compressed, complex_op, dev_a, None)
    if not complex_op:
        sq = a[:-1, :]
        prepared.pair_symmetric = bool(numpy.array_equal(sq, sq.T))
    return prepared


def _real_block_form(full: numpy.ndarray, complex_op: bool) -> numpy.ndarray:
    """The GEMM matrix of a complex (npair+1) x npair operator: itself when
    real, else the real 2 x 2 block form on planar (re, im) rows."""
    if not complex_op:
        return numpy.ascontiguousarray(full.real)
    m, k = full.shape
    a = numpy.zeros((2 * m, 2 * k))
    a[0::2, 0::2] = full.real
    a[0::2, 1::2] = -full.imag
    a[1::2, 0::2] = full.imag
    a[1::2, 1::2] = full.real
    return a


def prepare_spin_operator(tensors: Sequence[numpy.ndarray], norb: int,
                          device: torch.device) -> PreparedOperator:
    """Device form of a 1- or 2-tensor spin-orbital operator (2 norb
    spin orbitals, alpha first) applied to one (nalpha, nbeta) sector, in
    the reference's dvec form with separate alpha and beta dvecs
    (``src/fqe/fqe_data.py:533-583`` one-body, ``:712-741`` two-body):

        out = sum_ij h1'aa[i,j] Da[i,j] + h1'bb[i,j] Db[i,j]
              + sum_ij E^a_ij Ya[i,j] + E^b_ij Yb[i,j],
        Ya = V'aaaa Da + V'aabb Db,  Yb = V'bbaa Da + V'bbbb Db,

    with V' = -moveaxis(h2e, 1, 2) and h1' = h1e - sum_k V'[i,k,k,j] over
    all 2 norb spin orbitals.  Only the spin-conserving blocks enter, as in
    the reference; the alpha-beta blocks of h1e are not used.  The pair
    basis is the full norb^2 basis of each spin, alpha pairs first: a
    (2 norb^2 + 1) x 2 norb^2 GEMM matrix (the last row h1'), or a pair
    vector for a one-body operator."""
    n = norb
    m = 2 * n
    npair = n * n
    if len(tensors) == 1:
        h1 = numpy.asarray(tensors[0], dtype=numpy.complex128)
        if h1.shape != (m, m):
            raise ValueError("Inconsistent number of spin-orbitals in "
                             "operators and wavefunction.")
        vec = numpy.concatenate([h1[:n, :n].reshape(-1),
                                 h1[n:, n:].reshape(-1)])
        d = torch.from_numpy(vec.view(numpy.float64).copy()).to(device)
        return PreparedOperator(norb, False, bool(numpy.any(vec.imag)), None,
                                d, spin=True)
    if len(tensors) != 2:
        raise ValueError("the GPU engine applies 1- and 2-body spin-orbital "
                         "operators to spin-conserving sectors")
    h1e, h2e = tensors
    if h1e is None:
        h1e = numpy.zeros((m, m))
    h1e = numpy.asarray(h1e)
    h2e = numpy.asarray(h2e)
    assert h1e.shape == (m, m)
    assert h2e.shape == (m, m, m, m)
    v = -numpy.moveaxis(h2e.astype(numpy.complex128), 1, 2)
    h1p = h1e.astype(numpy.complex128) - numpy.einsum("ikkj->ij", v)
    vm = numpy.zeros((2 * npair, 2 * npair), dtype=numpy.complex128)
    for s in range(2):
        for t in range(2):
            vm[s * npair:(s + 1) * npair, t * npair:(t + 1) * npair] = \
                v[s * n:(s + 1) * n, s * n:(s + 1) * n,
                  t * n:(t + 1) * n, t * n:(t + 1) * n].reshape(npair, npair)
    h1row = numpy.concatenate([h1p[:n, :n].reshape(-1),
                               h1p[n:, n:].reshape(-1)])
    full = numpy.vstack([vm, h1row[None, :]])
    complex_op = bool(numpy.any(full.imag))
    a = _real_block_form(full, complex_op)
    dev_a = torch.from_numpy(numpy.ascontiguousarray(a, dtype=numpy.float64)
                             ).to(device)
    return PreparedOperator(norb, False, complex_op, dev_a, None, spin=True)


def precision_of(dtype: torch.dtype) -> int:
    if dtype == torch.complex128:
        return _lib.C128
    if dtype == torch.complex64:
        return _lib.C64
    raise TypeError(f"unsupported coefficient dtype {dtype}")


def sector_struct(graph, op: PreparedOperator,
                  precision: Optional[int] = None
                  ) -> Tuple[_lib.Sector, tuple]:
    """C descriptor of a sector for the pair basis of ``op``; the second
    element keeps the tables referenced.  For a pair-symmetric operator the
    beta-beta block is attached (split apply): its bfloat16 hi/lo split for
    complex64 (``settings.split``), its FP64 columns for complex128
    (``settings.split_fp64``)."""
    if op.spin:
        ta = graph.tables(True).spin_pair_table(True)
        tb = graph.tables(False).spin_pair_table(False)
    else:
        ta = graph.tables(True).pair_table(op.compressed)
        tb = graph.tables(False).pair_table(op.compressed)
    keep = [ta, tb]
    fwa = fwb = sla = None
    if settings.fused and not op.complex_op and op.h1 is None and not op.spin:
        fwa, sla = graph.tables(True).fused_tables(op.compressed)
        fwb, _ = graph.tables(False).fused_tables(op.compressed)
        keep += [fwa, fwb, sla]
    nslot = graph.nalpha() * (graph.norb() - graph.nalpha())
    hbb = hbb64 = None
    if fwa is not None and settings.split and precision == _lib.C64:
        hbb = op.beta_block(graph)
    # split kernel limits: gather depth npair + 1 <= 128 (112 for the FP64
    # shared-D-tile variant), at most 64 live pair rows (norb^2/4 + 1)
    kmax = 112 if (settings.contraction == "fp64" and
                   settings.split_tile_fp64 == _lib.TILE_SPLIT_F64_DTILE) else 128
    if (fwa is not None and settings.split_fp64 and precision == _lib.C128
            and op.npair + 1 <= kmax and graph.norb() ** 2 // 4 + 1 <= 64):
        blk = op.beta_block(graph)
        hbb64 = blk[4] if blk else None
    s = _lib.Sector(graph.norb(), graph.nalpha(), graph.nbeta(), op.npair,
                    graph.lena(), graph.lenb(), _lib.ptr(ta), _lib.ptr(tb),
                    _lib.ptr(fwa), _lib.ptr(fwb), _lib.ptr(sla), nslot,
                    _lib.ptr(hbb[0]) if hbb else None,
                    _lib.ptr(hbb[1]) if hbb else None,
                    hbb[2] if hbb else 0)
    if hbb:
        keep += [hbb[0], hbb[1]]
        s.tile = int(settings.split_tile)
    if hbb64 is not None:
        keep.append(hbb64)
        s.hbb_f64 = _lib.ptr(hbb64)
        if settings.contraction == "ozaki":
            s.tile = _lib.TILE_OZAKI
            if settings.ozaki_beta and graph.lenb() < 26214:
                hoz = op.beta_block_digits(graph)
                if hoz is not None:
                    keep.append(hoz)
                    s.hbb_oz = _lib.ptr(hoz)
        elif settings.contraction == "fp64":
            s.tile = int(settings.split_tile_fp64)
        else:
            raise ValueError(f"settings.contraction={settings.contraction!r}")
    return s, tuple(keep)


def workspace(sector: _lib.Sector, precision: int,
              device: torch.device) -> torch.Tensor:
    """Device scratch for one apply (dvec + Y planes); the one-body Taylor
    loop uses it as its second term buffer."""
    nbytes = int(_lib.lib().fqehip_apply_workspace_bytes(precision,
                                                          _as_ref(sector)))
    elem = 8 if precision == _lib.C128 else 4
    n = sector.lena * sector.lenb
    fused = (sector.nslot + 1) * 2 * n * elem
    if sector.tile == _lib.TILE_OZAKI and sector.hbb_f64:
        fused += 512 * sector.lenb        # beta gather table
        fused += 128 * (5 * 128 + 12)     # pair rows' digits
        fused += 8 * sector.lena          # row maxima of |C|
        if sector.hbb_oz:                 # C digit planes and row scales
            ldk = (sector.lenb + 63) // 64 * 64
            fused += 6 * 2 * sector.lena * ldk + 8 * sector.lena
    need = max((2 * sector.npair + 1) * 2 * n * elem, fused)
    if nbytes != need:
        raise _lib.FqeHipError(f"workspace size {nbytes} != {need}")
    nbytes = max(nbytes, 16 * sector.lena * sector.lenb, 16)
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def needs_panels(sector: _lib.Sector, op: PreparedOperator,
                 precision: int) -> bool:
    """True for a complex128 two-body apply on the unfused kernels whose
    dvec + Y exceed settings.max_workspace_bytes."""
    if precision != _lib.C128 or op.h1 is not None:
        return False
    import ctypes
    opc = op.c_struct()
    if _lib.lib().fqehip_fused_supported(precision, ctypes.byref(sector),
                                         ctypes.byref(opc)):
        return False
    n = sector.lena * sector.lenb
    return (2 * sector.npair + 1) * 2 * n * 8 > settings.max_workspace_bytes


def _as_ref(struct):
    import ctypes
    return ctypes.byref(struct)


def apply_sector(graph, coeff: torch.Tensor, op: PreparedOperator,
                 out: torch.Tensor, alpha: complex = 1.0, beta: complex = 0.0,
                 cin: Optional[torch.Tensor] = None,
                 acc: Optional[torch.Tensor] = None,
                 partials: Optional[torch.Tensor] = None,
                 done: Optional[torch.Tensor] = None,
                 ws: Optional[torch.Tensor] = None) -> None:
    """out = alpha * H coeff + beta * cin on one sector (device tensors).
    A complex128 sector whose unfused workspace exceeds
    ``settings.max_workspace_bytes`` runs in panels of orbital pairs."""
    prec = precision_of(coeff.dtype)
    s, keep = sector_struct(graph, op, prec)
    if ws is None and op.h1 is None and needs_panels(s, op, prec):
        nbytes = int(settings.max_workspace_bytes)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=coeff.device)
        opc = op.c_struct()
        _lib.call("fqehip_apply_panels", prec, _as_ref(s), _as_ref(opc),
                  _lib.ptr(coeff), float(numpy.real(alpha)),
                  float(numpy.imag(alpha)), float(numpy.real(beta)),
                  float(numpy.imag(beta)), _lib.ptr(cin), _lib.ptr(out),
                  _lib.ptr(acc), _lib.ptr(partials), _lib.ptr(done),
                  _lib.ptr(ws), nbytes, _lib.stream_ptr())
        del keep
        return
    if ws is None and op.h1 is None:
        ws = workspace(s, prec, coeff.device)
    opc = op.c_struct()
    _lib.call("fqehip_apply", prec, _as_ref(s), _as_ref(opc),
              _lib.ptr(coeff), float(numpy.real(alpha)),
              float(numpy.imag(alpha)), float(numpy.real(beta)),
              float(numpy.imag(beta)), _lib.ptr(cin), _lib.ptr(out),
              _lib.ptr(acc), _lib.ptr(partials), _lib.ptr(done),
              _lib.ptr(ws), _lib.stream_ptr())
    del keep


def norm_partials_count(n: int) -> int:
    return int(_lib.lib().fqehip_norm_partials_count(n))


_DOT_WS = {}


def dot128(a: torch.Tensor, b: torch.Tensor) -> complex:
    """<a|b> = sum conj(a) b over complex128 device vectors, FP64 with a
    fixed summation order (fqehip_block_dot)."""
    a = a.reshape(-1)
    b = b.reshape(-1)
    assert a.dtype == torch.complex128 and b.dtype == torch.complex128
    assert a.is_contiguous() and b.is_contiguous() and a.numel() == b.numel()
    dev = a.device
    if dev.index not in _DOT_WS:
        nbytes = int(_lib.call("fqehip_block_dot_workspace_bytes", 1))
        _DOT_WS[dev.index] = (torch.empty(nbytes, dtype=torch.uint8,
                                          device=dev),
                              torch.empty(1, dtype=torch.complex128,
                                          device=dev))
    ws, out = _DOT_WS[dev.index]
    _lib.call("fqehip_block_dot", a.numel(), 1, _lib.ptr(a), a.numel(),
              _lib.ptr(b), _lib.ptr(out), _lib.ptr(ws), _lib.stream_ptr())
    return complex(out.item())


def block_dots(rows: torch.Tensor, x: torch.Tensor) -> numpy.ndarray:
    """[sum_i conj(rows[j, i]) x[i] for j] over complex128 device data
    (fqehip_block_dot, FP64, fixed summation order)."""
    k, n = rows.shape
    assert rows.dtype == torch.complex128 and x.dtype == torch.complex128
    assert rows.is_contiguous() and x.is_contiguous() and x.numel() == n
    ws = torch.empty(max(int(_lib.call("fqehip_block_dot_workspace_bytes",
                                       k)), 16),
                     dtype=torch.uint8, device=rows.device)
    out = torch.empty(k, dtype=torch.complex128, device=rows.device)
    _lib.call("fqehip_block_dot", n, k, _lib.ptr(rows), n, _lib.ptr(x),
              _lib.ptr(out), _lib.ptr(ws), _lib.stream_ptr())
    return out.cpu().numpy()


def gram(a: torch.Tensor, b: torch.Tensor) -> numpy.ndarray:
    """G[p, q] = sum_i conj(a[p, i]) b[q, i] for complex128 device blocks
    a, b of shape [k, n] (fqehip_gram, FP64 matrix cores)."""
    k, n = a.shape
    assert b.shape == (k, n)
    assert a.dtype == torch.complex128 and b.dtype == torch.complex128
    assert a.is_contiguous() and b.is_contiguous()
    ws = torch.empty(max(int(_lib.call("fqehip_gram_workspace_bytes", n, k)),
                         16), dtype=torch.uint8, device=a.device)
    out = torch.empty((k, k), dtype=torch.complex128, device=a.device)
    _lib.call("fqehip_gram", n, k, _lib.ptr(a), _lib.ptr(b), _lib.ptr(out),
              _lib.ptr(ws), _lib.stream_ptr())
    return out.cpu().numpy()

