"""CPU oracle for the spin-orbital (``FqeDataSet``) apply and RDMs.

TEST INFRASTRUCTURE ONLY (same import rule as ``oracle/fqe_oracle.py``:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import it, and only as the checker).

An ``FqeDataSet`` holds every (nalpha, nbeta) sector of one total particle
number ``nele`` on ``norb`` spatial orbitals (reference
``src/fqe/fqe_data_set.py:35-60``).  Spin-orbital index p < norb is alpha
orbital p, p >= norb is beta orbital p - norb.  A one-body excitation
E_pq = a+_p a_q then either stays in the sector (same spin) or moves one
electron between spins, (na, nb) -> (na + 1, nb - 1) for a+_{i alpha}
a_{j beta} and (na, nb) -> (na - 1, nb + 1) for a+_{i beta} a_{j alpha}.

This is a NumPy restatement, written from the mathematics, of

* ``_calculate_dvec_with_coeff`` (reference ``src/fqe/fqe_data_set.py:
  737-796``): D[key][p, q] = sum over source sectors of E_pq C, with the
  spin-flip sign (-1)^nalpha (alpha created, beta removed) or
  (-1)^(nalpha - 1) (beta created, alpha removed) times the single-string
  parities of the maps ``FciGraph.find_mapping`` returns, which count the
  occupied orbitals ABOVE the orbital (reference
  ``src/fqe/fci_graph_set.py:137-150``; ``src/fqe/fci_graph.py:189-202``);
* ``_calculate_coeff_with_dvec`` (reference ``src/fqe/fqe_data_set.py:
  878-960``): out[key'] = sum_pq E_pq Y[key][p, q];
* ``_apply1`` / ``_apply12`` (reference ``src/fqe/fqe_data_set.py:
  153-209``), with the same dvec form as the restricted apply:
  V[i,j,k,l] = -h2[i,k,j,l], h1' = h1 - sum_k V[:,k,k,:];
* ``rdm1`` / ``rdm12`` (reference ``src/fqe/fqe_data_set.py:502-559``).

Every excitation is kept as a pair of small dense string matrices, so that
E_pq C = s * A @ C @ B.T with A acting on alpha strings and B on beta
strings.  The string tables and the same-spin maps are those of
``oracle/fqe_oracle.py``.

Sign convention.  Counting above rather than below makes the spin-flip
terms differ from a Jordan-Wigner operator on modes (alpha 0..norb-1, beta
norb..2 norb-1) by (-1)^(nele - 1): with nalpha + nbeta = nele, the
above-counts are (nalpha - below) and (nbeta - 1 - below) for alpha
created / beta removed, and (nalpha - 1 - below) and (nbeta - below) the
other way.  That is the same as giving sector (nalpha, nbeta) the phase
(-1)^((nele - 1) nbeta) against the Jordan-Wigner determinants; for odd
nele the two agree.

How it is pinned (tests/test_set_oracle.py): the reference's FqeDataSet
fixtures for nele=2, norb=2 (``tests/unittest_data/fqe_data_set/set0202``:
apply of h1, of h2 and of h1+h2, and the spin-orbital 1- and 2-RDMs, the
RDMs fixing the spin-flip sign); a Jordan-Wigner dense operator over all
sectors with that sector phase (``oracle/jw_dense.py``, which applies
fermion operators to combined occupation states and never uses the
per-spin sign split above); and the restricted oracle per sector when the
spin-orbital tensors are spin-diagonal.
"""

from __future__ import annotations

from typing import Dict, List, Tuple

import numpy

from .fqe_oracle import OracleGraph

Key = Tuple[int, int]


def set_keys(nele: int, norb: int) -> List[Key]:
    """Sectors of an FqeDataSet of nele electrons, nbeta ascending, as the
    reference's test loader builds them (reference
    ``tests/unittest_data/fqe_data_set_loader.py:26-37``); sectors with more
    than norb electrons of one spin are left out."""
    keys = []
    for nbeta in range(nele + 1):
        nalpha = nele - nbeta
        if nalpha <= norb and nbeta <= norb:
            keys.append((nalpha, nbeta))
    return keys


def _string_matrix(src: numpy.ndarray, tgt: numpy.ndarray, orb: int,
                   create: bool) -> numpy.ndarray:
    """Dense matrix of a single a+_orb (create) or a_orb on one spin's
    strings, with the parity of the reference's sector-to-sector maps:
    (-1)^(occupied orbitals above orb) (reference
    ``src/fqe/fci_graph_set.py:137-150``, ``count_bits_above``)."""
    index = {int(s): a for a, s in enumerate(tgt)}
    mat = numpy.zeros((len(tgt), len(src)))
    for a, s in enumerate(src):
        s = int(s)
        occupied = (s >> orb) & 1
        if occupied == (0 if create else 1):
            above = bin(s >> (orb + 1)).count("1")
            mat[index[s ^ (1 << orb)], a] = -1.0 if above % 2 else 1.0
    return mat


def _pair_matrix(maps, nstr: int, i: int, j: int) -> numpy.ndarray:
    """Dense matrix of a+_i a_j on one spin's strings from the oracle's
    (source, target, sign) maps."""
    mat = numpy.zeros((nstr, nstr))
    for src, tgt, sign in maps[(i, j)]:
        mat[tgt, src] += sign
    return mat


class SetGraph:
    """String tables of every sector of one (nele, norb) FqeDataSet and the
    excitation list of all E_pq, p, q in [0, 2 norb)."""

    def __init__(self, nele: int, norb: int) -> None:
        self.nele = nele
        self.norb = norb
        self.keys = set_keys(nele, norb)
        self.graphs: Dict[Key, OracleGraph] = {
            k: OracleGraph(k[0], k[1], norb) for k in self.keys}
        # (p, q, src, tgt, A, B): E_pq C_src contributes A @ C_src @ B.T
        # to sector tgt; the spin-flip sign is folded into A.
        self.excitations = []
        n = norb
        for key in self.keys:
            g = self.graphs[key]
            na, nb = key
            eye_a = numpy.eye(g.lena)
            eye_b = numpy.eye(g.lenb)
            for i in range(n):
                for j in range(n):
                    if na > 0:
                        self.excitations.append(
                            (i, j, key, key,
                             _pair_matrix(g.alpha_map, g.lena, i, j), eye_b))
                    if nb > 0:
                        self.excitations.append(
                            (i + n, j + n, key, key, eye_a,
                             _pair_matrix(g.beta_map, g.lenb, i, j)))
            up = (na + 1, nb - 1)
            if up in self.graphs:
                t = self.graphs[up]
                sign = -1.0 if na % 2 else 1.0
                for i in range(n):
                    amat = sign * _string_matrix(g.astr, t.astr, i, True)
                    for j in range(n):
                        bmat = _string_matrix(g.bstr, t.bstr, j, False)
                        self.excitations.append((i, j + n, key, up, amat,
                                                 bmat))
            down = (na - 1, nb + 1)
            if down in self.graphs:
                t = self.graphs[down]
                sign = -1.0 if (na - 1) % 2 else 1.0
                for j in range(n):
                    amat = sign * _string_matrix(g.astr, t.astr, j, False)
                    for i in range(n):
                        bmat = _string_matrix(g.bstr, t.bstr, i, True)
                        self.excitations.append((i + n, j, key, down, amat,
                                                 bmat))

    def shape(self, key: Key) -> Tuple[int, int]:
        g = self.graphs[key]
        return g.lena, g.lenb

    def zeros(self) -> Dict[Key, numpy.ndarray]:
        return {k: numpy.zeros(self.shape(k), dtype=numpy.complex128)
                for k in self.keys}


def set_dvec(coeffs: Dict[Key, numpy.ndarray],
             sg: SetGraph) -> Dict[Key, numpy.ndarray]:
    """D[key][p, q] = sum_src <key| E_pq |src> C_src (reference
    ``src/fqe/fqe_data_set.py:737-796``)."""
    m = 2 * sg.norb
    out = {k: numpy.zeros((m, m) + sg.shape(k), dtype=numpy.complex128)
           for k in sg.keys}
    for p, q, src, tgt, amat, bmat in sg.excitations:
        out[tgt][p, q] += amat @ coeffs[src] @ bmat.T
    return out


def set_coeff_from_dvec(y: Dict[Key, numpy.ndarray],
                        sg: SetGraph) -> Dict[Key, numpy.ndarray]:
    """out[tgt] = sum_pq E_pq Y[src][p, q] (reference
    ``src/fqe/fqe_data_set.py:878-960``)."""
    out = sg.zeros()
    for p, q, src, tgt, amat, bmat in sg.excitations:
        out[tgt] += amat @ y[src][p, q] @ bmat.T
    return out


def set_apply_1(coeffs, h1e, sg: SetGraph) -> Dict[Key, numpy.ndarray]:
    """out = sum_pq h1[p, q] E_pq C (reference
    ``src/fqe/fqe_data_set.py:153-181``)."""
    h1e = numpy.asarray(h1e)
    d = set_dvec(coeffs, sg)
    return {k: numpy.tensordot(h1e, d[k]) for k in sg.keys}


def set_apply_12(coeffs, h1e, h2e, sg: SetGraph) -> Dict[Key, numpy.ndarray]:
    """out = sum h1' D + sum_pq E_pq (V . D) with V[i,j,k,l] = -h2[i,k,j,l]
    and h1' = h1 - sum_k V[:,k,k,:] (reference
    ``src/fqe/fqe_data_set.py:183-209``)."""
    v = -numpy.moveaxis(numpy.asarray(h2e, dtype=numpy.complex128), 1, 2)
    h1p = numpy.asarray(h1e, dtype=numpy.complex128) - numpy.einsum(
        "ikkj->ij", v)
    d = set_dvec(coeffs, sg)
    out = {k: numpy.tensordot(h1p, d[k]) for k in sg.keys}
    y = {k: numpy.tensordot(v, d[k]) for k in sg.keys}
    extra = set_coeff_from_dvec(y, sg)
    return {k: out[k] + extra[k] for k in sg.keys}


def set_apply(coeffs, tensors, sg: SetGraph) -> Dict[Key, numpy.ndarray]:
    """``FqeDataSet.apply`` for 1- or 2-tensor tuples (reference
    ``src/fqe/fqe_data_set.py:130-151``)."""
    if len(tensors) == 1:
        return set_apply_1(coeffs, tensors[0], sg)
    if len(tensors) == 2:
        return set_apply_12(coeffs, tensors[0], tensors[1], sg)
    raise ValueError("oracle handles 1- and 2-body spin-orbital tensors")


def set_rdm12(coeffs, sg: SetGraph, bra=None):
    """Spin-orbital 1- and 2-RDMs (reference
    ``src/fqe/fqe_data_set.py:523-559``): rdm1 = (sum conj(D2) C)^T,
    rdm2[j, k, i, l] = -sum conj(D2[i, j]) D[k, l] + delta_ki rdm1[j, l],
    summed over sectors."""
    d = set_dvec(coeffs, sg)
    d2 = d if bra is None else set_dvec(bra, sg)
    ket = coeffs
    out1 = None
    out2 = None
    for k in sg.keys:
        t1 = numpy.tensordot(d2[k].conj(), ket[k]).T
        t2 = -numpy.tensordot(d2[k].conj(), d[k],
                              axes=((2, 3), (2, 3))).transpose(1, 2, 0, 3)
        out1 = t1 if out1 is None else out1 + t1
        out2 = t2 if out2 is None else out2 + t2
    for i in range(out1.shape[0]):
        out2[:, i, i, :] += out1
    return out1, out2
