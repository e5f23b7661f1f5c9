"""Brute-force second-quantised Hamiltonian in FQE's determinant basis.

TEST INFRASTRUCTURE ONLY (same import rule as ``oracle/fqe_oracle.py``).

Independent of the oracle's excitation tables: it applies the operator
strings of the reference's restricted Hamiltonian

    H = sum_ij h1[i,j] sum_s a+_{is} a_{js}
      + sum_ijkl h2[i,j,k,l] sum_{rho,eta} a+_{i rho} a+_{j eta} a_{k rho} a_{l eta}

(reference docstring ``src/fqe/wavefunction.py:370-372``) one fermion
operator at a time to occupation-number states, using only the canonical
anticommutation sign rule.  The determinant |Ia, Ib> is
prod_{i in Ia, ascending} a+_{i alpha} prod_{j in Ib, ascending} a+_{j beta}
|0> (reference ``src/fqe/fci_graph.py:99-105``), i.e. the occupation state
with alpha orbitals on modes 0..norb-1 and beta orbitals on modes
norb..2*norb-1, creation operators ordered by ascending mode.

Addresses are the lexicographic ranks of the ascending occupation tuples.

The second half builds general spin-orbital operators (modes = spin-orbital
indices, spin-flip terms included) over every sector of one particle number
(``dense_set_operator``, ``set_expectation``); ``set_sector_phase`` gives
the phases that turn these determinants into the reference's FqeDataSet
ones.
"""

from __future__ import annotations

import itertools
from typing import Optional, Tuple

import numpy


def _strings(norb: int, nele: int):
    out = []
    for occ in itertools.combinations(range(norb), nele):
        out.append(sum(1 << o for o in occ))
    return out


def _annihilate(state: int, mode: int) -> Optional[Tuple[int, int]]:
    if not (state >> mode) & 1:
        return None
    sign = -1 if bin(state & ((1 << mode) - 1)).count("1") % 2 else 1
    return state & ~(1 << mode), sign


def _create(state: int, mode: int) -> Optional[Tuple[int, int]]:
    if (state >> mode) & 1:
        return None
    sign = -1 if bin(state & ((1 << mode) - 1)).count("1") % 2 else 1
    return state | (1 << mode), sign


def _apply_string(state: int, ops) -> Optional[Tuple[int, int]]:
    """ops: sequence of (mode, is_creation), applied right to left."""
    sign = 1
    for mode, dag in reversed(ops):
        res = _create(state, mode) if dag else _annihilate(state, mode)
        if res is None:
            return None
        state, s = res
        sign *= s
    return state, sign


def dense_hamiltonian(h1e: numpy.ndarray, h2e: Optional[numpy.ndarray],
                      norb: int, nalpha: int, nbeta: int) -> numpy.ndarray:
    """Dense matrix of H on the (nalpha, nbeta) sector, rows/columns in the
    FQE flat order Ia * lenb + Ib."""
    astr = _strings(norb, nalpha)
    bstr = _strings(norb, nbeta)
    lenb = len(bstr)
    basis = {}
    for ia, sa in enumerate(astr):
        for ib, sb in enumerate(bstr):
            basis[sa | (sb << norb)] = ia * lenb + ib
    dim = len(basis)
    mat = numpy.zeros((dim, dim), dtype=numpy.complex128)

    def mode(orb: int, spin: int) -> int:
        return orb + spin * norb

    terms = []
    for i in range(norb):
        for j in range(norb):
            if h1e[i, j] != 0.0:
                for s in range(2):
                    terms.append((h1e[i, j], ((mode(i, s), True),
                                              (mode(j, s), False))))
    if h2e is not None:
        for i, j, k, l in itertools.product(range(norb), repeat=4):
            if h2e[i, j, k, l] == 0.0:
                continue
            for rho in range(2):
                for eta in range(2):
                    terms.append((h2e[i, j, k, l],
                                  ((mode(i, rho), True), (mode(j, eta), True),
                                   (mode(k, rho), False),
                                   (mode(l, eta), False))))

    for state, col in basis.items():
        for coef, ops in terms:
            res = _apply_string(state, ops)
            if res is None:
                continue
            new, sign = res
            mat[basis[new], col] += sign * coef
    return mat


# ---------------------------------------------------------------------------
# Spin-orbital operators over every sector of one particle number
# ---------------------------------------------------------------------------

def set_basis(norb: int, nele: int):
    """Combined basis of all (nalpha, nbeta) sectors with nalpha + nbeta =
    nele, sectors in ascending nbeta, each in the FQE flat order.

    Returns (index, blocks): index maps an occupation state (alpha on modes
    0..norb-1, beta on modes norb..2*norb-1) to its position, blocks lists
    ((nalpha, nbeta), offset, lena, lenb)."""
    index = {}
    blocks = []
    offset = 0
    for nbeta in range(nele + 1):
        nalpha = nele - nbeta
        if nalpha > norb or nbeta > norb:
            continue
        astr = _strings(norb, nalpha)
        bstr = _strings(norb, nbeta)
        for ia, sa in enumerate(astr):
            for ib, sb in enumerate(bstr):
                index[sa | (sb << norb)] = offset + ia * len(bstr) + ib
        blocks.append(((nalpha, nbeta), offset, len(astr), len(bstr)))
        offset += len(astr) * len(bstr)
    return index, blocks


def dense_set_operator(h1e: numpy.ndarray, h2e: Optional[numpy.ndarray],
                       norb: int, nele: int) -> numpy.ndarray:
    """Dense matrix of sum_pq h1[p,q] a+_p a_q + sum_pqrs h2[p,q,r,s]
    a+_p a+_q a_r a_s over the combined basis of ``set_basis``; p, q, r, s
    are spin-orbital modes (p < norb alpha, p >= norb beta)."""
    index, _ = set_basis(norb, nele)
    nmode = 2 * norb
    terms = []
    for p, q in itertools.product(range(nmode), repeat=2):
        if h1e[p, q] != 0.0:
            terms.append((h1e[p, q], ((p, True), (q, False))))
    if h2e is not None:
        for p, q, r, s in itertools.product(range(nmode), repeat=4):
            if h2e[p, q, r, s] != 0.0:
                terms.append((h2e[p, q, r, s],
                              ((p, True), (q, True), (r, False), (s, False))))
    mat = numpy.zeros((len(index), len(index)), dtype=numpy.complex128)
    for state, col in index.items():
        for coef, ops in terms:
            res = _apply_string(state, ops)
            if res is None:
                continue
            new, sign = res
            mat[index[new], col] += sign * coef
    return mat


def set_sector_phase(norb: int, nele: int) -> numpy.ndarray:
    """Diagonal phase (-1)^((nele - 1) nbeta) of each basis state of
    ``set_basis``: the reference's FqeDataSet determinants are the
    Jordan-Wigner ones times this phase (its spin-flip maps count occupied
    orbitals above the orbital, reference ``src/fqe/fci_graph_set.py:
    137-150``).  An operator M here reads P M P there."""
    _, blocks = set_basis(norb, nele)
    return numpy.concatenate([
        numpy.full(lena * lenb, -1.0 if ((nele - 1) * key[1]) % 2 else 1.0)
        for key, _, lena, lenb in blocks])


def set_expectation(vec: numpy.ndarray, ops, norb: int, nele: int) -> complex:
    """<vec| op string |vec> over the combined basis of ``set_basis``;
    ``ops`` is a sequence of (mode, is_creation), applied right to left."""
    index, _ = set_basis(norb, nele)
    total = 0.0j
    for state, col in index.items():
        res = _apply_string(state, ops)
        if res is None:
            continue
        new, sign = res
        total += numpy.conj(vec[index[new]]) * sign * vec[col]
    return total
