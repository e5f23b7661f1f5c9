"""Helpers of the apply path (reference ``src/fqe/util.py``)."""

from __future__ import annotations

from typing import Tuple

import numpy
import torch


def alpha_beta_electrons(nele: int, m_s: int) -> Tuple[int, int]:
    """reference ``src/fqe/util.py:29-50``"""
    if nele < 0:
        raise ValueError('Cannot have negative electrons')
    if nele < abs(m_s):
        raise ValueError('Spin quantum number exceeds physical limits')
    if (nele + m_s) % 2 != 0:
        raise ValueError('Parity of spin quantum number and number of '
                         'electrons is incompatible')
    nalpha = int(nele + m_s) // 2
    return nalpha, nele - nalpha


def validate_config(nalpha: int, nbeta: int, norb: int) -> None:
    """reference ``src/fqe/util.py:440-463``"""
    if nalpha < 0:
        raise ValueError("Cannot have negative number of alpha electrons")
    if nbeta < 0:
        raise ValueError("Cannot have negative number of beta electrons")
    if norb < 0:
        raise ValueError("Cannot have negative number of orbitals")
    if norb < nalpha or norb < nbeta:
        raise ValueError("Insufficient number of orbitals")


def validate_tuple(matrices) -> None:
    """reference ``src/fqe/util.py:466-477``"""
    assert isinstance(matrices, tuple)
    for rank, term in enumerate(matrices):
        assert isinstance(term, numpy.ndarray)
        assert 2 * (rank + 1) == term.ndim


def rand_wfn(adim: int, bdim: int) -> numpy.ndarray:
    """Unnormalised random coefficients (reference ``src/fqe/util.py:383-397``)."""
    return (numpy.random.randn(adim, bdim).astype(numpy.complex128) +
            numpy.random.randn(adim, bdim).astype(numpy.complex128) * 1.j)


def vdot(wfn1, wfn2) -> complex:
    """sum over common sectors of <wfn1|wfn2> (reference
    ``src/fqe/util.py:506-530``), on the device in complex128."""
    from . import engine
    keys = [k for k in wfn1.sectors() if k in wfn2.sectors()]
    total = 0.0 + 0.0j
    for key in keys:
        a = wfn1.sector(key).device_coeff().to(torch.complex128).contiguous()
        b = wfn2.sector(key).device_coeff().to(torch.complex128).contiguous()
        total += engine.dot128(a, b)
    return total
