"""Synthetic restricted Hamiltonians and wavefunctions of the benchmark
workloads (numpy, float64), with a fixed random-number convention so that
every run of a given seed sees the same inputs.

Per workload, from ``numpy.random.default_rng(seed)``, in this order:
1. coefficients: real parts, then imaginary parts, i.i.d. N(0, 1), then
   normalised;
2. h1: the strictly upper triangle N(0, 0.1) (mirrored), then the
   diagonal -2 + N(0, 0.5);
3. nfactors real symmetric matrices L_k (upper triangle incl. diagonal
   N(0, 0.1), mirrored), electron-repulsion integrals
   (ij|kl) = sum_k L_k[i,j] L_k[k,l].
The two-body tensor in FQE's convention is h2[i,j,k,l] = -1/2 (ik|jl)
(OpenFermion's physics tensor V[p,q,r,s] = (ps|qr) passed through the
reference's ``integrals_to_fqe_restricted``,
``src/fqe/openfermion_utils.py:390-402``).
"""

from __future__ import annotations

from math import comb

import numpy


def hamiltonian(norb: int, nfactors: int, rng: numpy.random.Generator):
    upper = numpy.triu(rng.normal(0.0, 0.1, size=(norb, norb)), 1)
    h1 = upper + upper.T
    h1[numpy.arange(norb), numpy.arange(norb)] = (
        -2.0 + rng.normal(0.0, 0.5, size=norb))
    eri = numpy.zeros((norb,) * 4)
    for _ in range(nfactors):
        lk = rng.normal(0.0, 0.1, size=(norb, norb))
        lk = numpy.triu(lk) + numpy.triu(lk, 1).T
        eri += lk[:, :, None, None] * lk[None, None, :, :]
    h2 = -0.5 * eri.transpose(0, 2, 1, 3)        # h2[i,j,k,l] = -(ik|jl)/2
    return h1, numpy.ascontiguousarray(h2)


def coefficients(lena: int, lenb: int,
                 rng: numpy.random.Generator) -> numpy.ndarray:
    re = rng.normal(size=(lena, lenb))
    im = rng.normal(size=(lena, lenb))
    c = re + 1.0j * im
    return c / numpy.linalg.norm(c)


def workload(norb: int, nalpha: int, nbeta: int, seed: int, nfactors: int):
    """(coefficients [lena, lenb] complex128, h1, h2) of one sector."""
    rng = numpy.random.default_rng(seed)
    c = coefficients(comb(norb, nalpha), comb(norb, nbeta), rng)
    h1, h2 = hamiltonian(norb, nfactors, rng)
    return c, h1, h2
