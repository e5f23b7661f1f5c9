"""Chebyshev propagator of ``Wavefunction.apply_generated_unitary``
(reference ``src/fqe/wavefunction.py:570-602``).

exp(-i t H) psi = exp(-i t E) [J_0(a t) T_0 + 2 sum_k (-i)^k J_k(a t) T_k]
psi with the spectrum mapped into [-wprime, wprime] by
H' = (H + E) / a, a = (spec_lim[1] - spec_lim[0]) / (2 wprime),
E = -(spec_lim[0] + a wprime), wprime = 0.9875, and the three-term
recurrence T_{k+1} = 2 H' T_k - T_{k-1}.  Every H application is the
engine's apply (e_0 included, as the reference's ``Wavefunction.apply``
does); the caller applies the final e_0 phase.  The series stops when
|coefficient| * ||T_k psi|| drops below ``accuracy``, else raises
RuntimeError after ``expansion - 1`` terms, as the reference does.
"""

from __future__ import annotations

import copy
from typing import Optional, Sequence

import numpy
from scipy.special import jv


def chebyshev_evolve(wfn, time: float, hamil, accuracy: float,
                     expansion: int, spec_lim: Optional[Sequence[float]]):
    assert spec_lim, ('Spectral range was not provided.'
                      ' Provide upper and lower limits.')
    wprime = 0.9875
    ascale = (spec_lim[1] - spec_lim[0]) / (2.0 * wprime)
    eshift = -(spec_lim[0] + ascale * wprime)

    time_evol = copy.deepcopy(wfn)
    time_evol.scale(jv(0, ascale * time))
    minus = copy.deepcopy(wfn)
    current = minus.apply(hamil)
    current.ax_plus_y(eshift, minus)
    current.scale(1.0 / ascale)
    time_evol.ax_plus_y(2.0 * jv(1, ascale * time) * (-1.j), current)
    for order in range(2, expansion):
        minus.scale(-1.0)
        minus.ax_plus_y(2.0 / ascale, current.apply(hamil))
        minus.ax_plus_y(2.0 * eshift / ascale, current)
        current, minus = minus, current
        coeff = 2.0 * jv(order, ascale * time) * (-1.j)**order
        time_evol.ax_plus_y(coeff, current)
        if current.norm() * numpy.abs(coeff) < accuracy:
            break
    else:
        raise RuntimeError("maximum chebyshev expansion limit reached")
    time_evol.scale(numpy.exp(eshift * time * 1.j))
    return time_evol
