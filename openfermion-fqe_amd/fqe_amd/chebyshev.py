"""Chebyshev propagator of ``Wavefunction.apply_generated_unitary``
(reference ``src/fqe/wavefunction.py:570-602``), run as a device-resident
three-term recurrence.

Mathematics.  With the spectrum of K = H + e_0 (the operator the
reference's ``Wavefunction.apply`` applies) inside [lo, hi], let
c = (lo + hi) / 2 and a = (hi - lo) / (2 w), w = 0.9875, so that
K' = (K - c) / a has its spectrum in [-w, w].  The Jacobi-Anger expansion
gives

    exp(-i t K) psi = exp(-i c t) sum_k b_k T_k(K') psi,
    b_0 = J_0(a t),  b_k = 2 (-i)^k J_k(a t),

with T_0 psi = psi, T_1 psi = K' psi and T_k psi = 2 K' T_{k-1} psi -
T_{k-2} psi.  The series stops after the first term k >= 2 with
|b_k| ||T_k psi|| < accuracy and raises RuntimeError when no term up to
expansion - 1 meets that test, as the reference does; w and the test must
be the reference's to reproduce its numbers.

Device design.  Per term and sector, one apply writes
(2/a) H T_{k-1} - T_{k-2} (the apply's own alpha / beta epilogue), and one
streaming kernel (``fqehip_cheb_update``) adds the shift
(2 (e_0 - c) / a) T_{k-1}, accumulates b_k T_k into the complex128 series
and reduces ||T_k||^2 per block; ``fqehip_norm_finalize`` then compares the
norm over all sectors with accuracy / |b_k| and sets a device flag that
makes every later kernel return at once.  Three term buffers rotate per
sector; nothing returns to the host until the series ends.  The caller
applies the final e_0 phase.
"""

from __future__ import annotations

import copy
import math
from typing import Optional, Sequence

import numpy
import torch
from scipy.special import jv

from . import _lib
from . import engine

# half-width of the mapped spectrum (the reference's choice)
SPECTRAL_MARGIN = 0.9875


def series_coefficients(a: float, time: float, nterms: int) -> numpy.ndarray:
    """b_k = (2 - delta_k0) (-i)^k J_k(a t), k < nterms."""
    k = numpy.arange(nterms)
    b = jv(k, a * time) * (-1.0j) ** k
    b[1:] *= 2.0
    return b


def chebyshev_evolve(wfn, time: float, hamil, accuracy: float,
                     expansion: int, spec_lim: Optional[Sequence[float]]):
    assert spec_lim, ('Spectral range was not provided.'
                      ' Provide upper and lower limits.')
    lo, hi = float(spec_lim[0]), float(spec_lim[1])
    centre = 0.5 * (lo + hi)
    a = (hi - lo) / (2.0 * SPECTRAL_MARGIN)
    b = series_coefficients(a, time, max(expansion, 2))
    shift = complex(hamil.e_0()) - centre           # K - c = H + shift
    from .wavefunction import _prepared              # shared operator cache
    op = _prepared(hamil, hamil.tensors(), wfn.norb())
    dev = _lib.device()
    out = copy.deepcopy(wfn)
    secs = list(out._civec.values())

    bufs, results, structs, wss, offs = [], [], [], [], [0]
    for sec in secs:
        c = sec.device_coeff()
        bufs.append([c.clone(), torch.empty_like(c), torch.empty_like(c)])
        results.append(torch.zeros(c.shape + (2,), dtype=torch.float64,
                                   device=dev))
        s, keep = engine.sector_struct(sec.get_fcigraph(), op,
                                       engine.precision_of(c.dtype))
        structs.append((s, keep))
        wss.append(engine.workspace(s, engine.precision_of(c.dtype), dev)
                   if op.h1 is None else None)
        offs.append(offs[-1] + engine.norm_partials_count(c.numel()))
    partials = torch.zeros(max(offs[-1], 1), dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)
    norm = torch.zeros(1, dtype=torch.float64, device=dev)
    opc = op.c_struct()
    done = status[1:]

    def apply(k, x, alpha, beta, cin, y):
        prec = engine.precision_of(x.dtype)
        _lib.call("fqehip_apply", prec, engine._as_ref(structs[k][0]),
                  engine._as_ref(opc), _lib.ptr(x), alpha.real, alpha.imag,
                  beta.real, beta.imag, _lib.ptr(cin), _lib.ptr(y), None,
                  None, _lib.ptr(done), _lib.ptr(wss[k]), _lib.stream_ptr())

    def update(k, tk, tkm1, gamma, coef, with_norm):
        prec = engine.precision_of(tk.dtype)
        _lib.call("fqehip_cheb_update", prec, tk.numel(), _lib.ptr(tk),
                  _lib.ptr(tkm1), gamma.real, gamma.imag, coef.real,
                  coef.imag, _lib.ptr(results[k]),
                  _lib.ptr(partials[offs[k]:]) if with_norm else None,
                  _lib.ptr(done), _lib.stream_ptr())

    # T_0 = psi enters with b_0; T_1 = (H + shift) psi / a with b_1
    for k in range(len(secs)):
        t0, t1, _ = bufs[k]
        update(k, t0, None, 0j, complex(b[0]), False)
        apply(k, t0, complex(1.0 / a), shift / a, t0, t1)
        update(k, t1, None, 0j, complex(b[1]), False)
    gamma = 2.0 * shift / a
    for order in range(2, expansion):
        for k in range(len(secs)):
            tkm2, tkm1, tk = bufs[k]
            apply(k, tkm1, complex(2.0 / a), complex(-1.0), tkm2, tk)
            update(k, tk, tkm1, gamma, complex(b[order]), True)
            bufs[k] = [tkm1, tk, tkm2]
        bmag = abs(b[order])
        limit = accuracy / bmag if bmag > 0.0 else math.inf
        _lib.call("fqehip_norm_finalize", _lib.ptr(partials), offs[-1],
                  _lib.ptr(norm), limit, order, _lib.ptr(done),
                  _lib.ptr(status[:1]), _lib.stream_ptr())
    st = status.cpu().numpy()
    if expansion <= 2 or st[1] == 0:
        raise RuntimeError("maximum chebyshev expansion limit reached")
    phase = numpy.exp(-1.0j * centre * time)
    for k, sec in enumerate(secs):
        res = torch.view_as_complex(results[k]) * phase
        sec.set_device_coeff(res.to(sec.device_coeff().dtype))
    out._chebyshev_terms = int(st[0])
    return out
