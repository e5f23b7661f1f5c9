"""Davidson-Liu eigensolver on the GPU engine.

Mirrors the reference's ``fqe.algorithm.davidson``
(``src/fqe/algorithm/davidson.py:140-359``): ``davidsonliu_fqe`` and
``davidson_diagonalization`` keep their signatures, exceptions, guess
vectors and convergence rule (change of the lowest ``nroots`` Ritz values
below ``epsilon``).  The method differs where the reference is quadratic:

* the preconditioner's diagonal <I|H|I> comes from the Slater-Condon
  closed form in one kernel (``fqehip_diagonal``) instead of one
  expectation value per determinant (``:160-180``);
* the subspace matrix grows by one column per new vector, one apply each,
  instead of being rebuilt from k^2 expectation values every iteration
  (``:211-216``);
* classical Gram-Schmidt in FP64 on the device, a second pass when the
  DGKS criterion asks for it
  (``fqehip_block_dot`` / ``fqehip_block_combine``), the Rayleigh-Ritz
  eigenproblem of the k x k subspace matrix in FP64 on the host.

``davidson`` is the residual-converged variant (||H u - theta u|| below
``tol``, thick restart to the Ritz vector when the subspace is full) with a
pluggable per-determinant correction of the diagonal preconditioner (the
learned preconditioner of ``fqe_amd.algorithm.mlp_precond``).
"""

from __future__ import annotations

import copy
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy
import torch

from .. import _lib, engine
from ..fci_graph import FciGraph
from ..hamiltonians import Hamiltonian, RestrictedHamiltonian


def diagonal(hamil: RestrictedHamiltonian, graph: FciGraph) -> torch.Tensor:
    """<I|H|I> (FP64, [lena, lenb], device) of a restricted Hamiltonian,
    e_0 included."""
    tens = hamil.tensors()
    norb = graph.norb()
    h1 = numpy.asarray(tens[0]) if len(tens) > 0 else numpy.zeros((norb,) * 2)
    h2 = (numpy.asarray(tens[1]) if len(tens) > 1 else
          numpy.zeros((norb,) * 4))
    dev = _lib.device()
    h1d = torch.from_numpy(numpy.ascontiguousarray(
        numpy.diagonal(h1).real, dtype=numpy.float64)).to(dev)
    jm = torch.from_numpy(numpy.ascontiguousarray(
        numpy.einsum("ijij->ij", h2).real, dtype=numpy.float64)).to(dev)
    km = torch.from_numpy(numpy.ascontiguousarray(
        numpy.einsum("ijji->ij", h2).real, dtype=numpy.float64)).to(dev)
    out = torch.empty((graph.lena(), graph.lenb()), dtype=torch.float64,
                      device=dev)
    _lib.call("fqehip_diagonal", norb, graph.lena(), graph.lenb(),
              _lib.ptr(graph.tables(True).strings),
              _lib.ptr(graph.tables(False).strings), _lib.ptr(h1d),
              _lib.ptr(jm), _lib.ptr(km), _lib.ptr(out), _lib.stream_ptr())
    e0 = numpy.real(hamil.e_0())
    if e0 != 0.0:
        out += float(e0)
    return out


class Subspace:
    """k orthonormal complex128 vectors V and their images AV on the
    device, with the k x k projected matrix G = V^H A V."""

    def __init__(self, n: int, maxk: int, device: torch.device) -> None:
        self.n = n
        self.maxk = maxk
        self.v = torch.empty((maxk, n), dtype=torch.complex128, device=device)
        self.av = torch.empty_like(self.v)
        self.g = numpy.zeros((maxk, maxk), dtype=numpy.complex128)
        self.k = 0
        self._ws = torch.empty(
            max(int(_lib.call("fqehip_block_dot_workspace_bytes", maxk)), 16),
            dtype=torch.uint8, device=device)
        self._dots = torch.empty(maxk, dtype=torch.complex128, device=device)
        self._coef = torch.empty(maxk, dtype=torch.complex128, device=device)

    def dots(self, x: torch.Tensor, k: Optional[int] = None,
             rows: Optional[torch.Tensor] = None) -> torch.Tensor:
        """V[:k]^H x on the device (complex128 [k])."""
        k = self.k if k is None else k
        base = self.v if rows is None else rows
        _lib.call("fqehip_block_dot", self.n, k, _lib.ptr(base), self.n,
                  _lib.ptr(x), _lib.ptr(self._dots), _lib.ptr(self._ws),
                  _lib.stream_ptr())
        return self._dots[:k]

    def combine(self, coef, x: torch.Tensor, beta: complex = 0.0,
                rows: Optional[torch.Tensor] = None) -> None:
        """x = beta x + sum_j coef_j rows_j (rows = V by default)."""
        k = len(coef)
        if isinstance(coef, torch.Tensor):
            c = coef
        else:
            self._coef[:k].copy_(torch.from_numpy(
                numpy.asarray(coef, dtype=numpy.complex128)))
            c = self._coef
        base = self.v if rows is None else rows
        _lib.call("fqehip_block_combine", self.n, k, _lib.ptr(base), self.n,
                  _lib.ptr(c), float(numpy.real(beta)),
                  float(numpy.imag(beta)), _lib.ptr(x), _lib.stream_ptr())

    def norm(self, x: torch.Tensor) -> float:
        _lib.call("fqehip_block_dot", self.n, 1, _lib.ptr(x), self.n,
                  _lib.ptr(x), _lib.ptr(self._dots), _lib.ptr(self._ws),
                  _lib.stream_ptr())
        return float(numpy.sqrt(max(self._dots[0].real.item(), 0.0)))

    def orthonormalize(self, x: torch.Tensor) -> float:
        """Classical Gram-Schmidt against V[:k] with the DGKS criterion
        (Daniel, Gragg, Kaufman and Stewart, Math. Comp. 30 (1976)): a
        second pass only when the first removed more than 1 - 1/sqrt(2) of
        the norm, then normalisation; returns the norm before
        normalisation."""
        if self.k:
            before = self.norm(x)
            c = self.dots(x).clone()
            self.combine(-c, x, beta=1.0)
            nrm = self.norm(x)
            if nrm < 0.7071067811865476 * before:
                c = self.dots(x).clone()
                self.combine(-c, x, beta=1.0)
                nrm = self.norm(x)
        else:
            nrm = self.norm(x)
        if nrm > 0.0:
            x.mul_(1.0 / nrm)
        return nrm

    def append(self, x: torch.Tensor, apply_fn) -> None:
        """Add the orthonormal vector x, its image and G's new column."""
        j = self.k
        self.v[j].copy_(x)
        apply_fn(self.v[j], self.av[j])
        self.k = j + 1
        col = self.dots(self.av[j]).cpu().numpy()
        self.g[:j + 1, j] = col
        self.g[j, :j + 1] = numpy.conj(col)
        self.g[j, j] = col[j].real

    def ritz(self, nroots: int):
        w, s = numpy.linalg.eigh(self.g[:self.k, :self.k])
        return w[:nroots], s[:, :nroots]

    def restart(self, coef: numpy.ndarray) -> None:
        """Replace the basis by V coef (coef: k x m with orthonormal
        columns); the images follow as AV coef and G as coef^H G coef, so
        no apply is needed."""
        k, m = coef.shape
        assert k == self.k and m <= self.maxk
        dev = self.v.device
        if getattr(self, "_tmp", None) is None or self._tmp.shape[1] < m:
            self._tmp = torch.empty((2, m, self.n), dtype=torch.complex128,
                                    device=dev)
        tmp = self._tmp
        if m <= 16:
            # all m new vectors (and images) in one pass over the rows
            cm = torch.from_numpy(numpy.ascontiguousarray(
                coef, dtype=numpy.complex128)).to(dev)
            for t, rows in ((tmp[0], self.v), (tmp[1], self.av)):
                _lib.call("fqehip_block_combine_multi", self.n, k,
                          _lib.ptr(rows), self.n, _lib.ptr(cm), m,
                          _lib.ptr(t), self.n, _lib.stream_ptr())
        else:
            for j in range(m):
                c = numpy.ascontiguousarray(coef[:, j])
                self.combine(c, tmp[0, j])
                self.combine(c, tmp[1, j], rows=self.av)
        self.v[:m].copy_(tmp[0, :m])
        self.av[:m].copy_(tmp[1, :m])
        g = coef.conj().T @ self.g[:k, :k] @ coef
        self.g[:] = 0.0
        self.g[:m, :m] = 0.5 * (g + g.conj().T)
        self.k = m


class _Applier:
    """H x for device vectors of one sector (engine apply, e_0 added)."""

    def __init__(self, hamil: RestrictedHamiltonian, graph: FciGraph,
                 dtype: torch.dtype) -> None:
        from ..wavefunction import _prepared
        self.graph = graph
        self.op = _prepared(hamil, hamil.tensors(), graph.norb())
        self.e0 = complex(hamil.e_0())
        self.dtype = dtype
        self.prec = engine.precision_of(dtype)
        self.s, self.keep = engine.sector_struct(graph, self.op, self.prec)
        self.ws = (engine.workspace(self.s, self.prec, _lib.device())
                   if self.op.h1 is None else None)
        self.napply = 0

    def __call__(self, x: torch.Tensor, out: torch.Tensor) -> None:
        shape = (self.graph.lena(), self.graph.lenb())
        if self.dtype == torch.complex128:
            src = x.view(shape)
            dst = out.view(shape)
        else:
            src = x.view(shape).to(self.dtype)
            dst = torch.empty_like(src)
        cin = src if self.e0 != 0 else None
        engine.apply_sector(self.graph, src, self.op, dst, 1.0, self.e0, cin,
                            ws=self.ws)
        if dst.data_ptr() != out.data_ptr():
            out.view(shape).copy_(dst)
        self.napply += 1


@dataclass
class DavidsonResult:
    eigenvalues: numpy.ndarray
    vectors: torch.Tensor            # [nroots, lena*lenb] complex128
    iterations: int
    applies: int
    residuals: List[float] = field(default_factory=list)
    converged: bool = False
    seconds: float = 0.0
    switch_iteration: int = 0        # mixed -> FP64 (dtype="mixed")


Correction = Callable[[float, torch.Tensor, torch.Tensor], Optional[torch.Tensor]]


def davidson(hamil: RestrictedHamiltonian, n_alpha: int, n_beta: int,
             tol: float = 1e-8, max_iter: int = 300, max_subspace: int = 40,
             guess: Optional[torch.Tensor] = None,
             correction: Optional[Correction] = None,
             dtype: torch.dtype = torch.complex128,
             eps: float = 1e-8, verbose: bool = False,
             diag: Optional[torch.Tensor] = None,
             recorder: Optional[Callable] = None,
             seed: int = 0, keep: int = 8,
             mixed_tol: float = 2e-5) -> DavidsonResult:
    """Lowest eigenpair to ||H u - theta u|| < tol.

    ``correction(theta, u, r)`` may return an FP32 [n] device tensor added
    to 1/(H_dd - theta) in the preconditioner; None keeps the plain
    diagonal preconditioner.  ``dtype`` selects the engine of the applies
    (complex128: FP64; complex64: mixed precision, whose apply error bounds
    the reachable residual; "mixed": mixed-precision applies until the
    residual falls below ``mixed_tol`` * max(1, |theta|), or stalls within
    50 times that (the mixed apply's error, about 4e-6 of ||H u||, is the
    floor), then a
    thick restart whose
    kept vectors get FP64 images, and FP64 applies to the end).  ``diag`` reuses a computed <I|H|I>;
    ``recorder(it, theta, u, r)`` sees every iteration's Ritz pair and
    residual (training data of the learned preconditioner).  A full
    subspace restarts thickly: the ``keep`` lowest Ritz vectors plus the
    previous iteration's Ritz vector (orthogonalised), images and projected
    matrix carried over without applies.  Without a
    guess the start vector is the lowest-diagonal determinant plus 1e-3 of
    a seeded random vector, so that no symmetry sector of the sector's
    eigenvectors is missing from it."""
    norb = hamil.dim()
    graph = FciGraph(n_alpha, n_beta, norb)
    n = graph.lena() * graph.lenb()
    dev = _lib.device()
    t0 = time.perf_counter()
    diag = (diagonal(hamil, graph) if diag is None else diag).reshape(-1)
    two_phase = dtype == "mixed"
    apply_fn = _Applier(hamil, graph,
                        torch.complex64 if two_phase else dtype)
    applies = 0
    switch_it = 0
    maxk = max(2, min(max_subspace, n))
    sub = Subspace(n, maxk, dev)
    x = torch.zeros(n, dtype=torch.complex128, device=dev)
    if guess is not None:
        x.copy_(guess.reshape(-1).to(torch.complex128))
    else:
        rng = numpy.random.default_rng(seed)
        noise = rng.normal(size=n) + 1j * rng.normal(size=n)
        noise *= 1e-3 / numpy.linalg.norm(noise)
        x.copy_(torch.from_numpy(noise))
        x[int(torch.argmin(diag).item())] += 1.0
    sub.orthonormalize(x)
    sub.append(x, apply_fn)
    u = torch.empty_like(x)
    au = torch.empty_like(x)
    r = torch.empty_like(x)
    res = []
    theta = 0.0
    converged = False
    it = 0
    prev = None                      # previous Ritz coefficients
    keep = max(1, min(keep, maxk - 2))
    for it in range(1, max_iter + 1):
        w, s = sub.ritz(1)
        theta = float(w[0])
        coef = s[:, 0]
        sub.combine(coef, u)
        sub.combine(coef, au, rows=sub.av)
        r.copy_(au)
        r.add_(u, alpha=-theta)
        rn = sub.norm(r)
        res.append(rn)
        if verbose:
            print(f"davidson it {it} theta {theta:.12f} |r| {rn:.3e}")
        if recorder is not None:
            recorder(it, theta, u, r)
        floor = mixed_tol * max(1.0, abs(theta))
        if two_phase and (rn < floor or
                          (rn < 50.0 * floor and len(res) > 30 and
                           min(res[-15:]) > 0.5 * min(res[-30:-15]))):
            # leave the mixed phase: thick restart on the lowest Ritz
            # vectors with their images recomputed by the FP64 engine
            two_phase = False
            switch_it = it
            applies += apply_fn.napply
            apply_fn = _Applier(hamil, graph, torch.complex128)
            _, allv = numpy.linalg.eigh(sub.g[:sub.k, :sub.k])
            m = min(keep, sub.k)
            sub.restart(allv[:, :m])
            for j in range(m):
                apply_fn(sub.v[j], sub.av[j])
            col = numpy.stack([sub.dots(sub.av[j], k=m).cpu().numpy()
                               for j in range(m)], axis=1)
            sub.g[:] = 0.0
            sub.g[:m, :m] = 0.5 * (col + col.conj().T)
            prev = None
            continue
        if rn < tol or sub.k >= n:
            converged = rn < tol
            break
        corr = correction(theta, u, r) if correction is not None else None
        _lib.call("fqehip_davidson_precondition", n, _lib.ptr(r),
                  _lib.ptr(diag), theta, eps, _lib.ptr(corr), _lib.ptr(x),
                  _lib.stream_ptr())
        if sub.k == maxk:
            _, allv = numpy.linalg.eigh(sub.g[:sub.k, :sub.k])
            cols = [allv[:, j] for j in range(keep)]
            if prev is not None:
                pv = numpy.zeros(sub.k, dtype=numpy.complex128)
                pv[:len(prev)] = prev
                for _ in range(2):
                    for c in cols:
                        pv -= c * numpy.vdot(c, pv)
                nrm_p = numpy.linalg.norm(pv)
                # a nearly dependent previous vector would only add
                # rounding noise to the basis (a residual floor)
                if nrm_p > 1e-3:
                    cols.append(pv / nrm_p)
            sub.restart(numpy.stack(cols, axis=1))
            prev = None
        else:
            prev = coef.copy()
        nrm = sub.orthonormalize(x)
        if nrm < 1e-14:
            # the correction lies in the subspace: perturb with the residual
            x.copy_(r)
            sub.orthonormalize(x)
        sub.append(x, apply_fn)
    torch.cuda.synchronize()
    return DavidsonResult(numpy.array([theta]), u.reshape(1, -1).clone(), it,
                          applies + apply_fn.napply, res, converged,
                          time.perf_counter() - t0, switch_it)


# ---------------------------------------------------------------------------
# reference API
# ---------------------------------------------------------------------------

def davidsonliu_fqe(hmat: Hamiltonian, nroots: int, guess_vecs, nele, sz,
                    norb, epsilon: float = 1.0e-8, verbose=False):
    """Lowest ``nroots`` eigenpairs (reference
    ``src/fqe/algorithm/davidson.py:140-307``): returns (eigenvalues,
    list of Wavefunction)."""
    if nroots < 1 or nroots > 2**(hmat.dim() - 1):
        raise ValueError("Number of roots is incorrectly specified")
    gv_sector = list(guess_vecs[0].sectors())[0]
    for gv in guess_vecs:
        if list(gv.sectors())[0] != gv_sector:
            raise TypeError("Sectors don't match for guess vectors")
    sec0 = guess_vecs[0].sector(gv_sector)
    graph = sec0.get_fcigraph()
    n = graph.lena() * graph.lenb()
    dev = _lib.device()
    diag = diagonal(hmat, graph).reshape(-1)
    apply_fn = _Applier(hmat, graph, torch.complex128)
    maxk = max(len(guess_vecs), n // 2 + nroots + 1)
    maxk = min(maxk, n)
    sub = Subspace(n, maxk, dev)
    x = torch.empty(n, dtype=torch.complex128, device=dev)
    for gv in guess_vecs:
        x.copy_(gv.sector(gv_sector).device_coeff().reshape(-1).to(
            torch.complex128))
        if sub.k < maxk and sub.orthonormalize(x) > 1e-14:
            sub.append(x, apply_fn)
    old = numpy.array([numpy.inf] * nroots)
    u = torch.empty_like(x)
    au = torch.empty_like(x)
    while True:
        k = sub.k
        w, s = sub.ritz(min(nroots, k))
        if verbose:
            print("subspace", k, "thetas", w)
        if len(w) == nroots and numpy.linalg.norm(w - old) < epsilon:
            break
        if k > n / 2 or k >= maxk:
            break
        old = w if len(w) == nroots else old
        grew = False
        for i in range(len(w)):
            sub.combine(s[:, i], u)
            sub.combine(s[:, i], au, rows=sub.av)
            au.add_(u, alpha=-float(w[i]))          # residual
            _lib.call("fqehip_davidson_precondition", n, _lib.ptr(au),
                      _lib.ptr(diag), float(w[i]), 1e-12, None, _lib.ptr(x),
                      _lib.stream_ptr())
            if sub.k < maxk and sub.orthonormalize(x) > 1e-14:
                sub.append(x, apply_fn)
                grew = True
        if not grew:
            break
    w, s = sub.ritz(min(nroots, sub.k))
    eigfuncs = []
    for i in range(len(w)):
        sub.combine(s[:, i], u)
        wfn = copy.deepcopy(guess_vecs[0])
        wfn.sector(gv_sector).set_device_coeff(
            u.reshape(graph.lena(), graph.lenb()).to(
                wfn.sector(gv_sector).device_coeff().dtype).clone())
        eigfuncs.append(wfn)
    return w, eigfuncs


def davidson_diagonalization(hamiltonian: RestrictedHamiltonian,
                             n_alpha: int, n_beta: int, nroots=1,
                             guess_vecs=None):
    """Reference ``src/fqe/algorithm/davidson.py:310-359``: guesses are the
    Hartree-Fock determinant, the determinant with every electron one
    orbital up, and a random normalised wavefunction."""
    from .. import wavefunction as W
    norb = hamiltonian.dim()
    nele = n_alpha + n_beta
    sz = n_alpha - n_beta
    wfn = W.Wavefunction([[nele, sz, norb]])
    graph = wfn.sector((nele, sz)).get_fcigraph()
    if guess_vecs is None:
        g1 = numpy.zeros((graph.lena(), graph.lenb()), dtype=numpy.complex128)
        g2 = numpy.zeros_like(g1)
        alpha_hf = (1 << n_alpha) - 1
        beta_hf = (1 << n_beta) - 1
        g1[graph.index_alpha(alpha_hf), graph.index_beta(beta_hf)] = 1.0
        g2[graph.index_alpha(alpha_hf << 1), graph.index_beta(beta_hf << 1)] = 1.0
        w1 = copy.deepcopy(wfn)
        w2 = copy.deepcopy(wfn)
        w1.set_wfn(strategy="from_data", raw_data={(nele, sz): g1})
        w2.set_wfn(strategy="from_data", raw_data={(nele, sz): g2})
        rnd = W.Wavefunction([[nele, sz, norb]])
        rnd.set_wfn(strategy="random")
        rnd.normalize()
        guess_vecs = [w1, w2, rnd]
    return davidsonliu_fqe(hamiltonian, nroots, guess_vecs, nele=nele, sz=sz,
                           norb=norb)
