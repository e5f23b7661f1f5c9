"""Learned correction of the Davidson diagonal preconditioner.

The plain Davidson update (reference ``src/fqe/algorithm/davidson.py:
282-291``) scales the residual by 1/(theta - H_II).  Here a small MLP
(10 -> 128 -> 128 -> 1, tanh; ``fqehip_mlp_*`` kernels) adds a
per-determinant correction c_I, so the update is
t_I = -(1/(H_II - theta) + c_I) r_I (``fqehip_davidson_precondition``).

Training data come from the method itself: small Hamiltonians are solved
to a tight residual with the plain preconditioner, and at every recorded
iteration (theta, u, r) the exact correction t* = psi/<u|psi> - u is
known; the least-squares best diagonal multiplier of r towards t* is
m_I = -Re(conj(r_I) t*_I)/|r_I|^2, and the network learns
m_I - 1/(H_II - theta) with weight |r_I|^2 (so the loss is the squared
distance of the preconditioned residual from t* over the sampled
determinants).  Training runs on the GPU with the project's forward,
back-propagation and Adam kernels; the plain diagonal preconditioner
remains the fallback (``correction=None``) and the residual norm remains
the convergence test.
"""

from __future__ import annotations

import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy
import torch

from .. import _lib
from ..fci_graph import FciGraph
from . import davidson as dv

NFEAT = 10
WIDTH = 128
NPARAMS = WIDTH * NFEAT + WIDTH + WIDTH * WIDTH + WIDTH + WIDTH + 1


def init_params(seed: int) -> numpy.ndarray:
    """Glorot-uniform weights (the output layer scaled by 0.1), zero
    biases, in the kernels' parameter order (float32)."""
    rng = numpy.random.default_rng(seed)
    lim1 = numpy.sqrt(6.0 / (WIDTH + NFEAT))
    lim2 = numpy.sqrt(6.0 / (2 * WIDTH))
    lim3 = numpy.sqrt(6.0 / (WIDTH + 1))
    parts = [rng.uniform(-lim1, lim1, WIDTH * NFEAT), numpy.zeros(WIDTH),
             rng.uniform(-lim2, lim2, WIDTH * WIDTH), numpy.zeros(WIDTH),
             0.1 * rng.uniform(-lim3, lim3, WIDTH), numpy.zeros(1)]
    return numpy.concatenate(parts).astype(numpy.float32)


def reference_strings(diag: torch.Tensor, graph: FciGraph) -> Tuple[int, int]:
    """Alpha and beta strings of the lowest-diagonal determinant."""
    idx = int(torch.argmin(diag.reshape(-1)).item())
    ia, ib = divmod(idx, graph.lenb())
    return int(graph.string_alpha(ia)), int(graph.string_beta(ib))


class SectorFeatures:
    """Everything the feature kernel needs for one Hamiltonian / sector."""

    def __init__(self, hamil, graph: FciGraph,
                 diag: Optional[torch.Tensor] = None) -> None:
        self.graph = graph
        self.diag = (dv.diagonal(hamil, graph) if diag is None else diag
                     ).reshape(-1).contiguous()
        self.dmin = float(torch.min(self.diag).item())
        self.ref_a, self.ref_b = reference_strings(self.diag, graph)
        self.n = graph.lena() * graph.lenb()

    def args(self, theta: float):
        g = self.graph
        return (g.norb(), g.lena(), g.lenb(), _lib.ptr(g.tables(True).strings),
                _lib.ptr(g.tables(False).strings), _lib.ptr(self.diag),
                float(theta), self.dmin, self.ref_a, self.ref_b)

    def features(self, theta: float, idx: torch.Tensor) -> torch.Tensor:
        m = idx.numel()
        x = torch.empty((m, NFEAT), dtype=torch.float32, device=idx.device)
        _lib.call("fqehip_mlp_features", *self.args(theta), m, _lib.ptr(idx),
                  _lib.ptr(x), _lib.stream_ptr())
        return x


class MLPPreconditioner:
    """Parameters on the device plus inference / training drivers."""

    def __init__(self, params: Optional[numpy.ndarray] = None,
                 seed: int = 0) -> None:
        p = init_params(seed) if params is None else numpy.asarray(
            params, dtype=numpy.float32)
        assert p.size == NPARAMS
        self.params = torch.from_numpy(p.copy()).to(_lib.device())
        self.history: List[float] = []

    # -- inference ---------------------------------------------------------
    def correction_for(self, feats: SectorFeatures):
        """Callable(theta, u, r) -> FP32 [n] correction, for davidson()."""
        xwork = torch.empty((feats.n, NFEAT), dtype=torch.float32,
                            device=self.params.device)
        corr = torch.empty(feats.n, dtype=torch.float32,
                           device=self.params.device)

        def fn(theta, u, r):
            _lib.call("fqehip_mlp_correction", _lib.ptr(self.params),
                      *feats.args(theta), _lib.ptr(xwork), _lib.ptr(corr),
                      _lib.stream_ptr())
            return corr
        return fn

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        _lib.call("fqehip_mlp_forward", _lib.ptr(self.params), x.shape[0],
                  _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr())
        return y

    # -- training ----------------------------------------------------------
    def loss_grad(self, x, t, w, work=None):
        m = x.shape[0]
        if work is None:
            work = torch.empty(int(_lib.call(
                "fqehip_mlp_train_workspace_floats", m)),
                dtype=torch.float32, device=x.device)
        g = torch.empty(NPARAMS + 1, dtype=torch.float32, device=x.device)
        _lib.call("fqehip_mlp_loss_grad", _lib.ptr(self.params), m,
                  _lib.ptr(x), _lib.ptr(t), _lib.ptr(w), _lib.ptr(g),
                  _lib.ptr(work), _lib.stream_ptr())
        return g

    def train(self, x: torch.Tensor, t: torch.Tensor, w: torch.Tensor,
              epochs: int = 30, batch: int = 4096, lr: float = 1e-3,
              seed: int = 0) -> List[float]:
        """Adam on mini-batches of the weighted squared error; the batch
        weights are renormalised to sum 1.  Returns the epoch losses."""
        dev = x.device
        m = x.shape[0]
        mo = torch.zeros(NPARAMS, dtype=torch.float32, device=dev)
        vo = torch.zeros_like(mo)
        work = torch.empty(int(_lib.call("fqehip_mlp_train_workspace_floats",
                                         batch)), dtype=torch.float32,
                           device=dev)
        gen = torch.Generator(device="cpu").manual_seed(seed)
        step = 0
        losses = []
        for _ in range(epochs):
            perm = torch.randperm(m, generator=gen).to(dev)
            total = 0.0
            for b0 in range(0, m - batch + 1, batch):
                sel = perm[b0:b0 + batch]
                xb = x[sel].contiguous()
                tb = t[sel].contiguous()
                wb = w[sel]
                wb = (wb / wb.sum().clamp_min(1e-30)).contiguous()
                g = self.loss_grad(xb, tb, wb, work)
                step += 1
                _lib.call("fqehip_mlp_adam", _lib.ptr(self.params),
                          _lib.ptr(g), _lib.ptr(mo), _lib.ptr(vo), step,
                          float(lr), 0.9, 0.999, 1e-8, _lib.stream_ptr())
                total += float(g[NPARAMS].item())
            losses.append(total / max(1, m // batch))
        self.history += losses
        return losses

    def numpy_params(self) -> numpy.ndarray:
        return self.params.cpu().numpy()


# ---------------------------------------------------------------------------
# training data from the method's own solves
# ---------------------------------------------------------------------------

def correction_samples(hamil, n_alpha: int, n_beta: int,
                       per_snapshot: int = 512, tol: float = 1e-10,
                       max_iter: int = 200, seed: int = 0):
    """(features, target, weight) device tensors from one Hamiltonian:
    a plain-preconditioned Davidson run to ``tol`` is recorded, psi is its
    final vector, and each recorded iteration contributes the
    ``per_snapshot`` determinants of largest |r_I| plus as many drawn
    uniformly at random."""
    norb = hamil.dim()
    graph = FciGraph(n_alpha, n_beta, norb)
    feats = SectorFeatures(hamil, graph)
    snaps = []

    def record(it, theta, u, r):
        snaps.append((theta, u.clone(), r.clone()))

    res = dv.davidson(hamil, n_alpha, n_beta, tol=tol, max_iter=max_iter,
                      recorder=record)
    psi = res.vectors[0]
    dev = psi.device
    gen = torch.Generator(device="cpu").manual_seed(seed)
    xs, ts, ws = [], [], []
    n = feats.n
    for theta, u, r in snaps[:-1]:
        ovl = torch.vdot(u, psi)
        if abs(ovl.item()) < 1e-12:
            continue
        tstar = psi / ovl - u
        r2 = (r.real ** 2 + r.imag ** 2)
        if float(r2.sum().item()) <= 0.0:
            continue
        m = -(r.conj() * tstar).real / r2.clamp_min(1e-300)
        den = feats.diag - theta
        den = torch.where(den.abs() < 1e-8, torch.full_like(den, 1e-8), den)
        tgt = (m - 1.0 / den).clamp(-5.0, 5.0)
        k = min(per_snapshot, n)
        top = torch.topk(r2, k).indices
        rnd = torch.randint(0, n, (k,), generator=gen).to(dev)
        idx = torch.cat([top, rnd])
        xs.append(feats.features(theta, idx))
        ts.append(tgt[idx].float())
        wv = r2[idx] / r2.sum()
        ws.append(wv.float())
    if not xs:
        return None
    return torch.cat(xs), torch.cat(ts), torch.cat(ws), res.iterations


def solve_both(hamil, n_alpha: int, n_beta: int,
               model: Optional[MLPPreconditioner], tol: float = 1e-8,
               max_iter: int = 300, dtype=torch.complex128) -> Dict[str, object]:
    """Davidson with the plain and with the learned preconditioner on the
    same Hamiltonian and start vector."""
    graph = FciGraph(n_alpha, n_beta, hamil.dim())
    feats = SectorFeatures(hamil, graph)
    t0 = time.perf_counter()
    plain = dv.davidson(hamil, n_alpha, n_beta, tol=tol, max_iter=max_iter,
                        dtype=dtype, diag=feats.diag)
    t1 = time.perf_counter()
    learned = None
    if model is not None:
        learned = dv.davidson(hamil, n_alpha, n_beta, tol=tol,
                              max_iter=max_iter, dtype=dtype, diag=feats.diag,
                              correction=model.correction_for(feats))
    t2 = time.perf_counter()
    return {"plain": plain, "learned": learned, "t_plain": t1 - t0,
            "t_learned": t2 - t1}


def solve_mixed(hamil, n_alpha: int, n_beta: int, tol: float = 1e-8,
                max_iter: int = 300, diag: Optional[torch.Tensor] = None):
    """Plain-preconditioned Davidson with mixed-precision applies until the
    mixed floor, FP64 applies after (``davidson(dtype="mixed")``)."""
    return dv.davidson(hamil, n_alpha, n_beta, tol=tol, max_iter=max_iter,
                       dtype="mixed", diag=diag)
