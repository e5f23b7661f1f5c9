"""CPU oracle of the learned Davidson preconditioner (test
infrastructure: only tests and the bench's CPU leg import it; the product
path runs the HIP kernels of ``fqehip_mlp_*``).

There is no counterpart in the reference: its Davidson preconditioner is
the plain diagonal ``reciprocal(theta - diag) * residual``
(``src/fqe/algorithm/davidson.py:282-291``).  This module restates, in
float64 numpy, exactly what the kernels compute:

* ``features``: the 10 norb-independent inputs of a determinant I;
* ``forward``: y = w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3;
* ``loss_and_grad``: L = sum_s w_s (y_s - t_s)^2 and its gradient by
  back-propagation, pinned by central finite differences in
  ``tests/test_oracle.py`` (a dropped term or transposed weight fails it);
* ``targets``: the per-determinant multiplier of the exact correction
  vector, m_I = -Re(conj(r_I) t*_I) / |r_I|^2 with t* = psi/<u|psi> - u,
  the least-squares best diagonal scaling of the residual r towards the
  Jacobi-Davidson exact correction t*.

Parameter vector (float32 on the device): W1 [128][10], b1 [128],
W2 [128][128], b2 [128], w3 [128], b3 [1].
"""

from __future__ import annotations

import numpy

NFEAT = 10
WIDTH = 128
FEAT_SOFT = 0.05


def param_shapes():
    return [("w1", (WIDTH, NFEAT)), ("b1", (WIDTH,)), ("w2", (WIDTH, WIDTH)),
            ("b2", (WIDTH,)), ("w3", (WIDTH,)), ("b3", (1,))]


def nparams() -> int:
    return sum(int(numpy.prod(s)) for _, s in param_shapes())


def unpack(theta: numpy.ndarray):
    out, off = {}, 0
    for name, shape in param_shapes():
        size = int(numpy.prod(shape))
        out[name] = theta[off:off + size].reshape(shape)
        off += size
    return out


def pack(p) -> numpy.ndarray:
    return numpy.concatenate([numpy.asarray(p[name]).reshape(-1)
                              for name, _ in param_shapes()])


def init_params(seed: int) -> numpy.ndarray:
    """Glorot-uniform weights, zero biases (float32)."""
    rng = numpy.random.default_rng(seed)
    p = {}
    for name, shape in param_shapes():
        if name.startswith("w") and len(shape) == 2:
            lim = numpy.sqrt(6.0 / (shape[0] + shape[1]))
            p[name] = rng.uniform(-lim, lim, size=shape)
        elif name == "w3":
            lim = numpy.sqrt(6.0 / (WIDTH + 1))
            p[name] = rng.uniform(-lim, lim, size=shape) * 0.1
        else:
            p[name] = numpy.zeros(shape)
    return pack(p).astype(numpy.float32)


def _homo_lumo(ref: int, norb: int):
    occ = [i for i in range(norb) if (ref >> i) & 1]
    vir = [i for i in range(norb) if not (ref >> i) & 1]
    return (max(occ) if occ else -1), (min(vir) if vir else -1)


def features(astr, bstr, diag, theta, dmin, ref_a, ref_b, norb):
    """[lena*lenb, 10] features of every determinant (row-major Ia, Ib):
    d - theta, 1/(|d - theta| + 0.05), d - dmin, theta - dmin, alpha and
    beta excitation levels against the reference strings, occupations of
    the reference HOMO and LUMO of each spin."""
    astr = numpy.asarray(astr, dtype=numpy.int64)
    bstr = numpy.asarray(bstr, dtype=numpy.int64)
    d = numpy.asarray(diag, dtype=numpy.float64).reshape(len(astr), len(bstr))
    ha, la = _homo_lumo(int(ref_a), norb)
    hb, lb = _homo_lumo(int(ref_b), norb)

    def popc(v):
        return numpy.array([bin(int(s)).count("1") for s in v], dtype=numpy.float64)

    def bit(v, i):
        if i < 0:
            return numpy.zeros(len(v))
        return ((v >> i) & 1).astype(numpy.float64)

    exa = popc(astr & ~numpy.int64(ref_a))
    exb = popc(bstr & ~numpy.int64(ref_b))
    x = numpy.zeros((len(astr), len(bstr), NFEAT))
    delta = d - theta
    x[..., 0] = delta
    x[..., 1] = 1.0 / (numpy.abs(delta) + FEAT_SOFT)
    x[..., 2] = d - dmin
    x[..., 3] = theta - dmin
    x[..., 4] = exa[:, None]
    x[..., 5] = exb[None, :]
    x[..., 6] = bit(astr, ha)[:, None]
    x[..., 7] = bit(astr, la)[:, None]
    x[..., 8] = bit(bstr, hb)[None, :]
    x[..., 9] = bit(bstr, lb)[None, :]
    return x.reshape(-1, NFEAT)


def forward(theta_p: numpy.ndarray, x: numpy.ndarray, keep: bool = False):
    p = unpack(numpy.asarray(theta_p, dtype=numpy.float64))
    h1 = numpy.tanh(x @ p["w1"].T + p["b1"])
    h2 = numpy.tanh(h1 @ p["w2"].T + p["b2"])
    y = h2 @ p["w3"] + p["b3"][0]
    if keep:
        return y, h1, h2
    return y


def loss_and_grad(theta_p, x, target, weight):
    """L = sum_s w_s (y_s - t_s)^2 and dL/dtheta (float64)."""
    p = unpack(numpy.asarray(theta_p, dtype=numpy.float64))
    y, h1, h2 = forward(theta_p, x, keep=True)
    res = y - target
    loss = float(numpy.sum(weight * res * res))
    d3 = 2.0 * weight * res                       # dL/dy
    d2 = numpy.outer(d3, p["w3"]) * (1.0 - h2 * h2)  # dL/da2
    d1 = (d2 @ p["w2"]) * (1.0 - h1 * h1)            # dL/da1
    g = {"w1": d1.T @ x, "b1": d1.sum(0), "w2": d2.T @ h1, "b2": d2.sum(0),
         "w3": h2.T @ d3, "b3": numpy.array([d3.sum()])}
    return loss, pack(g)


def targets(u, r, psi, inv_delta, clip: float = 5.0):
    """(target correction, weight) per determinant: the network learns
    m_I - 1/(d_I - theta), weighted by |r_I|^2 / sum |r|^2."""
    u = numpy.asarray(u).reshape(-1)
    r = numpy.asarray(r).reshape(-1)
    psi = numpy.asarray(psi).reshape(-1)
    tstar = psi / numpy.vdot(u, psi) - u
    r2 = numpy.abs(r) ** 2
    m = numpy.where(r2 > 0, -numpy.real(numpy.conj(r) * tstar) /
                    numpy.maximum(r2, 1e-300), 0.0)
    tgt = numpy.clip(m - inv_delta, -clip, clip)
    w = r2 / max(r2.sum(), 1e-300)
    return tgt, w


def adam_step(theta_p, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta_p - lr * mh / (numpy.sqrt(vh) + eps), m, v
