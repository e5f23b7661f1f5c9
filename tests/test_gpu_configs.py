"""The configured workloads at their full sizes on the GPU.

* Batched uneven sectors (256 sectors, norb 6..12, nalpha / nbeta uniform
  in [1, norb - 1], seed 3; one apply + a 10-term Taylor series each): the
  FP64 graph against the per-sector sequential FP64 path to 1e-12, the
  norb <= 8 sectors against the oracle to 1e-12, and the complex64 graph
  against the FP64 path within its stated bounds (apply 1e-4, Taylor
  1e-3: BF16x3 products with FP32 accumulation, errors of about 5e-6 per
  apply compounding over ten terms), and the complex128 graph with every
  split-capable sector on the Ozaki-digit kernels against the FP64 path
  to 1e-9.
* Davidson at norb=14 7a7b (held-out seed-4 Hamiltonian, FP64 applies):
  residual 1e-8, recomputed from a fresh apply; the eigenvalue of a second
  solve from another start vector agrees to 1e-9.
* The production-size plain method (norb=12 6a6b, seed 1, 12 factors): 100
  chained FP64 applies (normalised) and time_evolve(0.05) against the
  reference's compiled C apply chain and its Taylor loop, to 1e-11, with
  the same term count and a unit norm.
"""

import math

import numpy
import pytest

from oracle import fqe_oracle as fo
from oracle import ref_lm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return float(numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300))


def _oracle_taylor(spec, t, nterms):
    g = fo.OracleGraph(spec.nalpha, spec.nbeta, spec.norb)
    c = spec.coeff.astype(numpy.complex128)
    out = fo.apply_12(c, spec.h1, spec.h2, g)
    term, res = c.copy(), c.copy()
    for k in range(1, nterms + 1):
        term = (-1j * t / k) * fo.apply_12(term, spec.h1, spec.h2, g)
        res = res + term
    return out, res


def test_batched_256_sectors_seed3(fqe):
    import torch
    from fqe_amd import batched
    specs = batched.random_specs(256, seed=3, norb_range=(6, 12))
    print("256 sectors, FP64 graph", flush=True)
    fp = batched.BatchedTaylor(specs, 0.1, nterms=10,
                               dtype=torch.complex128, nstreams=8)
    fp.build()
    fp.launch()
    graph_res = fp.results()
    fp.run_sequential()
    seq = fp.results()
    fp.close()
    del fp
    worst = 0.0
    for (ga, ge), (sa, se) in zip(graph_res, seq):
        worst = max(worst, rel(ga, sa), rel(ge, se))
    print(f"  FP64 graph vs sequential: {worst:.2e}", flush=True)
    assert worst < 1e-12
    small = [k for k, sp in enumerate(specs) if sp.norb <= 8]
    assert len(small) > 20
    for k in small:
        ra, re = _oracle_taylor(specs[k], 0.1, 10)
        assert rel(graph_res[k][0], ra) < 1e-12, k
        assert rel(graph_res[k][1], re) < 1e-12, k
    print(f"  {len(small)} sectors with norb <= 8 match the oracle",
          flush=True)
    mx = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=torch.complex64,
                               nstreams=8)
    mx.build()
    mx.launch()
    mres = mx.results()
    mx.close()
    ea = max(rel(a, b[0]) for (a, _), b in zip(mres, seq))
    ee = max(rel(e, b[1]) for (_, e), b in zip(mres, seq))
    print(f"  complex64 graph vs FP64: apply {ea:.2e}, Taylor {ee:.2e}",
          flush=True)
    assert ea < 1e-4 and ee < 1e-3
    # the mixed engine at <= 1e-9: complex128 graph, every split-capable
    # sector on the Ozaki-digit split kernels
    from fqe_amd import _lib
    tiles = {f"{sp.norb},{sp.nalpha},{sp.nbeta}": "split" for sp in specs}
    oz = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=torch.complex128,
                               nstreams=8, tiles=tiles, contraction="ozaki")
    ndigit = sum(1 for p in oz.sectors if p.s.tile == _lib.TILE_OZAKI)
    oz.build()
    oz.launch()
    ores = oz.results()
    oz.close()
    ea = max(rel(a, b[0]) for (a, _), b in zip(ores, seq))
    ee = max(rel(e, b[1]) for (_, e), b in zip(ores, seq))
    print(f"  Ozaki graph ({ndigit} digit sectors) vs FP64: apply {ea:.2e}, "
          f"Taylor {ee:.2e}", flush=True)
    assert ndigit > 100
    assert ea < 1e-9 and ee < 1e-9


def test_davidson_norb14_seed4(fqe):
    import torch
    from fqe_amd import _lib, engine, synthetic
    from fqe_amd.algorithm import davidson as dv
    rng = numpy.random.default_rng(4)
    h1, h2 = synthetic.hamiltonian(14, 14, rng)
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    print("norb=14 Davidson, start 1", flush=True)
    r1 = dv.davidson(ham, 7, 7, tol=1e-8, max_iter=800)
    assert r1.converged
    print(f"  {r1.iterations} iterations, theta {r1.eigenvalues[0]:.12f}",
          flush=True)
    graph = fqe.FciGraph(7, 7, 14)
    op = engine.prepare_operator((h1, h2), 14, _lib.device())
    u = r1.vectors[0].reshape(graph.lena(), graph.lenb()).contiguous()
    hu = torch.empty_like(u)
    engine.apply_sector(graph, u, op, hu)
    theta = float(r1.eigenvalues[0])
    res = torch.linalg.vector_norm(hu - theta * u).item()
    nu = torch.linalg.vector_norm(u).item()
    print(f"  fresh residual {res:.2e}, |u| = {nu:.15f}", flush=True)
    assert abs(nu - 1.0) < 1e-12
    assert res < 1.5e-8
    print("norb=14 Davidson, start 2", flush=True)
    r2 = dv.davidson(ham, 7, 7, tol=1e-8, max_iter=800, seed=11)
    assert r2.converged
    print(f"  {r2.iterations} iterations, theta {r2.eigenvalues[0]:.12f}",
          flush=True)
    assert abs(r2.eigenvalues[0] - theta) < 1e-9


def _ref_taylor(c, h1, h2, g, t, accuracy=1e-15, expansion=30):
    """The reference's Taylor loop (src/fqe/wavefunction.py:556-568) over
    its compiled apply with the tensors of iht(t)."""
    a1, a2 = -1j * t * h1, -1j * t * h2
    work, result = c.copy(), c.copy()
    for order in range(1, expansion):
        work = ref_lm.apply_12_lm(work, a1, a2, g)
        coef = 1.0 / math.factorial(order)
        result += coef * work
        if numpy.linalg.norm(work) * coef < accuracy:
            return result, order
    raise RuntimeError("maximum taylor expansion limit reached")


@pytest.mark.skipif(not ref_lm.available(), reason="oracle/_ref not built")
def test_norb12_chain_and_evolve_vs_reference_c(fqe):
    c, h1, h2 = fo.synthetic_workload(12, 6, 6, 1, 12)
    key = (12, 0)
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    wfn = fqe.Wavefunction([[12, 0, 12]])
    wfn.set_wfn(strategy="from_data", raw_data={key: c})
    print("norb=12: 100 chained applies on the GPU", flush=True)
    x = wfn
    for _ in range(100):
        x = x.apply(ham)
        x.normalize()
    gpu_chain = x.get_coeff(key)
    ev = wfn.time_evolve(0.05, ham)
    g = ref_lm.RefGraph(6, 6, 12)
    print("norb=12: 100 chained applies with the reference C", flush=True)
    y = c.copy()
    for _ in range(100):
        y = ref_lm.apply_12_lm(y, h1, h2, g)
        y /= numpy.linalg.norm(y)
    e_chain = rel(gpu_chain, y)
    ref_ev, nterms = _ref_taylor(c, h1, h2, g, 0.05)
    got = ev.get_coeff(key)
    e_ev = rel(got, ref_ev)
    print(f"  chain {e_chain:.2e}, time_evolve {e_ev:.2e}, terms "
          f"{ev._taylor_terms} / {nterms}", flush=True)
    assert e_chain < 1e-11
    assert e_ev < 1e-11
    assert ev._taylor_terms == nterms
    assert abs(numpy.linalg.norm(got) - 1.0) < 1e-12
