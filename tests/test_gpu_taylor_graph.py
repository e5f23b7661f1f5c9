"""time_evolve with the whole Taylor series recorded into one HIP graph
(settings.taylor_graph) against the eager series: the same kernels with
the same arguments, so FP64 results agree to 1e-13 (the split kernels are
deterministic) and the term counts are equal; replays with eager work
between them give the same result each time; a series that does not
converge still raises RuntimeError."""

import numpy
import pytest

from oracle import fqe_oracle as fo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return float(numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300))


def _evolve(fqe, c, h1, h2, na, nb, norb, t, graph, dtype, contraction):
    from fqe_amd import settings
    saved = settings.taylor_graph, settings.contraction
    settings.taylor_graph, settings.contraction = graph, contraction
    try:
        key = (na + nb, na - nb)
        wfn = fqe.Wavefunction([[na + nb, na - nb, norb]], dtype=dtype)
        wfn.set_wfn(strategy="from_data", raw_data={key: c})
        ham = fqe.get_restricted_hamiltonian((h1, h2), e_0=0.3)
        outs = []
        for _ in range(3 if graph else 1):
            ev = wfn.time_evolve(t, ham)
            outs.append((ev.get_coeff(key), ev._taylor_terms))
            import torch
            torch.zeros(1000, device="cuda").sum().item()
        return outs
    finally:
        settings.taylor_graph, settings.contraction = saved


@pytest.mark.parametrize("na,nb,norb,dtype,contraction", [
    (4, 4, 8, numpy.complex128, "fp64"),
    (4, 4, 8, numpy.complex128, "ozaki"),
    (3, 2, 7, numpy.complex64, "fp64"),
    (6, 6, 12, numpy.complex128, "fp64"),
])
def test_taylor_graph_matches_eager(fqe, na, nb, norb, dtype, contraction):
    print(f"taylor graph {na} {nb} {norb} {dtype.__name__} {contraction}",
          flush=True)
    c, h1, h2 = fo.synthetic_workload(norb, na, nb, norb, norb)
    eager = _evolve(fqe, c, h1, h2, na, nb, norb, 0.1, False, dtype,
                    contraction)[0]
    graph = _evolve(fqe, c, h1, h2, na, nb, norb, 0.1, True, dtype,
                    contraction)
    tol = 1e-13 if dtype == numpy.complex128 else 1e-6
    for out, terms in graph:
        assert terms == eager[1]
        assert rel(out, eager[0]) < tol
    if norb <= 8 and dtype == numpy.complex128:
        g = fo.OracleGraph(na, nb, norb)
        ref, nterms = fo.taylor_evolve(c, (h1, h2), 0.3, 0.1, g)
        assert rel(graph[0][0], ref) < 1e-9
        assert graph[0][1] == nterms


def test_taylor_graph_not_converged_raises(fqe):
    from fqe_amd import settings
    c, h1, h2 = fo.synthetic_workload(6, 3, 3, 1, 6)
    saved = settings.taylor_graph
    settings.taylor_graph = True
    try:
        wfn = fqe.Wavefunction([[6, 0, 6]])
        wfn.set_wfn(strategy="from_data", raw_data={(6, 0): c})
        ham = fqe.get_restricted_hamiltonian((h1, h2))
        with pytest.raises(RuntimeError):
            wfn.apply_generated_unitary(1.0, "taylor", ham, expansion=4)
    finally:
        settings.taylor_graph = saved
