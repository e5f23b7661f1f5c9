"""GPU tests of the Davidson eigensolver kernels and drivers against the
oracle, FQE's LiH fixture (tests/davidson_test.py:26-40,
tests/wavefunction_test.py:537-548) and dense diagonalisation."""

import copy
import os

import numpy
import pytest

from oracle import fqe_oracle as fo

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
E_FCI_LIH = -8.877719570384043
E_HF_LIH = -8.857341498221992


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def _lih():
    z = numpy.load(os.path.join(GOLDEN, "lih_energy.npz"))
    return z["h1e"], z["h2e"], z["wfn"]


@pytest.mark.parametrize("na,nb,norb,kind", [(2, 2, 6, "lih"),
                                             (3, 2, 7, "symmetric"),
                                             (2, 3, 6, "real"),
                                             (0, 3, 5, "symmetric"),
                                             (4, 4, 4, "real")])
def test_diagonal_matches_slater_condon_oracle(fqe, na, nb, norb, kind):
    from fqe_amd.algorithm import davidson as dv
    print(f"diagonal {na} {nb} {norb} {kind}", flush=True)
    rng = numpy.random.default_rng(3 * norb + na)
    if kind == "lih":
        h1, h2, _ = _lih()
    elif kind == "symmetric":
        h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    else:
        h1 = rng.normal(size=(norb, norb))
        h2 = rng.normal(size=(norb,) * 4)
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    graph = fqe.FciGraph(na, nb, norb)
    d = dv.diagonal(ham, graph).cpu().numpy()
    g = fo.OracleGraph(na, nb, norb)
    ref = fo.slater_condon_diagonal(h1, h2, g)
    assert numpy.max(numpy.abs(d - ref)) < 1e-12 * max(1.0, numpy.abs(ref).max())
    if kind == "lih":
        assert abs(d[0, 0] - E_HF_LIH) < 1e-12


def test_block_dot_and_combine(fqe):
    import torch
    from fqe_amd import _lib
    from fqe_amd.algorithm.davidson import Subspace
    rng = numpy.random.default_rng(5)
    n, k = 100003, 11
    vs = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    sub = Subspace(n, 16, _lib.device())
    sub.v[:k].copy_(torch.from_numpy(vs))
    sub.k = k
    xd = torch.from_numpy(x).to(_lib.device())
    dots = sub.dots(xd).cpu().numpy()
    ref = vs.conj() @ x
    assert numpy.max(numpy.abs(dots - ref)) < 1e-10 * numpy.abs(ref).max()
    c = rng.normal(size=k) + 1j * rng.normal(size=k)
    sub.combine(c, xd, beta=0.5 - 0.25j)
    assert numpy.allclose(xd.cpu().numpy(), (0.5 - 0.25j) * x + c @ vs,
                          rtol=1e-13, atol=1e-11)


@pytest.mark.parametrize("k,m,n", [(11, 9, 100003), (40, 16, 5000), (1, 1, 7),
                                   (5, 3, 1)])
def test_block_combine_multi(fqe, k, m, n):
    """fqehip_block_combine_multi (a thick restart's new basis) against
    numpy: out[j] = sum_l c[l][j] v[l], every output of the call."""
    import torch
    from fqe_amd import _lib
    rng = numpy.random.default_rng(k * 7 + m)
    vs = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    c = rng.normal(size=(k, m)) + 1j * rng.normal(size=(k, m))
    dev = _lib.device()
    vd = torch.from_numpy(vs).to(dev)
    cd = torch.from_numpy(c).to(dev)
    out = torch.full((m + 1, n), 3.0 + 3.0j, dtype=torch.complex128, device=dev)
    _lib.call("fqehip_block_combine_multi", n, k, _lib.ptr(vd), n,
              _lib.ptr(cd), m, _lib.ptr(out), n, _lib.stream_ptr())
    got = out.cpu().numpy()
    ref = c.T @ vs
    assert numpy.max(numpy.abs(got[:m] - ref)) < 1e-12 * max(1.0, numpy.abs(ref).max())
    assert numpy.all(got[m] == 3.0 + 3.0j)


def test_davidson_lih_ground_state(fqe):
    """Residual-converged Davidson reaches FQE's LiH FCI energy and
    eigenvector (tests/davidson_test.py:26-40)."""
    from fqe_amd.algorithm import davidson as dv
    h1, h2, wfn = _lih()
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    res = dv.davidson(ham, 2, 2, tol=1e-9)
    print(f"LiH davidson: {res.iterations} iterations, E={res.eigenvalues[0]}",
          flush=True)
    assert res.converged
    assert abs(res.eigenvalues[0] - E_FCI_LIH) < 1e-10
    u = res.vectors[0].cpu().numpy()
    ovl = abs(numpy.vdot(wfn.reshape(-1), u)) / numpy.linalg.norm(wfn)
    assert abs(ovl - 1.0) < 1e-9


def test_davidson_matches_dense_eigh(fqe):
    from fqe_amd.algorithm import davidson as dv
    rng = numpy.random.default_rng(21)
    norb, na, nb = 6, 3, 2
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    w = numpy.linalg.eigvalsh(fo.dense_matrix(h1, h2, g))
    ham = fqe.get_restricted_hamiltonian((h1, h2), e_0=0.25)
    res = dv.davidson(ham, na, nb, tol=1e-9, max_subspace=8)
    assert res.converged
    assert abs(res.eigenvalues[0] - (w[0] + 0.25)) < 1e-10


def test_davidson_diagonalization_reference_api(fqe):
    """fqe.algorithm.davidson.davidson_diagonalization on LiH: the
    eigenvalue-change rule of the reference (epsilon=1e-8)."""
    from fqe_amd.algorithm import davidson as dv
    h1, h2, _ = _lih()
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    w, vecs = dv.davidson_diagonalization(ham, 2, 2, nroots=2)
    g = fo.OracleGraph(2, 2, 6)
    ref = numpy.linalg.eigvalsh(fo.dense_matrix(h1, h2, g))
    assert abs(w[0] - E_FCI_LIH) < 1e-7
    assert abs(w[1] - ref[1]) < 1e-6
    assert len(vecs) == 2
    e = vecs[0].expectationValue(ham)
    assert abs(e - w[0]) < 1e-7


def test_davidsonliu_errors(fqe):
    from fqe_amd.algorithm import davidson as dv
    h1, h2, _ = _lih()
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    w1 = fqe.Wavefunction([[4, 0, 6]])
    w1.set_wfn(strategy="random")
    w2 = fqe.Wavefunction([[4, 2, 6]])
    w2.set_wfn(strategy="random")
    with pytest.raises(ValueError):
        dv.davidsonliu_fqe(ham, 0, [w1], 4, 0, 6)
    with pytest.raises(TypeError):
        dv.davidsonliu_fqe(ham, 1, [w1, w2], 4, 0, 6)
    ok = copy.deepcopy(w1)
    ok.normalize()


def test_davidson_thick_restart_matches_dense_eigh(fqe):
    """A subspace of 6 vectors forces many thick restarts (3 kept Ritz
    vectors + the previous one); the solver still reaches the exact lowest
    eigenvalue."""
    from fqe_amd.algorithm import davidson as dv
    rng = numpy.random.default_rng(23)
    norb, na, nb = 7, 3, 3
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    w = numpy.linalg.eigvalsh(fo.dense_matrix(h1, h2, g))
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    res = dv.davidson(ham, na, nb, tol=1e-9, max_subspace=6, keep=3,
                      max_iter=800)
    print(f"thick restart: {res.iterations} iterations", flush=True)
    assert res.converged
    assert abs(res.eigenvalues[0] - w[0]) < 1e-10


def test_davidson_mixed_then_fp64(fqe):
    """Mixed-precision applies to the mixed floor, then FP64: the same
    eigenvalue as dense diagonalisation to 1e-10 and a 1e-9 residual."""
    from fqe_amd.algorithm import davidson as dv
    rng = numpy.random.default_rng(29)
    norb, na, nb = 7, 3, 3
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    w = numpy.linalg.eigvalsh(fo.dense_matrix(h1, h2, g))
    ham = fqe.get_restricted_hamiltonian((h1, h2), e_0=-0.5)
    res = dv.davidson(ham, na, nb, tol=1e-9, dtype="mixed")
    print(f"mixed-then-FP64: {res.iterations} iterations", flush=True)
    assert res.converged
    assert abs(res.eigenvalues[0] - (w[0] - 0.5)) < 1e-10


@pytest.mark.parametrize("k,n", [(37, 10001), (16, 4096), (5, 3), (144, 9000)])
def test_gram_matches_numpy(fqe, k, n):
    """fqehip_gram (the 2-RDM's dvec overlap) against numpy, with ragged
    k (not a multiple of 16) and n (not a multiple of 4 or of a chunk)."""
    import torch
    from fqe_amd import _lib, engine
    rng = numpy.random.default_rng(k + n)
    a = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    b = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    dev = _lib.device()
    g = engine.gram(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev))
    ref = a.conj() @ b.T
    assert numpy.max(numpy.abs(g - ref)) < 1e-11 * max(1.0, numpy.abs(ref).max())
