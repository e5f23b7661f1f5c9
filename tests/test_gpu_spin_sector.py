"""Spin-orbital operators on spin-conserving wavefunctions, on the GPU:
``FqeData.apply`` with (2 norb)-sized tensors (the reference's
``_apply_array_spin1`` / ``_apply_array_spin12``), ``Wavefunction.apply``
and ``time_evolve`` with ``SSOHamiltonian`` / ``General`` on sectors of
fixed spin.  Checked against the oracle (itself pinned by the reference's
one-column fixtures, the restricted apply and a Jordan-Wigner operator),
against the restricted engine for spin-orbital copies of restricted
tensors, and against exp(-iHt) of the Jordan-Wigner matrix.  FP64
throughout: relative error <= 1e-12 (1e-11 for time evolution)."""

import numpy
import pytest
import scipy.linalg

from oracle import fqe_oracle as fo
from oracle import jw_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return float(numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300))


def _general(n, rng, cplx=True, mixing=True):
    """Random spin-orbital tensors; with ``mixing`` the alpha-beta blocks
    of h1 and the spin-changing blocks of h2 are filled too (the sector
    apply must ignore them, as the reference does)."""
    m = 2 * n

    def draw(shape):
        x = rng.normal(size=shape)
        return x + 1j * rng.normal(size=shape) if cplx else x + 0j

    h1 = draw((m, m)) if mixing else numpy.zeros((m, m), complex)
    h2 = draw((m,) * 4) if mixing else numpy.zeros((m,) * 4, complex)
    for blk in (slice(0, n), slice(n, m)):
        h1[blk, blk] = draw((n, n))
    return h1, h2


def _sector(fqe, na, nb, n, c):
    d = fqe.FqeData(na, nb, n)
    d.set_wfn(strategy="from_data", raw_data=c)
    return d


def test_spin1_one_column_fixtures(fqe):
    """reference tests/fqe_data_test.py:486-518"""
    for col, rows, want in ((0, (0, 2, 3), [[1, 1], [0, 0]]),
                            (3, (0, 1, 3), [[0, 1], [0, 1]])):
        h1 = numpy.zeros((4, 4), dtype=numpy.complex128)
        for r in rows:
            h1[r, col] = 1.0
        d = fqe.FqeData(1, 1, 2)
        d.set_wfn(strategy="ones")
        d.apply_inplace((h1,))
        assert numpy.allclose(d.coeff, numpy.array(want, dtype=complex),
                              rtol=0, atol=1e-14)


@pytest.mark.parametrize("na,nb,n,cplx", [
    (1, 1, 2, True), (2, 1, 4, True), (2, 2, 4, False), (1, 3, 5, True),
    (3, 3, 6, True), (4, 2, 7, False), (0, 3, 5, True), (4, 4, 8, True),
])
def test_spin_sector_apply_matches_oracle(fqe, na, nb, n, cplx):
    print(f"spin sector apply {na} {nb} {n} cplx={cplx}", flush=True)
    rng = numpy.random.default_rng(31 * n + 7 * na + nb)
    g = fo.OracleGraph(na, nb, n)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    h1, h2 = _general(n, rng, cplx)
    out = _sector(fqe, na, nb, n, c).apply((h1, h2)).coeff
    assert rel(out, fo.apply_spin12(c, h1, h2, g)) < 1e-12
    out1 = _sector(fqe, na, nb, n, c).apply((h1,)).coeff
    assert rel(out1, fo.apply_spin1(c, h1, g)) < 1e-12


@pytest.mark.parametrize("na,nb,n", [(2, 2, 4), (3, 2, 6), (4, 4, 8)])
def test_spin_orbital_copy_of_restricted_equals_restricted(fqe, na, nb, n):
    rng = numpy.random.default_rng(50 + n)
    g = fo.OracleGraph(na, nb, n)
    h1, h2 = fo.synthetic_hamiltonian(n, n, rng)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    h1so, h2so = fo.restricted_as_spin_orbital(h1, h2)
    spin = _sector(fqe, na, nb, n, c).apply((h1so, h2so)).coeff
    restricted = _sector(fqe, na, nb, n, c).apply((h1, h2)).coeff
    assert rel(spin, restricted) < 1e-12


def test_wavefunction_apply_sso_and_general_with_e0(fqe):
    n = 5
    rng = numpy.random.default_rng(8)
    h1, h2 = _general(n, rng)
    wfn = fqe.Wavefunction([[4, 0, n], [4, 2, n], [3, 1, n]])
    wfn.set_wfn(strategy="random")
    for ham in (fqe.SSOHamiltonian((h1, h2), e_0=0.3 - 0.1j),
                fqe.General((h1, h2), e_0=0.3 - 0.1j)):
        out = wfn.apply(ham)
        for key in ((4, 0), (4, 2), (3, 1)):
            nele, sz = key
            na, nb = (nele + sz) // 2, (nele - sz) // 2
            g = fo.OracleGraph(na, nb, n)
            c = wfn.get_coeff(key)
            want = fo.apply_spin12(c, h1, h2, g) + (0.3 - 0.1j) * c
            assert rel(out.get_coeff(key), want) < 1e-12
    quad = fqe.SSOHamiltonian((h1,))
    out = wfn.apply(quad)
    g = fo.OracleGraph(2, 2, n)
    assert rel(out.get_coeff((4, 0)),
               fo.apply_spin1(wfn.get_coeff((4, 0)), h1, g)) < 1e-12


def _hermitian_general(n, rng):
    """Hermitian spin-conserving tensors whose terms all lie in the blocks
    the sector apply uses (so exp(-iHt) of the Jordan-Wigner matrix is the
    exact evolution)."""
    m = 2 * n
    h1 = numpy.zeros((m, m), complex)
    for blk in (slice(0, n), slice(n, m)):
        x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        h1[blk, blk] = 0.5 * (x + x.conj().T)
    h2 = numpy.zeros((m,) * 4, complex)
    for rho in range(2):
        for eta in range(2):
            r = slice(rho * n, (rho + 1) * n)
            e = slice(eta * n, (eta + 1) * n)
            h2[r, e, r, e] = 0.1 * rng.normal(size=(n,) * 4)
    # Hermitian: h2[p,q,r,s] = conj(h2[s,r,q,p]) (a+p a+q a r a s)
    h2 = 0.5 * (h2 + numpy.conj(h2.transpose(3, 2, 1, 0)))
    return h1, h2


@pytest.mark.parametrize("graph", [False, True])
def test_time_evolve_sso_matches_expm(fqe, graph):
    from fqe_amd import settings
    n, na, nb, t = 4, 2, 1, 0.3
    rng = numpy.random.default_rng(21)
    h1, h2 = _hermitian_general(n, rng)
    g = fo.OracleGraph(na, nb, n)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    mat = jw_dense.dense_set_operator(h1, h2, n, na + nb)
    _, blocks = jw_dense.set_basis(n, na + nb)
    off = [o for key, o, la, lb in blocks if key == (na, nb)][0]
    sz = g.lena * g.lenb
    block = mat[off:off + sz, off:off + sz]
    want = (scipy.linalg.expm(-1j * t * block) @ c.reshape(-1)).reshape(c.shape)
    saved = settings.taylor_graph
    settings.taylor_graph = graph
    try:
        wfn = fqe.Wavefunction([[na + nb, na - nb, n]])
        wfn.set_wfn(strategy="from_data", raw_data={(na + nb, na - nb): c})
        ev = wfn.time_evolve(t, fqe.SSOHamiltonian((h1, h2)))
    finally:
        settings.taylor_graph = saved
    assert rel(ev.get_coeff((na + nb, na - nb)), want) < 1e-11


def test_quadratic_sso_evolve_matches_two_body_form(fqe):
    """reference tests/wavefunction_test.py:322-355: an SSO Hamiltonian with
    only h1 evolves like the same h1 with a zero h2."""
    norb, nalpha, nbeta, time = 4, 2, 2, 0.1
    rng = numpy.random.default_rng(826283)
    h1e = numpy.zeros((2 * norb, 2 * norb))
    h2e = numpy.zeros((2 * norb,) * 4)
    h1a = rng.uniform(-1, 1, size=(norb, norb))
    h1a += h1a.T
    h1b = rng.uniform(-1, 1, size=(norb, norb))
    h1b += h1b.T
    h1e[:norb, :norb] = h1a
    h1e[norb:, norb:] = h1b
    wfn = fqe.Wavefunction([[nalpha + nbeta, nalpha - nbeta, norb]])
    wfn.set_wfn(strategy="random")
    wfn.normalize()
    ref = wfn.time_evolve(time, fqe.SSOHamiltonian((h1e, h2e)))
    out = wfn.time_evolve(time, fqe.SSOHamiltonian((h1e,)))
    key = (nalpha + nbeta, nalpha - nbeta)
    assert rel(out.get_coeff(key), ref.get_coeff(key)) < 1e-12
