"""The oracle's spin-orbital apply on one spin-conserving sector
(``fqe_oracle.apply_spin1`` / ``apply_spin12``), pinned by:

* the reference's literal expected outputs of ``_apply_array_spin1`` with
  one nonzero column (``tests/fqe_data_test.py:486-518``);
* the restricted apply: spin-orbital tensors built from a restricted
  (h1, h2) must give the spatial apply, itself pinned by FQE's golden data;
* an exact Jordan-Wigner operator of unrestricted tensors (different
  alpha-alpha, beta-beta and alpha-beta parts, complex), whose every term
  lies in the spin-conserving blocks the reference uses;
* a planted mistake (the alpha-beta and beta-alpha blocks exchanged) that
  the Jordan-Wigner check must catch.
"""

import numpy
import pytest

from oracle import fqe_oracle as fo
from oracle import jw_dense


def _unrestricted(n, rng, cplx=True):
    m = 2 * n

    def draw(shape):
        x = rng.normal(size=shape)
        return x + 1j * rng.normal(size=shape) if cplx else x

    h1 = numpy.zeros((m, m), complex)
    for blk in (slice(0, n), slice(n, m)):
        h1[blk, blk] = draw((n, n))
    h2 = numpy.zeros((m,) * 4, complex)
    for rho in range(2):
        for eta in range(2):
            r = slice(rho * n, (rho + 1) * n)
            e = slice(eta * n, (eta + 1) * n)
            h2[r, e, r, e] = draw((n,) * 4)
    return h1, h2


def _jw_sector(h1, h2, na, nb, n):
    nele = na + nb
    mat = jw_dense.dense_set_operator(h1, h2, n, nele)
    _, blocks = jw_dense.set_basis(n, nele)
    for key, off, lena, lenb in blocks:
        if key == (na, nb):
            return mat[off:off + lena * lenb, off:off + lena * lenb]
    raise KeyError((na, nb))


def test_spin1_one_column_fixtures():
    g = fo.OracleGraph(1, 1, 2)
    c = numpy.ones((2, 2), dtype=numpy.complex128)
    h1 = numpy.zeros((4, 4), dtype=numpy.complex128)
    h1[0, 0] = h1[2, 0] = h1[3, 0] = 1.0
    assert numpy.array_equal(fo.apply_spin1(c, h1, g),
                             numpy.array([[1, 1], [0, 0]], dtype=complex))
    h1 = numpy.zeros((4, 4), dtype=numpy.complex128)
    h1[0, 3] = h1[1, 3] = h1[3, 3] = 1.0
    assert numpy.array_equal(fo.apply_spin1(c, h1, g),
                             numpy.array([[0, 1], [0, 1]], dtype=complex))


@pytest.mark.parametrize("na,nb,n", [(2, 1, 4), (2, 2, 4), (1, 3, 5),
                                     (3, 3, 6)])
def test_spin12_of_restricted_tensors_is_the_spatial_apply(na, nb, n):
    rng = numpy.random.default_rng(100 + n + na)
    g = fo.OracleGraph(na, nb, n)
    h1, h2 = fo.synthetic_hamiltonian(n, n, rng)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    h1so, h2so = fo.restricted_as_spin_orbital(h1, h2)
    ref = fo.apply_12(c, h1, h2, g)
    out = fo.apply_spin12(c, h1so, h2so, g)
    assert numpy.abs(out - ref).max() < 1e-12 * numpy.abs(ref).max()
    ref1 = fo.apply_1(c, h1, g)
    out1 = fo.apply_spin1(c, h1so, g)
    assert numpy.abs(out1 - ref1).max() < 1e-12 * numpy.abs(ref1).max()


@pytest.mark.parametrize("na,nb,n,cplx", [(2, 1, 3, True), (1, 2, 4, True),
                                          (2, 2, 4, False), (0, 2, 3, True)])
def test_spin12_matches_jordan_wigner(na, nb, n, cplx):
    rng = numpy.random.default_rng(7 * n + na + 3 * nb)
    g = fo.OracleGraph(na, nb, n)
    h1, h2 = _unrestricted(n, rng, cplx)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = (_jw_sector(h1, h2, na, nb, n) @ c.reshape(-1)).reshape(c.shape)
    out = fo.apply_spin12(c, h1, h2, g)
    assert numpy.abs(out - ref).max() < 1e-12 * numpy.abs(ref).max()
    ref1 = (_jw_sector(h1, None, na, nb, n) @ c.reshape(-1)).reshape(c.shape)
    out1 = fo.apply_spin1(c, h1, g)
    assert numpy.abs(out1 - ref1).max() < 1e-12 * numpy.abs(ref1).max()


def test_planted_block_exchange_is_caught():
    """Exchanging the alpha-beta and beta-alpha two-body blocks (a plausible
    indexing mistake) must fail the Jordan-Wigner check."""
    na, nb, n = 2, 1, 3
    rng = numpy.random.default_rng(11)
    g = fo.OracleGraph(na, nb, n)
    h1, h2 = _unrestricted(n, rng)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = (_jw_sector(h1, h2, na, nb, n) @ c.reshape(-1)).reshape(c.shape)
    a, b = slice(0, n), slice(n, 2 * n)
    bad = h2.copy()
    bad[a, b, a, b], bad[b, a, b, a] = h2[b, a, b, a], h2[a, b, a, b]
    out = fo.apply_spin12(c, h1, bad, g)
    assert numpy.abs(out - ref).max() > 1e-3 * numpy.abs(ref).max()
