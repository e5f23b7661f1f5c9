"""Pins the spin-orbital (FqeDataSet) oracle, CPU only.

* the reference's FqeDataSet(nele=2, norb=2) fixtures (reference
  ``tests/fqe_data_set_test.py:122-158,229-255``, data of
  ``tests/unittest_data/fqe_data_set/set0202``, copied as arrays into
  ``tests/golden/fqe_data_set_golden.npz``);
* a Jordan-Wigner dense operator over all sectors of one particle number,
  for random general (non-Hermitian, spin-mixing) h1 and h2, read in the
  reference's determinant phases (sector (nalpha, nbeta) carries
  (-1)^((nele - 1) nbeta) against the Jordan-Wigner determinants);
* brute-force <a+_p a_q> and <a+_p a+_q a_r a_s> for the RDMs;
* the restricted oracle per sector when the tensors are spin-diagonal;
* spin-flip signs planted wrong in the oracle (a dropped (-1)^nalpha, and
  parities counted below instead of above the orbital) are caught by the
  fixtures and the brute force.
"""

import os

import numpy
import pytest

from oracle import fqe_oracle as fo
from oracle import fqe_set_oracle as so
from oracle import jw_dense

GOLDEN = os.path.join(os.path.dirname(__file__), "golden",
                      "fqe_data_set_golden.npz")


def _golden():
    z = numpy.load(GOLDEN)
    return {k: z[k] for k in z.files}


def _tag(key, norb):
    return f"{key[0]:02d}{key[1]:02d}{norb:02d}"


def _flat(vecs, sg):
    return numpy.concatenate([vecs[k].ravel() for k in sg.keys])


def _unflat(vec, sg):
    out, off = {}, 0
    for k in sg.keys:
        n = numpy.prod(sg.shape(k))
        out[k] = vec[off:off + n].reshape(sg.shape(k))
        off += n
    return out


def _random_set(sg, rng):
    return {k: rng.standard_normal(sg.shape(k))
            + 1j * rng.standard_normal(sg.shape(k)) for k in sg.keys}


@pytest.mark.parametrize("suffix", ["1", "2", "12"])
def test_set_apply_matches_reference_fixture(suffix):
    g = _golden()
    norb = 2
    sg = so.SetGraph(2, norb)
    assert sg.keys == [(2, 0), (1, 1), (0, 2)]
    coeffs = {k: g[f"c_{_tag(k, norb)}"] for k in sg.keys}
    if suffix == "1":
        tensors = (g["h1"],)
    elif suffix == "2":
        tensors = (numpy.zeros_like(g["h1"]), g["h2"])
    else:
        tensors = (g["h1"], g["h2"])
    out = so.set_apply(coeffs, tensors, sg)
    for k in sg.keys:
        numpy.testing.assert_allclose(out[k], g[f"out{suffix}_{_tag(k, norb)}"],
                                      rtol=1e-12, atol=1e-12)


def test_set_rdm12_matches_reference_fixture():
    g = _golden()
    norb = 2
    sg = so.SetGraph(2, norb)
    coeffs = {k: g[f"c_{_tag(k, norb)}"] for k in sg.keys}
    d1, d2 = so.set_rdm12(coeffs, sg)
    numpy.testing.assert_allclose(d1, g["rdm1"], rtol=1e-12, atol=1e-12)
    numpy.testing.assert_allclose(d2, g["rdm2"], rtol=1e-12, atol=1e-12)


def test_fixture_hamiltonian_is_spin_diagonal_restricted():
    # The fixture's h1/h2 are the spatial closed forms placed on the
    # spin blocks; the spin-orbital apply on them equals the restricted
    # apply of the spatial tensors on every sector.
    g = _golden()
    norb = 2
    sg = so.SetGraph(2, norb)
    coeffs = {k: g[f"c_{_tag(k, norb)}"] for k in sg.keys}
    h1 = g["h1"][:norb, :norb]
    h2 = g["h2"][:norb, :norb, :norb, :norb]
    out = so.set_apply(coeffs, (g["h1"], g["h2"]), sg)
    for k in sg.keys:
        ref = fo.apply_12(coeffs[k], h1, h2, fo.OracleGraph(k[0], k[1], norb))
        numpy.testing.assert_allclose(out[k], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nele,norb", [(2, 2), (1, 3), (2, 3), (3, 3), (3, 4)])
def test_set_apply_matches_jordan_wigner(nele, norb):
    rng = numpy.random.default_rng(100 * nele + norb)
    m = 2 * norb
    sg = so.SetGraph(nele, norb)
    h1 = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    h2 = rng.standard_normal((m,) * 4) + 1j * rng.standard_normal((m,) * 4)
    coeffs = _random_set(sg, rng)
    ph = jw_dense.set_sector_phase(norb, nele)
    mat = jw_dense.dense_set_operator(h1, h2, norb, nele)
    want = _unflat(ph * (mat @ (ph * _flat(coeffs, sg))), sg)
    got = so.set_apply(coeffs, (h1, h2), sg)
    for k in sg.keys:
        numpy.testing.assert_allclose(got[k], want[k], rtol=1e-11, atol=1e-11)
    mat1 = jw_dense.dense_set_operator(h1, None, norb, nele)
    want1 = _unflat(ph * (mat1 @ (ph * _flat(coeffs, sg))), sg)
    got1 = so.set_apply(coeffs, (h1,), sg)
    for k in sg.keys:
        numpy.testing.assert_allclose(got1[k], want1[k], rtol=1e-11,
                                      atol=1e-11)


@pytest.mark.parametrize("nele,norb", [(2, 2), (3, 3), (2, 3)])
def test_set_rdm12_matches_brute_force(nele, norb):
    rng = numpy.random.default_rng(7 + nele + norb)
    m = 2 * norb
    sg = so.SetGraph(nele, norb)
    coeffs = _random_set(sg, rng)
    vec = jw_dense.set_sector_phase(norb, nele) * _flat(coeffs, sg)
    d1, d2 = so.set_rdm12(coeffs, sg)
    for p in range(m):
        for q in range(m):
            want = jw_dense.set_expectation(vec, ((p, True), (q, False)),
                                            norb, nele)
            assert abs(d1[p, q] - want) < 1e-10
    for p, q, r, s in numpy.ndindex(m, m, m, m):
        want = jw_dense.set_expectation(
            vec, ((p, True), (q, True), (r, False), (s, False)), norb, nele)
        assert abs(d2[p, q, r, s] - want) < 1e-10


def _plant_dropped_nalpha_sign(sg):
    """Negate the spin-flip excitations of odd-nalpha sources: what a
    dropped (-1)^nalpha (or (-1)^(nalpha - 1)) factor would give."""
    planted = []
    for p, q, src, tgt, amat, bmat in sg.excitations:
        if src != tgt and src[0] % 2 == 1:
            amat = -amat
        planted.append((p, q, src, tgt, amat, bmat))
    sg.excitations = planted


def _plant_below_parity(monkeypatch):
    """Single-string parities counted below the orbital (the restricted
    maps' habit) instead of above it."""
    def below(src, tgt, orb, create):
        index = {int(s): a for a, s in enumerate(tgt)}
        mat = numpy.zeros((len(tgt), len(src)))
        for a, s in enumerate(src):
            s = int(s)
            if ((s >> orb) & 1) == (0 if create else 1):
                n = bin(s & ((1 << orb) - 1)).count("1")
                mat[index[s ^ (1 << orb)], a] = -1.0 if n % 2 else 1.0
        return mat
    monkeypatch.setattr(so, "_string_matrix", below)


def _caught(sg, nele, norb, seed):
    """True when the planted oracle disagrees with the reference fixture
    (nele=2, norb=2) or with the phased Jordan-Wigner operator."""
    if (nele, norb) == (2, 2):
        g = _golden()
        coeffs = {k: g[f"c_{_tag(k, norb)}"] for k in sg.keys}
        d1, d2 = so.set_rdm12(coeffs, sg)
        return not (numpy.allclose(d1, g["rdm1"])
                    and numpy.allclose(d2, g["rdm2"]))
    rng = numpy.random.default_rng(seed)
    m = 2 * norb
    h1 = rng.standard_normal((m, m))
    h2 = rng.standard_normal((m,) * 4)
    coeffs = _random_set(sg, rng)
    ph = jw_dense.set_sector_phase(norb, nele)
    mat = jw_dense.dense_set_operator(h1, h2, norb, nele)
    want = ph * (mat @ (ph * _flat(coeffs, sg)))
    got = _flat(so.set_apply(coeffs, (h1, h2), sg), sg)
    return not numpy.allclose(got, want)


@pytest.mark.parametrize("nele,norb", [(2, 2), (3, 3), (2, 3)])
def test_planted_dropped_nalpha_sign_is_caught(nele, norb):
    sg = so.SetGraph(nele, norb)
    _plant_dropped_nalpha_sign(sg)
    assert _caught(sg, nele, norb, 5)


@pytest.mark.parametrize("nele,norb", [(2, 2), (2, 3)])
def test_planted_below_parity_is_caught(monkeypatch, nele, norb):
    # Below- and above-counts differ by a sector phase that is trivial
    # for odd nele, so even electron counts are the ones that see it.
    _plant_below_parity(monkeypatch)
    sg = so.SetGraph(nele, norb)
    assert _caught(sg, nele, norb, 6)
