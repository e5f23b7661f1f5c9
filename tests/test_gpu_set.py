"""GPU parity of the spin-orbital set path (FqeDataSet apply and RDMs, and
General / GSOHamiltonian operators on spin-broken wavefunctions).

Every test runs fqehip_set_apply / fqehip_set_dvec through the C ABI and
checks it against the reference's set0202 fixtures, the spin-orbital oracle
(pinned on the CPU by tests/test_set_oracle.py) or the exponential of the
phased Jordan-Wigner matrix.  The path computes in complex128 (gathers and
one FP64 matrix-core GEMM, no atomics): results agree with the oracle to a
normwise relative error of 1e-12 and are bit-identical run to run.
"""

import os

import numpy
import pytest

from oracle import fqe_set_oracle as so

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden",
                      "fqe_data_set_golden.npz")
TOL = 1e-12


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return numpy.linalg.norm(numpy.asarray(a) - numpy.asarray(b)) / max(
        numpy.linalg.norm(numpy.asarray(b)), 1e-300)


def make_set(fqe, nele, norb, coeffs):
    data = {}
    for key, c in coeffs.items():
        sec = fqe.FqeData(key[0], key[1], norb)
        sec.coeff = numpy.array(c, dtype=numpy.complex128)
        data[key] = sec
    return fqe.FqeDataSet(nele, norb, data)


def random_coeffs(sg, rng):
    return {k: rng.standard_normal(sg.shape(k))
            + 1j * rng.standard_normal(sg.shape(k)) for k in sg.keys}


def random_tensors(norb, rng, hermitian=False):
    m = 2 * norb
    h1 = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    h2 = rng.standard_normal((m,) * 4) + 1j * rng.standard_normal((m,) * 4)
    if hermitian:
        h1 = h1 + h1.conj().T
        # a+_p a+_q a_r a_s is the adjoint of a+_s a+_r a_q a_p
        h2 = h2 + h2.conj().transpose(3, 2, 1, 0)
    return h1, h2


def print_case(name, **kw):
    print(f"[set] {name} {kw}", flush=True)


@pytest.mark.parametrize("suffix", ["1", "2", "12"])
def test_set_apply_matches_reference_fixture(fqe, suffix):
    print_case("fixture apply", suffix=suffix)
    z = numpy.load(GOLDEN)
    norb = 2
    keys = [(2, 0), (1, 1), (0, 2)]
    tag = {k: f"{k[0]:02d}{k[1]:02d}{norb:02d}" for k in keys}
    wset = make_set(fqe, 2, norb, {k: z[f"c_{tag[k]}"] for k in keys})
    if suffix == "1":
        tensors = (z["h1"],)
    elif suffix == "2":
        tensors = (numpy.zeros_like(z["h1"]), z["h2"])
    else:
        tensors = (z["h1"], z["h2"])
    out = wset.apply(tensors)
    for k in keys:
        want = z[f"out{suffix}_{tag[k]}"]
        assert numpy.allclose(out.sectors()[k].coeff, want, rtol=1e-12,
                              atol=1e-12)
    # the input is unchanged
    for k in keys:
        assert numpy.array_equal(wset.sectors()[k].coeff, z[f"c_{tag[k]}"])


def test_set_rdm12_matches_reference_fixture(fqe):
    print_case("fixture rdm12")
    z = numpy.load(GOLDEN)
    norb = 2
    keys = [(2, 0), (1, 1), (0, 2)]
    wset = make_set(fqe, 2, norb, {
        k: z[f"c_{k[0]:02d}{k[1]:02d}{norb:02d}"] for k in keys})
    d1, d2 = wset.rdm12()
    assert numpy.allclose(d1, z["rdm1"], rtol=1e-12, atol=1e-12)
    assert numpy.allclose(d2, z["rdm2"], rtol=1e-12, atol=1e-12)
    (r1,) = wset.rdm1()
    assert numpy.allclose(r1, z["rdm1"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nele,norb", [(1, 1), (2, 1), (3, 3), (2, 4),
                                       (4, 4), (5, 4), (6, 5), (7, 6)])
def test_set_apply_matches_oracle(fqe, nele, norb):
    print_case("oracle apply", nele=nele, norb=norb)
    rng = numpy.random.default_rng(1000 + 31 * nele + norb)
    sg = so.SetGraph(nele, norb)
    coeffs = random_coeffs(sg, rng)
    h1, h2 = random_tensors(norb, rng)
    wset = make_set(fqe, nele, norb, coeffs)
    got12 = wset.apply((h1, h2))
    got1 = wset.apply((h1,))
    want12 = so.set_apply(coeffs, (h1, h2), sg)
    want1 = so.set_apply(coeffs, (h1,), sg)
    for k in sg.keys:
        assert rel(got12.sectors()[k].coeff, want12[k]) < TOL
        assert rel(got1.sectors()[k].coeff, want1[k]) < TOL


@pytest.mark.parametrize("nele,norb", [(3, 3), (4, 4), (3, 5)])
def test_set_rdms_match_oracle(fqe, nele, norb):
    print_case("oracle rdm", nele=nele, norb=norb)
    rng = numpy.random.default_rng(77 + nele * norb)
    sg = so.SetGraph(nele, norb)
    ket = random_coeffs(sg, rng)
    bra = random_coeffs(sg, rng)
    wket = make_set(fqe, nele, norb, ket)
    wbra = make_set(fqe, nele, norb, bra)
    d1, d2 = wket.rdm12()
    w1, w2 = so.set_rdm12(ket, sg)
    assert rel(d1, w1) < TOL and rel(d2, w2) < TOL
    t1, t2 = wket.rdm12(wbra)
    v1, v2 = so.set_rdm12(ket, sg, bra=bra)
    assert rel(t1, v1) < TOL and rel(t2, v2) < TOL


def test_set_apply_is_deterministic_and_hermitian_at_norb10(fqe):
    # C(20, 10) = 184756 determinants over 11 sectors; (2 norb)^2 = 400
    # pairs.  The oracle would take minutes here: check <x|H y> = <H x|y>
    # for a Hermitian spin-mixing H, linearity, and bit-identical reruns.
    nele, norb = 10, 10
    print_case("hermitian", nele=nele, norb=norb)
    rng = numpy.random.default_rng(2024)
    keys = [(nele - nb, nb) for nb in range(nele + 1)]
    from math import comb
    shapes = {k: (comb(norb, k[0]), comb(norb, k[1])) for k in keys}
    x = {k: rng.standard_normal(s) + 1j * rng.standard_normal(s)
         for k, s in shapes.items()}
    y = {k: rng.standard_normal(s) + 1j * rng.standard_normal(s)
         for k, s in shapes.items()}
    h1, h2 = random_tensors(norb, rng, hermitian=True)
    wx = make_set(fqe, nele, norb, x)
    wy = make_set(fqe, nele, norb, y)
    hx = wx.apply((h1, h2))
    hy = wy.apply((h1, h2))
    hy2 = wy.apply((h1, h2))
    lhs = sum(numpy.vdot(x[k], hy.sectors()[k].coeff) for k in keys)
    rhs = sum(numpy.vdot(hx.sectors()[k].coeff, y[k]) for k in keys)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    for k in keys:
        assert numpy.array_equal(hy.sectors()[k].coeff, hy2.sectors()[k].coeff)
    # linearity: H (x + 2y) = Hx + 2 Hy
    xy = make_set(fqe, nele, norb, {k: x[k] + 2.0 * y[k] for k in keys})
    hxy = xy.apply((h1, h2))
    num = sum(numpy.linalg.norm(hxy.sectors()[k].coeff - hx.sectors()[k].coeff
                                - 2.0 * hy.sectors()[k].coeff) ** 2
              for k in keys) ** 0.5
    den = sum(numpy.linalg.norm(hxy.sectors()[k].coeff) ** 2
              for k in keys) ** 0.5
    assert num <= 1e-12 * den


def test_set_errors(fqe):
    print_case("errors")
    norb = 2
    data = {(2, 0): fqe.FqeData(2, 0, norb), (1, 1): fqe.FqeData(1, 1, norb)}
    wset = fqe.FqeDataSet(2, norb, data)
    with pytest.raises(ValueError, match="needs the sectors"):
        wset.apply((numpy.eye(4),))
    full = make_set(fqe, 2, norb, {(2, 0): numpy.ones((1, 1)),
                                   (1, 1): numpy.ones((2, 2)),
                                   (0, 2): numpy.ones((1, 1))})
    with pytest.raises(ValueError, match="h1 must have shape"):
        full.apply((numpy.eye(3),))
    m = 2 * norb
    with pytest.raises(NotImplementedError):
        full.apply((numpy.eye(m), numpy.zeros((m,) * 4),
                    numpy.zeros((m,) * 6)))


def test_set_too_large_raises_memory_error(fqe):
    """norb = nele = 14: 40 M determinants over 15 sectors, so the two-body
    apply's (2 norb)^2-pair workspace (about 1 TB) and the RDM dvec exceed
    HBM; both raise MemoryError before allocating anything of that size."""
    print_case("too large", nele=14, norb=14)
    norb = nele = 14
    data = {(nele - nb, nb): fqe.FqeData(nele - nb, nb, norb)
            for nb in range(nele + 1)}
    wset = fqe.FqeDataSet(nele, norb, data)
    m = 2 * norb
    with pytest.raises(MemoryError, match="no chunking over pairs"):
        wset.apply((numpy.zeros((m, m)), numpy.zeros((m,) * 4)))
    with pytest.raises(MemoryError, match="spin-orbital dvec"):
        wset.rdm1()


# --------------------------------------------------------------------------
# Wavefunction-level routing: General / GSOHamiltonian on spin-broken
# wavefunctions go through FqeDataSet per particle number (reference
# src/fqe/wavefunction.py:315-364, :366-399, :442-471).
# --------------------------------------------------------------------------

def spin_broken_wfn(fqe, nele, norb, coeffs):
    keys = sorted(coeffs)
    wfn = fqe.Wavefunction([[nele, k[0] - k[1], norb] for k in keys],
                           broken=['spin'])
    for k in keys:
        wfn.sector((nele, k[0] - k[1])).coeff = numpy.array(
            coeffs[k], dtype=numpy.complex128)
    return wfn


def test_gso_apply_on_wavefunction_matches_fixture(fqe):
    print_case("wavefunction gso apply")
    z = numpy.load(GOLDEN)
    norb = 2
    keys = [(2, 0), (1, 1), (0, 2)]
    coeffs = {k: z[f"c_{k[0]:02d}{k[1]:02d}{norb:02d}"] for k in keys}
    wfn = spin_broken_wfn(fqe, 2, norb, coeffs)
    e0 = 0.3 - 0.1j
    out = fqe.apply(fqe.get_gso_hamiltonian((z["h1"], z["h2"]), e_0=e0), wfn)
    for k in keys:
        want = z[f"out12_{k[0]:02d}{k[1]:02d}{norb:02d}"] + e0 * coeffs[k]
        got = out.sector((2, k[0] - k[1])).coeff
        assert numpy.allclose(got, want, rtol=1e-12, atol=1e-12)
    # a bare tuple of spin-orbital tensors builds a General Hamiltonian
    out2 = wfn.apply((z["h1"], z["h2"]))
    for k in keys:
        assert numpy.allclose(out2.sector((2, k[0] - k[1])).coeff,
                              z[f"out12_{k[0]:02d}{k[1]:02d}{norb:02d}"],
                              rtol=1e-12, atol=1e-12)


def test_gso_time_evolve_matches_dense_exponential(fqe):
    print_case("wavefunction gso time_evolve")
    import scipy.linalg
    from oracle import jw_dense
    nele, norb, t = 3, 3, 0.4
    rng = numpy.random.default_rng(31)
    sg = so.SetGraph(nele, norb)
    coeffs = random_coeffs(sg, rng)
    norm = numpy.sqrt(sum(numpy.linalg.norm(c) ** 2 for c in coeffs.values()))
    coeffs = {k: c / norm for k, c in coeffs.items()}
    h1, h2 = random_tensors(norb, rng, hermitian=True)
    h1, h2 = 0.2 * h1, 0.05 * h2
    wfn = spin_broken_wfn(fqe, nele, norb, coeffs)
    out = fqe.time_evolve(wfn, t, fqe.get_gso_hamiltonian((h1, h2)))
    ph = jw_dense.set_sector_phase(norb, nele)
    mat = jw_dense.dense_set_operator(h1, h2, norb, nele)
    mat = ph[:, None] * mat * ph[None, :]
    vec = numpy.concatenate([coeffs[k].ravel() for k in sg.keys])
    want = scipy.linalg.expm(-1j * t * mat) @ vec
    got = numpy.concatenate([out.sector((nele, k[0] - k[1])).coeff.ravel()
                             for k in sg.keys])
    assert rel(got, want) < 1e-11


def test_gso_routing_errors(fqe):
    print_case("wavefunction gso errors")
    norb = 2
    m = 2 * norb
    ham = fqe.get_gso_hamiltonian((numpy.eye(m),))
    partial = fqe.Wavefunction([[2, 2, norb], [2, 0, norb]], broken=['spin'])
    with pytest.raises(ValueError, match="spin complete"):
        fqe.apply(ham, partial)
    # on a spin-conserving wavefunction the sectors take the spin-orbital
    # apply one by one, as in the reference: the identity h1 counts the
    # electrons; 3-body tensors are outside the path
    conserving = fqe.Wavefunction([[2, 0, norb]])
    conserving.set_wfn(strategy="random")
    out = fqe.apply(ham, conserving)
    assert numpy.allclose(out.get_coeff((2, 0)),
                          2.0 * conserving.get_coeff((2, 0)), atol=1e-13)
    with pytest.raises(NotImplementedError):
        fqe.apply(fqe.get_gso_hamiltonian((numpy.zeros((m, m)),
                                           numpy.zeros((m,) * 4),
                                           numpy.zeros((m,) * 6))),
                  conserving)
    full = fqe.Wavefunction([[2, 2, norb], [2, 0, norb], [2, -2, norb]],
                            broken=['spin'])
    with pytest.raises(ValueError, match="incorrect size"):
        fqe.apply(fqe.get_gso_hamiltonian((numpy.eye(3),)), full)
