"""Pins the CPU oracle before anything is checked against it (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails at least one of
them: literal tables from the reference tests, FQE's float64 golden apply
outputs, a closed-form eigenvector, an independent Jordan-Wigner matrix,
exact exponentials, and the reference's own compiled C.
"""

import json
import math
import os

import itertools

import numpy
import pytest
import scipy.linalg

from oracle import fqe_oracle as fo
from oracle import jw_dense
from oracle import ref_lm

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def literal():
    with open(os.path.join(GOLDEN, "literal_tables.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def golden():
    return numpy.load(os.path.join(GOLDEN, "fqe_data_golden.npz"))


# --------------------------------------------------------------------------
# (a) string tables
# --------------------------------------------------------------------------

def test_strings_match_reference_literal(literal):
    """tests/fci_graph_test.py:33-112 (norb=8, nalpha=4)."""
    strings = fo.build_strings(8, 4)
    assert strings.tolist() == literal["strings_n8_k4"]
    index = {str(int(s)): a for a, s in enumerate(strings)}
    assert index == literal["index_n8_k4"]
    # addresses quoted in tests/fci_graph_test.py:120-121
    z = fo.z_matrix(8, 4)
    assert fo.z_address(0b1111, z) == 0
    assert fo.z_address((1 << 1) | (1 << 2) | (1 << 3) | (1 << 7), z) == 38


def test_maps_match_reference_literal(literal):
    """tests/fci_graph_test.py:151-190 (norb=4, nalpha=2, nbeta=1)."""
    g = fo.OracleGraph(2, 1, 4)
    for key, rows in literal["maps_n4_k2"].items():
        i, j = map(int, key.split(","))
        assert g.alpha_map[(i, j)].tolist() == [list(r) for r in rows]
    for key, rows in literal["maps_n4_k1"].items():
        i, j = map(int, key.split(","))
        assert g.beta_map[(i, j)].tolist() == [list(r) for r in rows]


@pytest.mark.parametrize("norb,nele", [(1, 1), (5, 0), (6, 3), (9, 4),
                                       (10, 7), (12, 6)])
def test_z_matrix_address_equals_lexicographic_rank(norb, nele):
    strings = fo.build_strings(norb, nele)
    z = fo.z_matrix(norb, nele)
    assert strings.size == math.comb(norb, nele)
    for addr, s in enumerate(strings):
        assert fo.z_address(int(s), z) == addr


def test_dexc_shape_and_order():
    g = fo.OracleGraph(3, 2, 6)
    for dexc, nele in ((g.dexca, 3), (g.dexcb, 2)):
        assert dexc.shape[1] == nele * (6 - nele + 1)
        assert numpy.all(numpy.diff(dexc[:, :, 1], axis=1) > 0)


needs_ref = pytest.mark.skipif(not ref_lm.available(),
                               reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("na,nb,norb", [(4, 3, 8), (4, 4, 6), (0, 3, 7),
                                        (2, 0, 6), (5, 5, 10), (3, 6, 9)])
def test_tables_bit_exact_vs_reference_c(na, nb, norb):
    g = fo.OracleGraph(na, nb, norb)
    r = ref_lm.RefGraph(na, nb, norb)
    assert numpy.array_equal(g.astr, r.astr)
    assert numpy.array_equal(g.bstr, r.bstr)
    for key in g.alpha_map:
        assert numpy.array_equal(g.alpha_map[key], r.alpha_map[key])
        assert numpy.array_equal(g.beta_map[key], r.beta_map[key])
    assert numpy.array_equal(g.dexca, r.dexca)
    assert numpy.array_equal(g.dexcb, r.dexcb)


# --------------------------------------------------------------------------
# apply
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb", [(2, 3, 6), (2, 1, 4), (1, 1, 2)])
def test_apply_matches_fqe_golden(golden, na, nb, norb):
    """FQE's float64 outputs of apply((h1,)), apply((0,h2)), apply((h1,h2))
    (tests/unittest_data/fqe_data/generate_data.py:68-95)."""
    tag = f"{na:02d}{nb:02d}{norb:02d}"
    g = fo.OracleGraph(na, nb, norb)
    c, h1, h2 = golden["c_" + tag], golden["h1_" + tag], golden["h2_" + tag]
    cases = {"1": fo.apply_1(c, h1, g),
             "2": fo.apply_12(c, numpy.zeros_like(h1), h2, g),
             "12": fo.apply_12(c, h1, h2, g)}
    for suffix, out in cases.items():
        ref = golden[f"out{suffix}_{tag}"]
        assert numpy.linalg.norm(out - ref) <= 1e-13 * numpy.linalg.norm(ref)


def test_apply_closed_form_eigenvector(literal):
    """tests/fqe_data_test.py:761-826: apply(h) == lambda * wfn."""
    norb = 4
    h2 = numpy.zeros((norb,) * 4)
    for i, j, k, l in numpy.ndindex(h2.shape):
        h2[i, j, k, l] = (i + k) * (j + l) * 0.02
    wfn = numpy.array(literal["eigvec_2_2_4"])
    wfn = wfn[..., 0] + 1j * wfn[..., 1]
    g = fo.OracleGraph(2, 2, norb)
    out = fo.apply_12(wfn, numpy.zeros((norb, norb)), h2, g)
    numpy.testing.assert_allclose(out, literal["eigval_2_2_4"] * wfn,
                                  atol=1e-12)


def _random_h(norb, rng, complex_h=True, symmetric=False):
    h1 = rng.normal(size=(norb, norb))
    h2 = rng.normal(size=(norb,) * 4)
    if complex_h:
        h1 = h1 + 1j * rng.normal(size=(norb, norb))
        h2 = h2 + 1j * rng.normal(size=(norb,) * 4)
    if symmetric:
        h1 = h1 + h1.T
        h2 = h2 + h2.transpose(2, 1, 0, 3)  # i <-> k
        h2 = h2 + h2.transpose(0, 3, 2, 1)  # j <-> l
        h2 = h2 + h2.transpose(1, 0, 3, 2)
    return h1, h2


@pytest.mark.parametrize("na,nb,norb", [(2, 2, 4), (1, 2, 3), (0, 2, 4),
                                        (3, 0, 4), (2, 3, 5), (4, 4, 4)])
@pytest.mark.parametrize("complex_h,symmetric", [(True, False),
                                                 (False, True)])
def test_apply_matches_jordan_wigner(na, nb, norb, complex_h, symmetric):
    rng = numpy.random.default_rng(17 + 100 * na + 10 * nb + norb)
    h1, h2 = _random_h(norb, rng, complex_h, symmetric)
    g = fo.OracleGraph(na, nb, norb)
    dense = jw_dense.dense_hamiltonian(h1, h2, norb, na, nb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    out = fo.apply_12(c, h1, h2, g)
    ref = (dense @ c.reshape(-1)).reshape(c.shape)
    assert numpy.linalg.norm(out - ref) <= 1e-12 * numpy.linalg.norm(ref)
    out1 = fo.apply_1(c, h1, g)
    ref1 = (jw_dense.dense_hamiltonian(h1, None, norb, na, nb)
            @ c.reshape(-1)).reshape(c.shape)
    assert numpy.linalg.norm(out1 - ref1) <= 1e-12 * numpy.linalg.norm(ref1)


def test_planted_transpose_is_caught():
    """A plausible oracle mistake (V built with the wrong axis swap) must
    fail the Jordan-Wigner check; the mistake is planted in a copy of the
    oracle's tensor preparation, on the CPU."""
    rng = numpy.random.default_rng(3)
    norb, na, nb = 4, 2, 1
    h1, h2 = _random_h(norb, rng, complex_h=True)
    g = fo.OracleGraph(na, nb, norb)
    dense = jw_dense.dense_hamiltonian(h1, h2, norb, na, nb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = (dense @ c.reshape(-1)).reshape(c.shape)
    good = fo.apply_12(c, h1, h2, g)
    bad = fo.apply_12(c, h1, numpy.moveaxis(h2, 1, 3), g)
    assert numpy.linalg.norm(good - ref) < 1e-12 * numpy.linalg.norm(ref)
    assert numpy.linalg.norm(bad - ref) > 1e-3 * numpy.linalg.norm(ref)


@needs_ref
@pytest.mark.parametrize("na,nb,norb,seed", [(4, 4, 8, 0), (3, 5, 8, 7),
                                             (5, 5, 10, 1), (0, 4, 7, 2)])
def test_apply_matches_reference_lm_kernels(na, nb, norb, seed):
    """The reference's default C path (lm kernels, complex128)."""
    c, h1, h2 = fo.synthetic_workload(norb, na, nb, seed, norb)
    g = fo.OracleGraph(na, nb, norb)
    out = fo.apply_12(c, h1, h2, g)
    ref = ref_lm.apply_12_lm(c, h1, h2, g)
    assert numpy.linalg.norm(out - ref) <= 1e-13 * numpy.linalg.norm(ref)
    # complex, non-symmetric h as well
    rng = numpy.random.default_rng(seed)
    h1c, h2c = _random_h(norb, rng, complex_h=True)
    out = fo.apply_12(c, h1c, h2c, g)
    ref = ref_lm.apply_12_lm(c, h1c, h2c, g)
    assert numpy.linalg.norm(out - ref) <= 1e-13 * numpy.linalg.norm(ref)


# --------------------------------------------------------------------------
# Taylor time evolution
# --------------------------------------------------------------------------

def test_taylor_matches_exact_exponential():
    norb, na, nb = 4, 2, 2
    rng = numpy.random.default_rng(5)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    dense = jw_dense.dense_hamiltonian(h1, h2, norb, na, nb)
    assert numpy.allclose(dense, dense.conj().T)
    e0 = 0.3 - 0.1j
    t = 0.1
    out, nterms = fo.taylor_evolve(c, (h1, h2), e0, t, g)
    exact = scipy.linalg.expm(-1j * t * dense) @ c.reshape(-1)
    exact *= numpy.exp(-1j * t * e0) ** 2  # the reference's doubled phase
    assert numpy.linalg.norm(out.reshape(-1) - exact) < 1e-12
    assert 2 <= nterms <= 29
    # unitarity of the propagator (e0 phase aside)
    out0, _ = fo.taylor_evolve(c, (h1, h2), 0.0, t, g)
    assert abs(numpy.linalg.norm(out0) - 1.0) < 1e-13


def test_taylor_raises_when_not_converged():
    g = fo.OracleGraph(1, 1, 2)
    h1 = numpy.eye(2)
    h2 = numpy.zeros((2,) * 4)
    c = numpy.ones((2, 2), dtype=complex) / 2
    with pytest.raises(RuntimeError):
        fo.taylor_evolve(c, (h1, h2), 0.0, 1.0, g, expansion=2)


# --------------------------------------------------------------------------
# Slater-Condon diagonal
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb", [(2, 2, 4), (3, 1, 5), (0, 2, 4)])
def test_slater_condon_matches_dense_diagonal(na, nb, norb):
    rng = numpy.random.default_rng(na * 7 + nb)
    h1, h2 = _random_h(norb, rng, complex_h=False)
    g = fo.OracleGraph(na, nb, norb)
    dense = jw_dense.dense_hamiltonian(h1, h2, norb, na, nb)
    diag = fo.slater_condon_diagonal(h1, h2, g)
    numpy.testing.assert_allclose(diag.reshape(-1), dense.diagonal().real,
                                  atol=1e-11)


# --------------------------------------------------------------------------
# learned-preconditioner oracle (oracle/mlp_oracle.py)
# --------------------------------------------------------------------------

def test_mlp_oracle_gradient_matches_finite_differences():
    """Back-propagation of the oracle against central differences in
    float64 over every parameter block (a transposed weight or a dropped
    tanh' factor fails it)."""
    from oracle import mlp_oracle as mo
    rng = numpy.random.default_rng(0)
    theta = mo.init_params(1).astype(numpy.float64)
    theta += 0.05 * rng.normal(size=theta.size)
    x = rng.normal(size=(7, mo.NFEAT))
    t = rng.normal(size=7)
    w = rng.uniform(0.1, 1.0, size=7)
    loss, grad = mo.loss_and_grad(theta, x, t, w)
    shapes = mo.param_shapes()
    off = 0
    for name, shape in shapes:
        size = int(numpy.prod(shape))
        picks = rng.choice(size, size=min(size, 12), replace=False)
        for q in picks:
            e = numpy.zeros_like(theta)
            e[off + q] = 1e-6
            num = (mo.loss_and_grad(theta + e, x, t, w)[0] -
                   mo.loss_and_grad(theta - e, x, t, w)[0]) / 2e-6
            assert abs(num - grad[off + q]) < 1e-6 * max(1.0, abs(num)), (name, q)
        off += size
    assert abs(loss - numpy.sum(w * (mo.forward(theta, x) - t) ** 2)) < 1e-12


def test_mlp_oracle_features_by_hand():
    """Features of two determinants of the norb=4, 2a1b sector, written out
    by hand."""
    from oracle import mlp_oracle as mo
    g = fo.OracleGraph(2, 1, 4)
    diag = numpy.arange(g.lena * g.lenb, dtype=float).reshape(g.lena, g.lenb)
    x = mo.features(g.astr, g.bstr, diag, theta=-0.5, dmin=0.0, ref_a=0b0011,
                    ref_b=0b0001, norb=4).reshape(g.lena, g.lenb, -1)
    # alpha strings: 0b0011, 0b0101, ...; beta strings 0b0001, 0b0010, ...
    ia = list(g.astr).index(0b0101)
    ib = list(g.bstr).index(0b0010)
    d = diag[ia, ib]
    expect = [d + 0.5, 1.0 / (d + 0.5 + 0.05), d, -0.5,
              1.0, 1.0,          # excitation levels
              0.0, 1.0,          # alpha HOMO (1) empty, LUMO (2) occupied
              0.0, 1.0]          # beta HOMO (0) empty, LUMO (1) occupied
    assert numpy.allclose(x[ia, ib], expect)
    assert numpy.allclose(x[0, 0, 4:], [0, 0, 1, 0, 1, 0])


def test_mlp_oracle_targets_and_adam():
    from oracle import mlp_oracle as mo
    rng = numpy.random.default_rng(3)
    psi = rng.normal(size=6) + 1j * rng.normal(size=6)
    u = psi + 0.1 * (rng.normal(size=6) + 1j * rng.normal(size=6))
    r = rng.normal(size=6) + 1j * rng.normal(size=6)
    inv = rng.uniform(0.2, 1.0, size=6)
    tgt, w = mo.targets(u, r, psi, inv, clip=1e9)
    tstar = psi / numpy.vdot(u, psi) - u
    # the target multiplier minimises |m r_I + t*_I| over real m
    for i in range(6):
        ms = numpy.linspace(-50, 50, 200001)
        best = ms[numpy.argmin(numpy.abs(ms * r[i] + tstar[i]))]
        assert abs(tgt[i] + inv[i] - best) < 1e-3
    assert abs(w.sum() - 1.0) < 1e-12
    # first Adam step moves every parameter by lr against its gradient sign
    p = numpy.zeros(4)
    gr = numpy.array([1.0, -2.0, 0.5, 0.0])
    q, _, _ = mo.adam_step(p, gr, numpy.zeros(4), numpy.zeros(4), 1, 0.01)
    assert numpy.allclose(q, [-0.01, 0.01, -0.01, 0.0], atol=1e-9)


def test_lih_fixture_energies():
    """FQE's LiH data: <psi|H|psi> is the FCI energy of
    tests/davidson_test.py:27, the lowest dense eigenvalue equals it, and
    the Slater-Condon diagonal at the HF determinant is the HF energy of
    tests/wavefunction_test.py:537-548."""
    z = numpy.load(os.path.join(os.path.dirname(__file__), "golden",
                                "lih_energy.npz"))
    h1, h2, wfn = z["h1e"], z["h2e"], z["wfn"]
    g = fo.OracleGraph(2, 2, 6)
    e = numpy.vdot(wfn, fo.apply_12(wfn, h1, h2, g)) / numpy.vdot(wfn, wfn)
    assert abs(e - (-8.877719570384043)) < 1e-10
    w = numpy.linalg.eigvalsh(fo.dense_matrix(h1, h2, g))
    assert abs(w[0] - (-8.877719570384043)) < 1e-10
    d = fo.slater_condon_diagonal(h1, h2, g)
    assert abs(d[0, 0] - (-8.857341498221992)) < 1e-12


def test_rdm12_oracle_matches_fqe_golden_and_energy(golden):
    """The oracle RDMs reproduce FQE's stored rdm1/rdm12 of FqeData(2,1,4)
    (tests/fqe_data_test.py:669-700) and contract with h1 / h2 to the
    energy of the apply."""
    g = fo.OracleGraph(2, 1, 4)
    c = golden["c_020104"]
    r1, r2 = fo.rdm12(c, g)
    assert numpy.max(numpy.abs(r1 - golden["rdm1_020104"])) < 1e-13
    assert numpy.max(numpy.abs(r2 - golden["rdm2_020104"])) < 1e-13
    rng = numpy.random.default_rng(4)
    h2 = rng.normal(size=(4,) * 4)
    h1 = rng.normal(size=(4, 4))
    e = numpy.vdot(c, fo.apply_12(c, h1, h2, g))
    assert abs(numpy.sum(h1 * r1) + numpy.einsum("ijkl,ijkl->", h2, r2) - e) < 1e-12


def _chemist_dense(h1, eri, norb, na, nb):
    """Dense H of the textbook chemist-notation Hamiltonian
    sum_pq h1[p,q] sum_s a+_ps a_qs
      + 1/2 sum_pqrs (pq|rs) sum_{s,t} a+_ps a+_rt a_st a_qs
    on FQE's (na, nb) sector basis, built operator string by operator string
    (independent of FQE's h2 convention)."""
    astr = jw_dense._strings(norb, na)
    bstr = jw_dense._strings(norb, nb)
    lenb = len(bstr)
    basis = {sa | (sb << norb): ia * lenb + ib
             for ia, sa in enumerate(astr) for ib, sb in enumerate(bstr)}
    mat = numpy.zeros((len(basis), len(basis)), dtype=numpy.complex128)
    terms = []
    for p_ in range(norb):
        for q in range(norb):
            for s_ in range(2):
                terms.append((h1[p_, q], ((p_ + s_ * norb, True),
                                          (q + s_ * norb, False))))
    for p_, q, r, t in itertools.product(range(norb), repeat=4):
        for s_ in range(2):
            for u in range(2):
                terms.append((0.5 * eri[p_, q, r, t],
                              ((p_ + s_ * norb, True), (r + u * norb, True),
                               (t + u * norb, False), (q + s_ * norb, False))))
    for state, col in basis.items():
        for coef, ops in terms:
            res = jw_dense._apply_string(state, ops)
            if res is not None:
                mat[basis[res[0]], col] += res[1] * coef
    return mat


@pytest.mark.parametrize("norb,na,nb,seed", [(3, 2, 1, 7), (4, 2, 1, 8),
                                             (4, 2, 2, 9)])
def test_physics_to_fqe_h2_is_the_chemist_hamiltonian(norb, na, nb, seed):
    """The conversion of the synthetic ERIs (ij|kl) to FQE's restricted h2
    (physics_to_fqe_h2, after the reference's integrals_to_fqe_restricted)
    gives the Hamiltonian of the textbook chemist-notation formula, built
    here from its own operator strings; a transposed index, a dropped 1/2
    or a flipped sign in the conversion fails this."""
    rng = numpy.random.default_rng(seed)
    h1, _ = fo.synthetic_hamiltonian(norb, 2, rng)
    # ERIs with the 8-fold symmetry of real orbitals, as in the workload
    eri = numpy.zeros((norb,) * 4)
    for _ in range(3):
        lk = rng.normal(0.0, 0.1, size=(norb, norb))
        lk = lk + lk.T
        eri += numpy.einsum("ij,kl->ijkl", lk, lk)
    ref = _chemist_dense(h1, eri, norb, na, nb)
    h2 = fo.physics_to_fqe_h2(eri)
    dense = jw_dense.dense_hamiltonian(h1, h2, norb, na, nb)
    assert numpy.max(numpy.abs(dense - ref)) < 1e-13
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    got = fo.apply_12(c, h1, h2, g).reshape(-1)
    assert numpy.linalg.norm(got - ref @ c.reshape(-1)) < 1e-12 * numpy.linalg.norm(got)


@pytest.mark.parametrize("norb,na,nb,seed,k", [(8, 4, 4, 0, 8), (6, 2, 3, 5, 3)])
def test_product_synthetic_inputs_equal_oracle_convention(norb, na, nb, seed, k):
    """The benchmark's input generator (fqe_amd.synthetic) draws exactly the
    arrays of the oracle's documented convention."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..",
                                    "openfermion-fqe_amd"))
    from fqe_amd import synthetic
    got = synthetic.workload(norb, na, nb, seed, k)
    ref = fo.synthetic_workload(norb, na, nb, seed, k)
    for a, b in zip(got, ref):
        assert numpy.allclose(a, b, rtol=0, atol=1e-15)
