"""GPU parity of the HIP engine against the oracle and the reference's
golden data.  Every test runs the HIP kernels through the C ABI.

Tolerances: FP64 engine results agree with the complex128 oracle to a
normwise relative error of 1e-12 (tables bit-exact); the mixed engine
(complex64 storage, BF16 hi/lo matrix-core GEMM) to 2e-4.
"""

import copy
import json
import os

import numpy
import pytest

from oracle import fqe_oracle as fo
from oracle import ref_lm

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FP64_TOL = 1e-12
MIXED_TOL = 2e-4


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return numpy.linalg.norm(numpy.asarray(a) - numpy.asarray(b)) / max(
        numpy.linalg.norm(numpy.asarray(b)), 1e-300)


# --------------------------------------------------------------------------
# (a) tables
# --------------------------------------------------------------------------

@pytest.mark.parametrize("norb,nele", [(1, 0), (1, 1), (4, 2), (6, 3),
                                       (8, 4), (7, 0), (7, 7), (9, 2),
                                       (10, 5)])
def test_string_tables_bit_exact(fqe, norb, nele):
    print(f"tables norb={norb} nele={nele}", flush=True)
    g = fqe.FciGraph(nele, 0, norb)
    ref = fo.OracleGraph(nele, 0, norb)
    assert numpy.array_equal(g._astr, ref.astr)
    assert g._aind == {int(s): a for a, s in enumerate(ref.astr)}
    if nele:
        assert numpy.array_equal(g._dexca, ref.dexca)
    for key, val in ref.alpha_map.items():
        assert numpy.array_equal(g._alpha_map[key], val), key


def test_string_tables_match_reference_literal(fqe):
    with open(os.path.join(GOLDEN, "literal_tables.json")) as f:
        lit = json.load(f)
    g = fqe.FciGraph(4, 0, 8)
    assert g._astr.tolist() == lit["strings_n8_k4"]
    assert {str(k): v for k, v in g._aind.items()} == lit["index_n8_k4"]
    g = fqe.FciGraph(2, 1, 4)
    for key, rows in lit["maps_n4_k2"].items():
        i, j = map(int, key.split(","))
        assert g.alpha_map(i, j).tolist() == [list(r) for r in rows]
    for key, rows in lit["maps_n4_k1"].items():
        i, j = map(int, key.split(","))
        assert g.beta_map(i, j).tolist() == [list(r) for r in rows]


@pytest.mark.parametrize("compressed", [False, True])
def test_pair_table(fqe, compressed):
    norb, nele = 8, 3
    g = fqe.FciGraph(nele, 0, norb)
    tab = g.tables(True).pair_table(compressed).cpu().numpy()
    ref = fo.OracleGraph(nele, 0, norb)
    npair = norb * (norb + 1) // 2 if compressed else norb * norb
    expect = numpy.zeros((npair, ref.lena), dtype=numpy.int32)
    for t in range(ref.lena):
        for src, ij, sign in ref.dexca[t]:
            i, j = divmod(int(ij), norb)
            p = (min(i, j) + max(i, j) * (max(i, j) + 1) // 2
                 if compressed else ij)
            assert expect[p, t] == 0
            expect[p, t] = sign * (src + 1)
    assert numpy.array_equal(tab, expect)


# --------------------------------------------------------------------------
# (b) dvec
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb", [(2, 3, 6), (3, 3, 6), (0, 2, 5),
                                        (4, 1, 7)])
def test_dvec_matches_oracle(fqe, na, nb, norb):
    rng = numpy.random.default_rng(na + 10 * nb + 100 * norb)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    d = fqe.FqeData(na, nb, norb)
    d.coeff = c
    assert rel(d.calculate_dvec_spatial(), fo.dvec(c, g)) < FP64_TOL


# --------------------------------------------------------------------------
# full apply
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb", [(2, 3, 6), (2, 1, 4), (1, 1, 2)])
def test_apply_matches_fqe_golden(fqe, na, nb, norb):
    """FQE's own outputs (tests/unittest_data/fqe_data) for apply((h1,)),
    apply((0, h2)), apply((h1, h2)) and the complex variant of
    tests/fqe_data_test.py:541-568."""
    z = numpy.load(os.path.join(GOLDEN, "fqe_data_golden.npz"))
    tag = f"{na:02d}{nb:02d}{norb:02d}"
    c, h1, h2 = z["c_" + tag], z["h1_" + tag], z["h2_" + tag]
    d = fqe.FqeData(na, nb, norb)
    d.coeff = c
    assert rel(d.apply((h1,)).coeff, z["out1_" + tag]) < FP64_TOL
    assert rel(d.apply((numpy.zeros_like(h1), h2)).coeff,
               z["out2_" + tag]) < FP64_TOL
    assert rel(d.apply((h1, h2)).coeff, z["out12_" + tag]) < FP64_TOL
    h1c = h1.astype(numpy.complex128) + 1e-7j
    out = d.apply((h1c, h2.astype(numpy.complex128))).coeff
    assert numpy.allclose(out, z["out12_" + tag])
    g = fo.OracleGraph(na, nb, norb)
    assert rel(out, fo.apply_12(c, h1c, h2, g)) < FP64_TOL


def test_apply_closed_form_eigenvector(fqe):
    """tests/fqe_data_test.py:761-826"""
    with open(os.path.join(GOLDEN, "literal_tables.json")) as f:
        lit = json.load(f)
    norb = 4
    h2 = numpy.zeros((norb,) * 4, dtype=numpy.complex128)
    for i, j, k, l in numpy.ndindex(h2.shape):
        h2[i, j, k, l] = (i + k) * (j + l) * 0.02
    wfn = numpy.array(lit["eigvec_2_2_4"])
    wfn = wfn[..., 0] + 1j * wfn[..., 1]
    d = fqe.FqeData(2, 2, norb)
    d.coeff = wfn
    out = d.apply((numpy.zeros((norb, norb), dtype=numpy.complex128), h2))
    assert numpy.allclose(out.coeff, lit["eigval_2_2_4"] * wfn, atol=1e-12)


def _h(norb, rng, kind):
    h1 = rng.normal(size=(norb, norb))
    h2 = rng.normal(size=(norb,) * 4)
    if kind == "complex":
        h1 = h1 + 1j * rng.normal(size=(norb, norb))
        h2 = h2 + 1j * rng.normal(size=(norb,) * 4)
    elif kind == "symmetric":
        return fo.synthetic_hamiltonian(norb, norb, rng)
    return h1, h2


@pytest.mark.parametrize("kind", ["real", "complex", "symmetric"])
@pytest.mark.parametrize("na,nb,norb", [(2, 2, 4), (3, 2, 6), (0, 3, 5),
                                        (4, 4, 4), (3, 4, 8)])
def test_apply_matches_oracle(fqe, kind, na, nb, norb):
    print(f"apply {kind} {na} {nb} {norb}", flush=True)
    rng = numpy.random.default_rng(norb * 31 + na * 7 + nb)
    h1, h2 = _h(norb, rng, kind)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    d = fqe.FqeData(na, nb, norb)
    d.coeff = c
    assert rel(d.apply((h1, h2)).coeff, fo.apply_12(c, h1, h2, g)) < FP64_TOL
    assert rel(d.apply((h1,)).coeff, fo.apply_1(c, h1, g)) < FP64_TOL


def test_wavefunction_apply_with_e0_multi_sector(fqe):
    norb = 5
    rng = numpy.random.default_rng(11)
    h1, h2 = fo.synthetic_hamiltonian(norb, 3, rng)
    e0 = 0.7 - 0.2j
    wfn = fqe.Wavefunction([[4, 0, norb], [4, 2, norb], [4, -2, norb]])
    wfn.set_wfn(strategy="random")
    ham = fqe.get_restricted_hamiltonian((h1, h2), e_0=e0)
    out = fqe.apply(ham, wfn)
    for key in wfn.sectors():
        sec = wfn.sector(key)
        g = fo.OracleGraph(sec.nalpha(), sec.nbeta(), norb)
        ref = fo.apply_hamiltonian(sec.coeff, (h1, h2), e0, g)
        assert rel(out.sector(key).coeff, ref) < FP64_TOL
    # expectation value and Hermiticity
    ev = wfn.expectationValue(ham)
    other = copy.deepcopy(wfn)
    other.set_wfn(strategy="random")
    lhs = fqe.vdot(other, wfn.apply(ham))
    rhs = numpy.conj(fqe.vdot(wfn, other.apply(ham)))
    assert abs(lhs - rhs - (e0 - numpy.conj(e0)) * fqe.vdot(other, wfn)) < 1e-12
    assert abs(ev.imag - e0.imag) < 1e-12


def test_apply_errors(fqe):
    wfn = fqe.Wavefunction([[2, 0, 3]])
    with pytest.raises(ValueError):
        wfn.apply(fqe.get_restricted_hamiltonian(
            (numpy.zeros((4, 4)), numpy.zeros((4,) * 4))))
    ham = fqe.get_restricted_hamiltonian((numpy.eye(3),
                                          numpy.zeros((3,) * 4)))
    ham._conserve_number = False
    with pytest.raises(TypeError):
        wfn.apply(ham)


# --------------------------------------------------------------------------
# Taylor time evolution
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb,e0", [(2, 2, 4, 0.0), (3, 2, 6, 0.4),
                                           (4, 4, 8, 0.0)])
def test_time_evolve_matches_oracle(fqe, na, nb, norb, e0):
    rng = numpy.random.default_rng(norb + na)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref, nterms = fo.taylor_evolve(c, (h1, h2), e0, 0.1, g)
    wfn = fqe.Wavefunction([[na + nb, na - nb, norb]])
    wfn.set_wfn(strategy="from_data", raw_data={(na + nb, na - nb): c})
    ham = fqe.get_restricted_hamiltonian((h1, h2), e_0=e0)
    out = wfn.time_evolve(0.1, ham)
    assert rel(out.get_coeff((na + nb, na - nb)), ref) < FP64_TOL
    assert out._taylor_terms == nterms


def test_time_evolve_multi_sector_shares_term_count(fqe):
    norb = 4
    rng = numpy.random.default_rng(2)
    h1, h2 = fo.synthetic_hamiltonian(norb, 2, rng)
    wfn = fqe.Wavefunction([[2, 0, norb], [2, 2, norb]])
    wfn.set_wfn(strategy="random")
    out = wfn.time_evolve(0.2, fqe.get_restricted_hamiltonian((h1, h2)))
    # the oracle on the stacked sectors: norms add over sectors
    totals = []
    for key in wfn.sectors():
        sec = wfn.sector(key)
        g = fo.OracleGraph(sec.nalpha(), sec.nbeta(), norb)
        totals.append((key, sec.coeff, g))
    iht = (-0.2j * h1, -0.2j * h2)
    work = {k: c.copy() for k, c, _ in totals}
    res = {k: c.copy() for k, c, _ in totals}
    import math
    for order in range(1, 30):
        nrm = 0.0
        for k, _, g in totals:
            work[k] = fo.apply_12(work[k], iht[0], iht[1], g)
            res[k] = res[k] + work[k] / math.factorial(order)
            nrm += numpy.linalg.norm(work[k]) ** 2
        if math.sqrt(nrm) / math.factorial(order) < 1e-15:
            break
    assert out._taylor_terms == order
    for k, _, _ in totals:
        assert rel(out.get_coeff(k), res[k]) < FP64_TOL


def test_time_evolve_not_converged_raises(fqe):
    wfn = fqe.Wavefunction([[2, 0, 3]])
    wfn.set_wfn(strategy="random")
    rng = numpy.random.default_rng(0)
    h1, h2 = fo.synthetic_hamiltonian(3, 2, rng)
    with pytest.raises(RuntimeError):
        wfn.apply_generated_unitary(1.0, 'taylor',
                                    fqe.get_restricted_hamiltonian((h1, h2)),
                                    expansion=3)


def test_time_evolve_quadratic(fqe):
    norb, na, nb = 5, 2, 2
    rng = numpy.random.default_rng(9)
    h1, _ = fo.synthetic_hamiltonian(norb, 1, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    wfn = fqe.Wavefunction([[4, 0, norb]])
    wfn.set_wfn(strategy="from_data", raw_data={(4, 0): c})
    out = wfn.time_evolve(0.1, fqe.get_restricted_hamiltonian((h1,)))
    dense = numpy.zeros((g.lena * g.lenb,) * 2, dtype=complex)
    for col in range(dense.shape[0]):
        e = numpy.zeros(dense.shape[0], dtype=complex)
        e[col] = 1
        dense[:, col] = fo.apply_1(e.reshape(g.lena, g.lenb), h1, g).ravel()
    import scipy.linalg
    exact = (scipy.linalg.expm(-0.1j * dense) @ c.ravel()).reshape(c.shape)
    assert rel(out.get_coeff((4, 0)), exact) < 1e-12


# --------------------------------------------------------------------------
# the plain-method workloads
# --------------------------------------------------------------------------

def test_small_workload_apply_and_evolve(fqe):
    """norb=8, 4a4b, seed 0, 8 factors: one apply + time_evolve(0.1)."""
    c, h1, h2 = fo.synthetic_workload(8, 4, 4, 0, 8)
    g = fo.OracleGraph(4, 4, 8)
    wfn = fqe.Wavefunction([[8, 0, 8]])
    wfn.set_wfn(strategy="from_data", raw_data={(8, 0): c})
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    out = wfn.apply(ham)
    assert rel(out.get_coeff((8, 0)), fo.apply_12(c, h1, h2, g)) < FP64_TOL
    if ref_lm.available():
        assert rel(out.get_coeff((8, 0)),
                   ref_lm.apply_12_lm(c, h1, h2, g)) < FP64_TOL
    ref, nterms = fo.taylor_evolve(c, (h1, h2), 0.0, 0.1, g)
    ev = wfn.time_evolve(0.1, ham)
    assert rel(ev.get_coeff((8, 0)), ref) < FP64_TOL
    assert ev._taylor_terms == nterms


@pytest.mark.skipif(not ref_lm.available(), reason="oracle/_ref not built")
def test_production_size_apply_vs_reference_c(fqe):
    """norb=12, 6a6b (853,776 determinants), seed 1, 12 factors: the
    engine's apply against the reference's own compiled lm kernels."""
    print("norb=12 apply vs reference C", flush=True)
    c, h1, h2 = fo.synthetic_workload(12, 6, 6, 1, 12)
    d = fqe.FqeData(6, 6, 12)
    d.coeff = c
    out = d.apply((h1, h2)).coeff
    rg = ref_lm.RefGraph(6, 6, 12)
    ref = ref_lm.apply_12_lm(c, h1, h2, rg)
    assert rel(out, ref) < FP64_TOL


# --------------------------------------------------------------------------
# mixed precision
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb", [(3, 3, 6), (4, 4, 8)])
def test_mixed_precision_apply_and_evolve(fqe, na, nb, norb):
    c, h1, h2 = fo.synthetic_workload(norb, na, nb, 5, norb)
    g = fo.OracleGraph(na, nb, norb)
    wfn = fqe.Wavefunction([[na + nb, na - nb, norb]], dtype=numpy.complex64)
    key = (na + nb, na - nb)
    wfn.set_wfn(strategy="from_data", raw_data={key: c})
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    out = wfn.apply(ham).get_coeff(key)
    assert out.dtype == numpy.complex64
    err = rel(out, fo.apply_12(c, h1, h2, g))
    print(f"mixed apply rel err {err:.3e}", flush=True)
    assert err < MIXED_TOL
    ref, _ = fo.taylor_evolve(c, (h1, h2), 0.0, 0.1, g)
    err = rel(wfn.time_evolve(0.1, ham).get_coeff(key), ref)
    print(f"mixed evolve rel err {err:.3e}", flush=True)
    assert err < MIXED_TOL


# --------------------------------------------------------------------------
# fused (gather + GEMM + slot scatter) and unfused kernels
# --------------------------------------------------------------------------

@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("na,nb,norb,kind", [(3, 0, 7, "symmetric"),
                                             (5, 5, 10, "symmetric"),
                                             (4, 4, 9, "real"),
                                             (7, 3, 10, "symmetric"),
                                             (2, 6, 8, "real"),
                                             (1, 1, 3, "real")])
def test_fused_and_unfused_paths(fqe, fused, dtype, na, nb, norb, kind):
    import ctypes
    from fqe_amd import settings, engine, _lib
    print(f"paths fused={fused} {dtype} {na} {nb} {norb} {kind}", flush=True)
    rng = numpy.random.default_rng(na * 100 + nb * 10 + norb)
    h1, h2 = _h(norb, rng, kind)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = fo.apply_12(c, h1, h2, g)
    npdt = numpy.complex128 if dtype == "c128" else numpy.complex64
    saved = settings.fused, settings.split
    settings.fused = fused
    settings.split = False   # the split apply has its own tests below
    try:
        d = fqe.FqeData(na, nb, norb, dtype=npdt)
        d.coeff = c
        op = engine.prepare_operator((h1, h2), norb, _lib.device())
        s, keep = engine.sector_struct(d.get_fcigraph(), op)
        prec = engine.precision_of(d.device_coeff().dtype)
        took = _lib.lib().fqehip_fused_supported(prec, ctypes.byref(s),
                                                 ctypes.byref(op.c_struct()))
        assert bool(took) == fused
        out = d.apply((h1, h2)).coeff
    finally:
        settings.fused, settings.split = saved
    tol = FP64_TOL if dtype == "c128" else MIXED_TOL
    assert rel(out, ref) < tol


def test_fused_taylor_with_e0(fqe):
    """Time evolution through the fused kernels (real symmetric h)."""
    norb, na, nb = 9, 4, 3
    c, h1, h2 = fo.synthetic_workload(norb, na, nb, 7, norb)
    g = fo.OracleGraph(na, nb, norb)
    ref, nterms = fo.taylor_evolve(c, (h1, h2), 0.3 + 0.1j, 0.05, g)
    key = (na + nb, na - nb)
    wfn = fqe.Wavefunction([[na + nb, na - nb, norb]])
    wfn.set_wfn(strategy="from_data", raw_data={key: c})
    out = wfn.time_evolve(0.05, fqe.get_restricted_hamiltonian((h1, h2),
                                                               e_0=0.3 + 0.1j))
    assert out._taylor_terms == nterms
    assert rel(out.get_coeff(key), ref) < FP64_TOL


def test_fused_taylor_mixed_split(fqe):
    """Mixed-precision time evolution through the split apply."""
    norb, na, nb = 9, 4, 3
    c, h1, h2 = fo.synthetic_workload(norb, na, nb, 8, norb)
    g = fo.OracleGraph(na, nb, norb)
    ref, nterms = fo.taylor_evolve(c, (h1, h2), 0.3 + 0.1j, 0.05, g)
    key = (na + nb, na - nb)
    wfn = fqe.Wavefunction([[na + nb, na - nb, norb]], dtype=numpy.complex64)
    wfn.set_wfn(strategy="from_data", raw_data={key: c})
    out = wfn.time_evolve(0.05, fqe.get_restricted_hamiltonian((h1, h2),
                                                               e_0=0.3 + 0.1j))
    err = rel(out.get_coeff(key), ref)
    print(f"split mixed evolve rel err {err:.3e}", flush=True)
    assert err < MIXED_TOL


# --------------------------------------------------------------------------
# split apply (complex64, pair-symmetric operator): beta-beta block as a
# dense GEMM, alpha-only row kernel
# --------------------------------------------------------------------------

def _bf16_to_f64(t):
    """int16 tensor of bfloat16 bit patterns -> float64 numpy."""
    u = t.cpu().numpy().view(numpy.uint16).astype(numpy.uint32) << 16
    return u.view(numpy.float32).astype(numpy.float64)


@pytest.mark.parametrize("nb,norb,kind", [(3, 7, "symmetric"),
                                          (1, 3, "symmetric"),
                                          (4, 9, "symmetric"),
                                          (2, 5, "real"),
                                          (0, 4, "symmetric"),
                                          (5, 5, "symmetric")])
def test_beta_block_matches_dense_sector(fqe, nb, norb, kind):
    """Hbb equals the dense CI matrix of the (0, nb) sector, in which no
    alpha string exists to excite (lena = 1), with the same bits on a
    second build; its bf16 hi + lo split represents it to 2^-16 relative
    and is zero-padded."""
    import torch
    from fqe_amd import _lib, engine
    print(f"beta block nb={nb} norb={norb} {kind}", flush=True)
    rng = numpy.random.default_rng(100 + 7 * nb + norb)
    h1, h2 = _h(norb, rng, kind)
    dense = fo.dense_matrix(h1, h2, fo.OracleGraph(0, nb, norb)).real
    graph = fqe.FciGraph(0, nb, norb)
    op = engine.prepare_operator((h1, h2), norb, _lib.device())
    assert op.pair_symmetric == (kind == "symmetric")
    lenb = graph.lenb()
    fwb, _ = graph.tables(False).fused_tables(op.compressed)
    ldh = int(_lib.call("fqehip_beta_block_ld", lenb))
    assert ldh % 128 == 0 and ldh >= lenb
    dev = _lib.device()
    hcol = torch.full((lenb, lenb), 7.0, dtype=torch.float64, device=dev)
    hhi = torch.full((ldh, ldh), 77, dtype=torch.int16, device=dev)
    hlo = torch.full((ldh, ldh), 77, dtype=torch.int16, device=dev)
    _lib.call("fqehip_beta_block", op.npair, lenb, _lib.ptr(fwb),
              _lib.ptr(op.a), _lib.ptr(hcol), _lib.ptr(hhi), _lib.ptr(hlo),
              ldh, _lib.stream_ptr())
    hbb = hcol.cpu().numpy().T
    assert rel(hbb, dense) < FP64_TOL
    # the build runs in a fixed order: a second build gives the same bits
    hcol2 = torch.full_like(hcol, 3.0)
    _lib.call("fqehip_beta_block", op.npair, lenb, _lib.ptr(fwb),
              _lib.ptr(op.a), _lib.ptr(hcol2), _lib.ptr(hhi), _lib.ptr(hlo),
              ldh, _lib.stream_ptr())
    assert torch.equal(hcol, hcol2)
    split = _bf16_to_f64(hhi) + _bf16_to_f64(hlo)
    assert numpy.all(split[lenb:, :] == 0) and numpy.all(split[:, lenb:] == 0)
    err = numpy.abs(split[:lenb, :lenb] - dense)
    assert numpy.all(err <= 2.0 ** -16 * numpy.abs(dense) + 1e-30)


@pytest.mark.parametrize("na,nb,norb", [(3, 0, 7), (0, 3, 7), (5, 5, 10),
                                        (7, 3, 10), (1, 1, 3), (4, 4, 9),
                                        (2, 6, 8), (6, 5, 12)])
def test_split_apply(fqe, na, nb, norb):
    """The split apply is taken for complex64 and a pair-symmetric operator
    and agrees with the oracle and with the row kernel's own result."""
    import ctypes
    from fqe_amd import _lib, engine, settings
    print(f"split apply {na} {nb} {norb}", flush=True)
    rng = numpy.random.default_rng(norb * 13 + na * 5 + nb)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = fo.apply_12(c, h1, h2, g)
    d = fqe.FqeData(na, nb, norb, dtype=numpy.complex64)
    d.coeff = c
    op = engine.prepare_operator((h1, h2), norb, _lib.device())
    assert op.pair_symmetric
    s, keep = engine.sector_struct(d.get_fcigraph(), op, _lib.C64)
    assert s.hbb_hi and s.hbb_lo and s.ldh >= g.lenb
    assert _lib.lib().fqehip_fused_supported(_lib.C64, ctypes.byref(s),
                                             ctypes.byref(op.c_struct()))
    split = d.apply((h1, h2)).coeff
    saved = settings.split
    settings.split = False
    try:
        rows = d.apply((h1, h2)).coeff
    finally:
        settings.split = saved
    err = rel(split, ref)
    print(f"split rel err {err:.3e} vs rows {rel(rows, ref):.3e}", flush=True)
    assert err < MIXED_TOL
    assert rel(split, rows) < MIXED_TOL


@pytest.mark.parametrize("lena,lenb", [(1, 1), (5, 130), (129, 127),
                                       (300, 257), (1000, 1000)])
def test_beta_gemm_f64_matches_numpy(fqe, lena, lenb):
    """fqehip_beta_gemm_f64: out = C hcol (hcol[Ib][Ib'] = Hbb[Ib'][Ib]) on
    the FP64 matrix cores, for sizes off the 128-tile and 16-step grid."""
    import torch
    from fqe_amd import _lib
    print(f"beta gemm f64 {lena}x{lenb}", flush=True)
    rng = numpy.random.default_rng(lena * 3 + lenb)
    c = rng.normal(size=(lena, lenb)) + 1j * rng.normal(size=(lena, lenb))
    h = rng.normal(size=(lenb, lenb))
    dev = _lib.device()
    cd = torch.from_numpy(c).to(dev)
    hd = torch.from_numpy(h).to(dev)
    out = torch.full((lena, lenb), 5.0 + 5.0j, dtype=torch.complex128,
                     device=dev)
    _lib.call("fqehip_beta_gemm_f64", lena, lenb, _lib.ptr(hd), _lib.ptr(cd),
              _lib.ptr(out), None, _lib.stream_ptr())
    ref = c @ h
    assert numpy.max(numpy.abs(out.cpu().numpy() - ref)) < 1e-12 * numpy.abs(ref).max()


@pytest.mark.parametrize("na,nb,norb", [(3, 0, 7), (0, 3, 7), (5, 5, 10),
                                        (7, 3, 10), (1, 1, 3), (4, 4, 9),
                                        (2, 6, 8), (6, 5, 12), (7, 7, 14)])
def test_split_apply_fp64(fqe, na, nb, norb):
    """The complex128 split apply (FP64 beta-beta GEMM + FP64 row kernel
    over the live alpha pair rows) is taken for a pair-symmetric operator
    and agrees with the oracle (sectors the oracle finishes quickly) and
    with the FP64 row kernel that scatters beta excitations itself."""
    import ctypes
    from fqe_amd import _lib, engine, settings
    print(f"split apply fp64 {na} {nb} {norb}", flush=True)
    rng = numpy.random.default_rng(norb * 11 + na * 7 + nb)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = fo.apply_12(c, h1, h2, g) if g.lena * g.lenb <= 200000 else None
    d = fqe.FqeData(na, nb, norb, dtype=numpy.complex128)
    d.coeff = c
    op = engine.prepare_operator((h1, h2), norb, _lib.device())
    s, keep = engine.sector_struct(d.get_fcigraph(), op, _lib.C128)
    assert s.hbb_f64 and not s.hbb_hi
    assert _lib.lib().fqehip_fused_supported(_lib.C128, ctypes.byref(s),
                                             ctypes.byref(op.c_struct()))
    split = d.apply((h1, h2)).coeff
    saved = settings.split_fp64
    settings.split_fp64 = False
    try:
        rows = d.apply((h1, h2)).coeff
    finally:
        settings.split_fp64 = saved
    assert rel(split, rows) < FP64_TOL
    if ref is not None:
        assert rel(split, ref) < FP64_TOL
    # the column-owned kernel (default) and the shared-D-tile kernel
    # accumulate the same products in the same order: identical bits for
    # one prepared operator (each preparation rebuilds H_bb with LDS
    # atomics, which may differ in the last bit)
    outs = []
    saved = settings.split_tile_fp64
    try:
        for tile in (0, 32):
            settings.split_tile_fp64 = tile
            e = copy.deepcopy(d)
            e.apply_inplace_scaled((h1, h2), op=op)
            outs.append(e.coeff)
    finally:
        settings.split_tile_fp64 = saved
    assert rel(outs[0], split) < FP64_TOL
    assert numpy.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("tile", [16])
@pytest.mark.parametrize("na,nb,norb", [(3, 0, 7), (5, 5, 10), (1, 1, 3),
                                        (2, 6, 8), (6, 5, 12), (7, 7, 14),
                                        (0, 4, 9), (8, 7, 15)])
def test_split_apply_live_rows(fqe, na, nb, norb, tile):
    """The split row kernel over the live pair rows (the default) computes
    the same products as the all-rows kernel (tile 16; same operands per
    output element), so their applies are identical, bit for bit, and
    agree with the oracle (sectors small enough for the oracle; the full-size sectors,
    norb 14 and 15 at half filling, are checked against the all-rows
    kernel only)."""
    from fqe_amd import _lib, engine, settings
    print(f"split live rows {tile}: {na} {nb} {norb}", flush=True)
    rng = numpy.random.default_rng(norb * 17 + na * 3 + nb)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = fo.apply_12(c, h1, h2, g) if g.lena * g.lenb <= 200000 else None
    d = fqe.FqeData(na, nb, norb, dtype=numpy.complex64)
    d.coeff = c
    # one prepared operator for both (each preparation rebuilds H_bb with
    # LDS atomics, which may differ in the last bit)
    op = engine.prepare_operator((h1, h2), norb, _lib.device())
    outs = []
    saved = settings.split_tile
    try:
        for t in (0, tile):
            settings.split_tile = t
            e = copy.deepcopy(d)
            e.apply_inplace_scaled((h1, h2), op=op)
            outs.append(e.coeff)
    finally:
        settings.split_tile = saved
    out, base = outs
    if ref is not None:
        assert rel(out, ref) < MIXED_TOL
    assert numpy.array_equal(out, base)


# --------------------------------------------------------------------------
# Chebyshev propagator (reference src/fqe/wavefunction.py:570-602; the
# reference's tests tests/evolution_test.py:777-808 and
# tests/wavefunction_test.py:649-676 compare it with the Taylor series to
# 1e-8; here also with the exact exponential of the dense oracle matrix)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("na,nb,norb", [(2, 1, 4), (2, 2, 5)])
def test_chebyshev_matches_taylor_and_expm(fqe, na, nb, norb):
    from scipy.linalg import expm
    rng = numpy.random.default_rng(40 + norb)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    dense = fo.dense_matrix(h1, h2, g)
    w = numpy.linalg.eigvalsh(dense)
    exact = (expm(-1j * 0.1 * dense) @ c.reshape(-1)).reshape(c.shape)
    key = (na + nb, na - nb)
    wfn = fqe.Wavefunction([[na + nb, na - nb, norb]])
    wfn.set_wfn(strategy="from_data", raw_data={key: c})
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    cheb = wfn.apply_generated_unitary(0.1, "chebyshev", ham, accuracy=1e-14,
                                       spec_lim=(w[0], w[-1]))
    tay = wfn.apply_generated_unitary(0.1, "taylor", ham)
    assert rel(cheb.get_coeff(key), exact) < 1e-11
    assert numpy.linalg.norm(cheb.get_coeff(key) - tay.get_coeff(key)) < 1e-8
    with pytest.raises(RuntimeError):
        wfn.apply_generated_unitary(0.1, "chebyshev", ham, expansion=3,
                                    spec_lim=(w[0], w[-1]))


def test_norm_and_vdot_fp64_kernels(fqe):
    rng = numpy.random.default_rng(77)
    norb, na, nb = 6, 3, 2
    g = fo.OracleGraph(na, nb, norb)
    a = fo.synthetic_coeff(g.lena, g.lenb, rng) * 3.0
    b = fo.synthetic_coeff(g.lena, g.lenb, rng)
    key = (na + nb, na - nb)
    wa = fqe.Wavefunction([[na + nb, na - nb, norb]])
    wb = fqe.Wavefunction([[na + nb, na - nb, norb]])
    wa.set_wfn(strategy="from_data", raw_data={key: a})
    wb.set_wfn(strategy="from_data", raw_data={key: b})
    assert abs(wa.norm() - numpy.linalg.norm(a)) < 1e-13
    assert abs(fqe.vdot(wa, wb) - numpy.vdot(a, b)) < 1e-13


# --------------------------------------------------------------------------
# reduced density matrices (reference src/fqe/fqe_data.py:1668-1782;
# FQE's own outputs tests/unittest_data/fqe_data/{dr,di}{1,2}020104.npy,
# tests/fqe_data_test.py:669-700)
# --------------------------------------------------------------------------

def test_rdm12_match_fqe_golden(fqe):
    z = numpy.load(os.path.join(GOLDEN, "fqe_data_golden.npz"))
    d = fqe.FqeData(2, 1, 4)
    d.coeff = z["c_020104"]
    (r1,) = d.rdm1()
    assert numpy.max(numpy.abs(r1 - z["rdm1_020104"])) < 1e-13
    o1, o2 = d.rdm12()
    assert numpy.max(numpy.abs(o1 - z["rdm1_020104"])) < 1e-13
    assert numpy.max(numpy.abs(o2 - z["rdm2_020104"])) < 1e-13


def test_rdm1_contracts_to_the_one_body_energy(fqe):
    """sum_pq h1[p,q] rdm1[p,q] = <C|H1|C> for a random one-body h1, and
    the transition RDM with a second state matches <bra|H1|ket>."""
    rng = numpy.random.default_rng(91)
    norb, na, nb = 6, 3, 2
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    b = fo.synthetic_coeff(g.lena, g.lenb, rng)
    h1 = rng.normal(size=(norb, norb)) + 1j * rng.normal(size=(norb, norb))
    ket = fqe.FqeData(na, nb, norb)
    ket.coeff = c
    bra = fqe.FqeData(na, nb, norb)
    bra.coeff = b
    (r1,) = ket.rdm1()
    assert abs(numpy.sum(h1 * r1) - numpy.vdot(c, fo.apply_1(c, h1, g))) < 1e-12
    (t1,) = ket.rdm1(bra)
    assert abs(numpy.sum(h1 * t1) - numpy.vdot(b, fo.apply_1(c, h1, g))) < 1e-12
    # two-body: <C|sum h2[i,j,k,l] a+_i a+_j a_k a_l|C> from rdm2
    h2 = rng.normal(size=(norb,) * 4)
    _, r2 = ket.rdm12()
    e2 = numpy.vdot(c, fo.apply_12(c, numpy.zeros((norb, norb)), h2, g))
    assert abs(numpy.einsum("ijkl,ijkl->", h2, r2) - e2) < 1e-11


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_wavefunction_save_read_round_trip(fqe, tmp_path, dtype):
    """Wavefunction.save / read (reference src/fqe/wavefunction.py:726-765)
    in the pickle-free format: same sectors, norb, conserved quantities and
    coefficients; a foreign file is refused."""
    wfn = fqe.Wavefunction([[4, 0, 5], [4, 2, 5], [4, -2, 5]], dtype=dtype)
    wfn.set_wfn(strategy="random")
    wfn.save("w.npz", path=str(tmp_path))
    other = fqe.Wavefunction([[2, 0, 3]])
    other.read("w.npz", path=str(tmp_path))
    assert other.norb() == 5 and set(other.sectors()) == set(wfn.sectors())
    assert other._conserved == wfn._conserved
    for key in wfn.sectors():
        got = other.get_coeff(key)
        assert got.dtype == numpy.dtype(dtype)
        assert numpy.array_equal(got, wfn.get_coeff(key))
    numpy.savez(str(tmp_path / "bad.npz"), x=numpy.zeros(3))
    with pytest.raises((ValueError, KeyError)):
        other.read("bad.npz", path=str(tmp_path))
