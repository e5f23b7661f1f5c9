"""GPU parity of the Ozaki-digit engine: the complex128 split apply whose
alpha-row contraction is formed exactly from 8-bit digits on the int8
matrix cores (``settings.contraction = "ozaki"``).

Tolerance: the metric's gate, a normwise relative error of at most 1e-9
against the complex128 references (the FP64 engine, the oracle, the
reference's compiled C lm kernels), for applies and for time evolution.
The digit scheme's own error is about 1e-11 (dropped low-order digit pairs
and the 2^-38 truncation of D'), so the gate holds with margin.
"""

import copy

import numpy
import pytest

from oracle import fqe_oracle as fo
from oracle import ref_lm

pytestmark = pytest.mark.gpu

GATE = 1e-9


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return numpy.linalg.norm(numpy.asarray(a) - numpy.asarray(b)) / max(
        numpy.linalg.norm(numpy.asarray(b)), 1e-300)


class _contraction:
    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        from fqe_amd import settings
        self.saved = settings.contraction
        settings.contraction = self.mode

    def __exit__(self, *exc):
        from fqe_amd import settings
        settings.contraction = self.saved


def _apply(fqe, c, h1, h2, na, nb, norb, mode, op=None):
    with _contraction(mode):
        d = fqe.FqeData(na, nb, norb, dtype=numpy.complex128)
        d.coeff = c
        if op is None:
            return d.apply((h1, h2)).coeff
        d.apply_inplace_scaled((h1, h2), op=op)
        return d.coeff


@pytest.mark.parametrize("na,nb,norb", [(1, 1, 3), (2, 2, 4), (4, 4, 8),
                                        (3, 5, 9), (5, 5, 10), (7, 3, 10),
                                        (2, 6, 8), (6, 6, 12), (5, 6, 11),
                                        (6, 7, 13)])
def test_ozaki_apply_matches_fp64_and_oracle(fqe, na, nb, norb):
    import ctypes
    from fqe_amd import _lib, engine
    print(f"ozaki apply {na} {nb} {norb}", flush=True)
    rng = numpy.random.default_rng(norb * 13 + na * 5 + nb)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    with _contraction("ozaki"):
        d = fqe.FqeData(na, nb, norb, dtype=numpy.complex128)
        op = engine.prepare_operator((h1, h2), norb, _lib.device())
        s, keep = engine.sector_struct(d.get_fcigraph(), op, _lib.C128)
        assert s.tile == _lib.TILE_OZAKI and s.hbb_f64
        assert _lib.lib().fqehip_fused_supported(
            _lib.C128, ctypes.byref(s), ctypes.byref(op.c_struct()))
        del keep
    oz = _apply(fqe, c, h1, h2, na, nb, norb, "ozaki")
    fp = _apply(fqe, c, h1, h2, na, nb, norb, "fp64")
    err = rel(oz, fp)
    print(f"  ozaki vs fp64 {err:.3e}", flush=True)
    assert err < GATE
    if g.lena * g.lenb <= 200000:
        assert rel(oz, fo.apply_12(c, h1, h2, g)) < GATE


def test_ozaki_peaked_vector(fqe):
    """A vector dominated by one determinant (a ground-state-like shape):
    the per-string D' scale must keep the small components accurate."""
    norb, na, nb = 10, 5, 5
    rng = numpy.random.default_rng(77)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = 1e-4 * fo.synthetic_coeff(g.lena, g.lenb, rng)
    c[0, 0] = 1.0
    c[3, 7] = 0.05
    oz = _apply(fqe, c, h1, h2, na, nb, norb, "ozaki")
    fp = _apply(fqe, c, h1, h2, na, nb, norb, "fp64")
    err = rel(oz, fp)
    print(f"peaked ozaki vs fp64 {err:.3e}", flush=True)
    assert err < GATE


def test_ozaki_small_workload_apply_and_evolve(fqe):
    """norb=8 4a4b seed 0: one apply and time_evolve(0.1) against the
    oracle's complex128 Taylor series (same term count)."""
    c, h1, h2 = fo.synthetic_workload(8, 4, 4, 0, 8)
    g = fo.OracleGraph(4, 4, 8)
    with _contraction("ozaki"):
        wfn = fqe.Wavefunction([[8, 0, 8]])
        wfn.set_wfn(strategy="from_data", raw_data={(8, 0): c})
        ham = fqe.get_restricted_hamiltonian((h1, h2))
        out = wfn.apply(ham)
        ev = wfn.time_evolve(0.1, ham)
    assert rel(out.get_coeff((8, 0)), fo.apply_12(c, h1, h2, g)) < GATE
    ref, nterms = fo.taylor_evolve(c, (h1, h2), 0.0, 0.1, g)
    err = rel(ev.get_coeff((8, 0)), ref)
    print(f"evolve ozaki vs oracle {err:.3e}", flush=True)
    assert err < GATE
    assert ev._taylor_terms == nterms


@pytest.mark.skipif(not ref_lm.available(), reason="oracle/_ref not built")
def test_ozaki_norb14_vs_fp64_and_reference_c(fqe):
    """The metric's workload (norb=14 7a7b, seed 2, 14 factors): the digit
    engine against the FP64 engine and the reference's compiled C lm
    kernels (complex128), at the 1e-9 gate."""
    print("norb=14 ozaki vs fp64 and reference C", flush=True)
    c, h1, h2 = fo.synthetic_workload(14, 7, 7, 2, 14)
    oz = _apply(fqe, c, h1, h2, 7, 7, 14, "ozaki")
    fp = _apply(fqe, c, h1, h2, 7, 7, 14, "fp64")
    ref = ref_lm.apply_12_lm(c, h1, h2, ref_lm.RefGraph(7, 7, 14))
    e1, e2 = rel(oz, fp), rel(oz, ref)
    print(f"  ozaki vs fp64 {e1:.3e}, vs reference C {e2:.3e}", flush=True)
    assert e1 < GATE and e2 < GATE


def test_ozaki_norb15_global_beta_rows(fqe):
    """norb=15 7a8b (41.4 M determinants), the largest sector the digit
    split kernel takes (121 of its 128 pair rows): row ia (103 KB) no longer
    fits in LDS beside the operator digits, so the beta gather reads it from
    global memory.  Against the FP64 engine (the column-owned FP64 split
    kernel at its widest gather, 121 of 128 rows) at the 1e-9 gate, and the
    FP64 split result against the FP64 unfused path (its LDS-resident pair
    GEMM, 8 row tiles) to 1e-12."""
    import ctypes
    import torch
    from fqe_amd import _lib, engine
    norb, na, nb = 15, 7, 8
    print("norb=15 ozaki vs fp64", flush=True)
    rng = numpy.random.default_rng(1507)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    with _contraction("ozaki"):
        d = fqe.FqeData(na, nb, norb, dtype=numpy.complex128)
        op = engine.prepare_operator((h1, h2), norb, _lib.device())
        s, keep = engine.sector_struct(d.get_fcigraph(), op, _lib.C128)
        assert s.tile == _lib.TILE_OZAKI and s.hbb_f64
        assert _lib.lib().fqehip_fused_supported(
            _lib.C128, ctypes.byref(s), ctypes.byref(op.c_struct()))
        del keep, d
    oz = _apply(fqe, c, h1, h2, na, nb, norb, "ozaki")
    torch.cuda.empty_cache()
    fp = _apply(fqe, c, h1, h2, na, nb, norb, "fp64")
    torch.cuda.empty_cache()
    err = rel(oz, fp)
    print(f"  ozaki vs fp64 {err:.3e}", flush=True)
    assert err < GATE
    from fqe_amd import settings
    saved = settings.split_fp64
    settings.split_fp64 = False
    try:
        unf = _apply(fqe, c, h1, h2, na, nb, norb, "fp64")
    finally:
        settings.split_fp64 = saved
    torch.cuda.empty_cache()
    err2 = rel(fp, unf)
    print(f"  fp64 split vs fp64 unfused {err2:.3e}", flush=True)
    assert err2 < 1e-12

