"""Edge cases and full-size properties of the apply, on every kernel path
(FP64 split / row / unfused, mixed split / row / unfused).

* edge sectors: no electrons of one spin, full shells, one orbital, one
  determinant, compared with the oracle;
* norb=14 7a7b (the full benchmark sector, 11,778,624 determinants):
  the FP64 engine against the reference's own compiled C (≤1e-12), the
  mixed engine against the FP64 engine (≤1e-5), and the size-independent
  properties of a Hermitian operator: <a|H b> = conj(<b|H a>) and
  linearity H(x a + y b) = x H a + y H b.
"""

import numpy
import pytest

from oracle import fqe_oracle as fo
from oracle import ref_lm

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-4


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300)


PATHS = [("c128", True, True), ("c128", True, False), ("c128", False, False),
         ("c64", True, True), ("c64", True, False), ("c64", False, False)]


# norb 15: 120 compressed pairs, past the FP64 row kernels' gather depth
# (112), so complex128 takes the unfused kernels; complex64 keeps its split
@pytest.mark.parametrize("dtype,fused,split", PATHS)
@pytest.mark.parametrize("na,nb,norb", [(1, 0, 1), (1, 1, 1), (0, 0, 3),
                                        (2, 2, 2), (3, 1, 3), (0, 2, 4),
                                        (4, 0, 4), (1, 3, 4), (1, 1, 15),
                                        (2, 1, 15)])
def test_edge_sectors(fqe, dtype, fused, split, na, nb, norb):
    from fqe_amd import settings
    print(f"edge {dtype} fused={fused} split={split} {na} {nb} {norb}",
          flush=True)
    rng = numpy.random.default_rng(7 * norb + 3 * na + nb)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    ref = fo.apply_12(c, h1, h2, g)
    saved = settings.fused, settings.split, settings.split_fp64
    settings.fused, settings.split, settings.split_fp64 = fused, split, split
    try:
        npdt = numpy.complex128 if dtype == "c128" else numpy.complex64
        d = fqe.FqeData(na, nb, norb, dtype=npdt)
        d.coeff = c
        out = d.apply((h1, h2)).coeff
    finally:
        settings.fused, settings.split, settings.split_fp64 = saved
    tol = FP64_TOL if dtype == "c128" else MIXED_TOL
    if numpy.linalg.norm(ref) < 1e-12:
        assert numpy.linalg.norm(out) < 1e-6
    else:
        assert rel(out, ref) < tol


@pytest.fixture(scope="module")
def big(fqe):
    c, h1, h2 = fo.synthetic_workload(14, 7, 7, 2, 14)
    rng = numpy.random.default_rng(12)
    b = fo.synthetic_coeff(c.shape[0], c.shape[1], rng)
    return c, b, h1, h2


def test_norb14_fp64_vs_reference_c_and_mixed(fqe, big):
    print("norb=14 FP64 and mixed vs reference C", flush=True)
    c, _, h1, h2 = big
    d = fqe.FqeData(7, 7, 14)
    d.coeff = c
    out = d.apply((h1, h2)).coeff
    ref = ref_lm.apply_12_lm(c, h1, h2, ref_lm.RefGraph(7, 7, 14))
    assert rel(out, ref) < FP64_TOL
    m = fqe.FqeData(7, 7, 14, dtype=numpy.complex64)
    m.coeff = c
    outm = m.apply((h1, h2)).coeff
    err = rel(outm.astype(numpy.complex128), ref)
    print(f"norb=14 mixed rel err {err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_norb14_hermiticity_and_linearity(fqe, big, dtype):
    print(f"norb=14 properties {dtype}", flush=True)
    c, b, h1, h2 = big
    npdt = numpy.complex128 if dtype == "c128" else numpy.complex64
    tol = 1e-12 if dtype == "c128" else 1e-5

    def H(x):
        d = fqe.FqeData(7, 7, 14, dtype=npdt)
        d.coeff = x
        return d.apply((h1, h2)).coeff.astype(numpy.complex128)

    hc, hb = H(c), H(b)
    lhs = numpy.vdot(b, hc)
    rhs = numpy.conj(numpy.vdot(c, hb))
    scale = numpy.linalg.norm(hb) * numpy.linalg.norm(c)
    assert abs(lhs - rhs) < tol * scale
    x, y = 0.3 - 0.7j, -1.1 + 0.2j
    hl = H(x * c + y * b)
    assert rel(hl, x * hc + y * hb) < tol
