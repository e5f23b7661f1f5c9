"""The apply in panels of orbital pairs (fqehip_apply_panels): the unfused
gather / GEMM / scatter kernels with the dvec and Y held for a few pairs at
a time, for sectors whose full dvec outgrows device memory.

Tolerance: FP64 throughout, 1e-12 normwise against the unchunked apply
(the same products summed in another grouping); at norb=16 8a8b (165.6M
determinants, where the unchunked dvec + Y would need 723 GB) the checks
are Hermiticity <y|H x> = conj(<x|H y>) and linearity H(a x + b y) =
a H x + b H y, both to 1e-12 of the scale.
"""

import ctypes

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


def _panels(graph, op, c, nbytes, alpha=1.0, beta=0.0, cin=None):
    import torch
    from fqe_amd import _lib, engine
    s, keep = engine.sector_struct(graph, op, _lib.C128)
    opc = op.c_struct()
    out = torch.empty_like(c)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=c.device)
    _lib.call("fqehip_apply_panels", _lib.C128, ctypes.byref(s),
              ctypes.byref(opc), _lib.ptr(c), alpha.real, alpha.imag,
              beta.real, beta.imag, _lib.ptr(cin), _lib.ptr(out), None, None,
              None, _lib.ptr(ws), nbytes, _lib.stream_ptr())
    torch.cuda.synchronize()
    del keep
    return out


@pytest.mark.parametrize("kind", ["real", "complex"])
@pytest.mark.parametrize("na,nb,norb", [(3, 2, 6), (6, 6, 12)])
def test_panels_match_unchunked(fqe, kind, na, nb, norb):
    import torch
    from fqe_amd import _lib, engine, settings
    print(f"panels {kind} {na} {nb} {norb}", flush=True)
    rng = numpy.random.default_rng(norb * 3 + na)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    if kind == "complex":
        h1 = h1 + 0.05j * rng.normal(size=h1.shape)
        h1 = 0.5 * (h1 + h1.conj().T)
        h2 = h2 + 0.01j * rng.normal(size=h2.shape)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    cin = fo.synthetic_coeff(g.lena, g.lenb, rng)
    dev = _lib.device()
    d = fqe.FqeData(na, nb, norb)
    graph = d.get_fcigraph()
    op = engine.prepare_operator((h1, h2), norb, dev)
    tc = torch.from_numpy(c).to(dev)
    tcin = torch.from_numpy(cin).to(dev)
    ref = torch.empty_like(tc)
    saved = settings.fused
    settings.fused = False
    try:
        engine.apply_sector(graph, tc, op, ref, 0.3 - 0.4j, 0.5, tcin)
    finally:
        settings.fused = saved
    ref = ref.cpu().numpy()
    plane = 16 * g.lena * g.lenb
    for planes in (3, 6, 41, 2 * op.npair + 2):
        out = _panels(graph, op, tc, planes * plane, 0.3 - 0.4j, 0.5, tcin)
        err = rel(out.cpu().numpy(), ref)
        print(f"  {planes} planes: {err:.2e}", flush=True)
        assert err < 1e-12
    if g.lena * g.lenb <= 5000:
        want = (0.3 - 0.4j) * fo.apply_12(c, h1, h2, g) + 0.5 * cin
        assert rel(ref, want) < 1e-12


def test_panels_norb14_match_split(fqe):
    """A budget of 30 planes (5.6 GB) at norb=14 7a7b against the default
    FP64 split apply."""
    import torch
    from fqe_amd import _lib, engine
    c, h1, h2 = fo.synthetic_workload(14, 7, 7, 2, 14)
    dev = _lib.device()
    d = fqe.FqeData(7, 7, 14)
    graph = d.get_fcigraph()
    op = engine.prepare_operator((h1, h2), 14, dev)
    tc = torch.from_numpy(c).to(dev)
    ref = torch.empty_like(tc)
    engine.apply_sector(graph, tc, op, ref)
    out = _panels(graph, op, tc, 30 * 16 * tc.numel())
    err = rel(out.cpu().numpy(), ref.cpu().numpy())
    print(f"norb=14 panels vs split {err:.2e}", flush=True)
    assert err < 1e-12


def test_norb16_hermiticity_and_linearity(fqe):
    """norb=16 8a8b: the engine routes the sector to the panel apply (its
    unfused workspace is 723 GB); Hermiticity and linearity to 1e-12."""
    import torch
    from fqe_amd import _lib, engine, settings
    norb, na, nb = 16, 8, 8
    rng = numpy.random.default_rng(16)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    dev = _lib.device()
    d = fqe.FqeData(na, nb, norb)
    graph = d.get_fcigraph()
    op = engine.prepare_operator((h1, h2), norb, dev)
    s, keep = engine.sector_struct(graph, op, _lib.C128)
    assert engine.needs_panels(s, op, _lib.C128)
    del keep
    gen = torch.Generator(device=dev)
    gen.manual_seed(16)
    shape = (graph.lena(), graph.lenb())

    def vec():
        v = torch.complex(torch.randn(shape, generator=gen, device=dev,
                                      dtype=torch.float64),
                          torch.randn(shape, generator=gen, device=dev,
                                      dtype=torch.float64))
        return v / torch.linalg.vector_norm(v)

    def H(x):
        out = torch.empty_like(x)
        engine.apply_sector(graph, x, op, out)
        return out

    print("norb=16 8a8b panels: H x", flush=True)
    x, y = vec(), vec()
    hx = H(x)
    print("norb=16 8a8b panels: H y", flush=True)
    hy = H(y)
    lhs = torch.vdot(y.reshape(-1), hx.reshape(-1)).item()
    rhs = torch.vdot(x.reshape(-1), hy.reshape(-1)).item().conjugate()
    scale = (torch.linalg.vector_norm(hx) * torch.linalg.vector_norm(y)).item()
    print(f"  hermiticity {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) < 1e-12 * scale
    a, b = 0.3 - 0.7j, -1.1 + 0.2j
    print("norb=16 8a8b panels: H (a x + b y)", flush=True)
    hl = H(a * x + b * y)
    want = a * hx + b * hy
    err = (torch.linalg.vector_norm(hl - want) /
           torch.linalg.vector_norm(want)).item()
    print(f"  linearity {err:.2e}", flush=True)
    assert err < 1e-12
