"""The no-allocation apply entry point, fqehip_apply_noalloc, against
fqehip_apply through the C ABI.  Both run the same kernels; the first takes
its workspace from the library's own device array instead of the caller.

Tolerance: bit-identical results.  Every path exercised here is
deterministic: the FP64 and complex64 split paths and the unfused path of a
complex operator accumulate in a fixed order.  Two calls in a row, the
second on a smaller sector, check that nothing of the first call's array
contents leaks into the second.
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


def _inputs(norb, na, nb, seed, kind):
    rng = numpy.random.default_rng(seed)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    if kind == "complex":
        h1 = h1 + 0.05j * rng.normal(size=h1.shape)
        h1 = 0.5 * (h1 + h1.conj().T)
        h2 = h2 + 0.01j * rng.normal(size=h2.shape)
    g = fo.OracleGraph(na, nb, norb)
    c = fo.synthetic_coeff(g.lena, g.lenb, rng)
    cin = fo.synthetic_coeff(g.lena, g.lenb, rng)
    return h1, h2, c, cin


def _both(fqe, norb, na, nb, seed, kind, dtype):
    """(fqehip_apply result, fqehip_apply_noalloc result) as numpy."""
    import torch
    from fqe_amd import _lib, engine
    print(f"noalloc {kind} {dtype} norb={norb} {na}a{nb}b", flush=True)
    h1, h2, c, cin = _inputs(norb, na, nb, seed, kind)
    d = fqe.FqeData(na, nb, norb, dtype=dtype)
    graph = d.get_fcigraph()
    dev = _lib.device()
    tdt = torch.complex128 if dtype == numpy.complex128 else torch.complex64
    tc = torch.from_numpy(c).to(dev).to(tdt)
    tcin = torch.from_numpy(cin).to(dev).to(tdt)
    op = engine.prepare_operator((h1, h2), norb, dev)
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.1j
    ref = torch.empty_like(tc)
    engine.apply_sector(graph, tc, op, ref, alpha, beta, tcin)
    prec = engine.precision_of(tdt)
    s, keep = engine.sector_struct(graph, op, prec)
    opc = op.c_struct()
    out = torch.full_like(tc, float("nan"))
    _lib.call("fqehip_apply_noalloc", prec, ctypes.byref(s),
              ctypes.byref(opc), _lib.ptr(tc), alpha.real, alpha.imag,
              beta.real, beta.imag, _lib.ptr(tcin), _lib.ptr(out), None,
              None, None, _lib.stream_ptr())
    torch.cuda.synchronize()
    del keep
    return ref.cpu().numpy(), out.cpu().numpy()


@pytest.mark.parametrize("kind,dtype,big,small", [
    ("real", numpy.complex128, (8, 4, 4), (6, 3, 2)),
    ("real", numpy.complex64, (10, 5, 5), (7, 3, 3)),
    ("complex", numpy.complex128, (7, 3, 3), (5, 2, 1)),
])
def test_noalloc_matches_apply_two_calls(fqe, kind, dtype, big, small):
    for k, (norb, na, nb) in enumerate((big, small)):
        ref, out = _both(fqe, norb, na, nb, 100 + k, kind, dtype)
        assert numpy.all(numpy.isfinite(out))
        assert numpy.array_equal(ref, out), (
            numpy.abs(ref - out).max(), numpy.abs(ref).max())


def test_noalloc_rejects_sector_above_the_fixed_size(fqe):
    """norb=12 6a6b in complex128 needs (36+1)*2*853776*8 = 505 MB of
    workspace on the fused path, above the 256 MiB array: the call returns
    FQEHIP_ERR_UNSUPPORTED before any device work."""
    import torch
    from fqe_amd import _lib, engine
    norb, na, nb = 12, 6, 6
    rng = numpy.random.default_rng(5)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    d = fqe.FqeData(na, nb, norb, dtype=numpy.complex128)
    op = engine.prepare_operator((h1, h2), norb, _lib.device())
    s, keep = engine.sector_struct(d.get_fcigraph(), op, _lib.C128)
    opc = op.c_struct()
    lib = _lib.lib()
    rc = lib.fqehip_apply_noalloc(_lib.C128, ctypes.byref(s),
                                  ctypes.byref(opc), 16, 1.0, 0.0, 0.0, 0.0,
                                  None, 16, None, None, None, None)
    assert rc == _lib.ERR_UNSUPPORTED
    msg = lib.fqehip_last_error().decode()
    assert "505" in msg and "workspace" in msg, msg
    torch.cuda.synchronize()
    del keep
