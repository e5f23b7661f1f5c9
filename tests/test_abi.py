"""C-ABI library checks that need no GPU: it loads, exports every entry
point the public header declares, and the host layer refuses to run
without a GPU instead of falling back to the CPU."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fqehip.h")
LIB = os.path.join(ROOT, "openfermion-fqe_amd", "fqe_amd", "libfqehip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fqehip_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfqehip.so not built")
def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (binds the library to torch's HIP runtime)
    handle = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfqehip.so not built")
def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import fqe_amd
    from fqe_amd import _lib
    assert _lib.lib().fqehip_version() >= 1
    with pytest.raises(_lib.FqeHipError):
        fqe_amd.Wavefunction([[2, 0, 2]])


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfqehip.so not built")
def test_argument_errors_are_reported_without_gpu_work():
    """Invalid sizes are rejected on the host with a status code and a
    message (the reference exits or prints instead)."""
    import torch  # noqa: F401
    from fqe_amd import _lib
    lib = _lib.lib()
    rc = lib.fqehip_string_tables(40, 3, 9880, None, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"norb" in lib.fqehip_last_error()
    rc = lib.fqehip_string_tables(6, 3, 21, None, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"C(6,3)" in lib.fqehip_last_error()
    assert lib.fqehip_norm_partials_count(0) == 0
    assert lib.fqehip_norm_partials_count(257) == 2


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfqehip.so not built")
def test_set_sizes_and_argument_errors_without_gpu_work():
    """The spin-orbital set entry points size the concatenated sectors as
    sum over nbeta of C(norb, nele - nbeta) C(norb, nbeta) = C(2 norb, nele)
    and reject bad sizes and pointers on the host."""
    import math
    import torch  # noqa: F401
    from fqe_amd import _lib
    lib = _lib.lib()
    for norb in range(1, 9):
        for nele in range(0, 2 * norb + 1):
            n = lib.fqehip_set_length(norb, nele)
            assert n == math.comb(2 * norb, nele)
            m2 = (2 * norb) ** 2
            assert lib.fqehip_set_workspace_bytes(norb, nele) == (
                ((4 * m2 + 2) * n + (2 * m2 + 2) * 2 * m2) * 8)
    assert lib.fqehip_set_length(17, 2) == -1
    assert lib.fqehip_set_length(3, 7) == -1
    rc = lib.fqehip_set_apply(17, 2, None, None, None, None, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"norb" in lib.fqehip_last_error()
    rc = lib.fqehip_set_apply(3, 3, None, None, None, None, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"null" in lib.fqehip_last_error()
    rc = lib.fqehip_set_apply(3, 3, 8, 8, 16, 32, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"workspace" in lib.fqehip_last_error()
    rc = lib.fqehip_set_apply(3, 3, 8, None, 16, 16, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"alias" in lib.fqehip_last_error()


def _header_prototypes():
    """name -> number of parameters of each declared entry point."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(fqehip_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = m.group(2).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        protos[m.group(1)] = n
    return protos


def test_ctypes_signatures_match_header():
    """Every binding in fqe_amd._lib passes exactly the parameters the C
    prototype declares (a missing argument shifts every later one)."""
    from fqe_amd import _lib
    protos = _header_prototypes()
    assert set(_lib._SIGNATURES) == set(protos), (
        set(_lib._SIGNATURES) ^ set(protos))
    for name, (_, argtypes) in _lib._SIGNATURES.items():
        assert len(argtypes) == protos[name], name
