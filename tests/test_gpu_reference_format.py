"""Wavefunction.read of a file in the reference's ``save`` layout (written
by tests/reference_pickles.py), on the GPU: the sectors come back bit for
bit, the header fields are kept, and the wavefunction applies a
RestrictedHamiltonian like one built from the same arrays (oracle to
1e-12)."""

import math

import numpy
import pytest

from oracle import fqe_oracle as fo

import reference_pickles as rp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


@pytest.mark.parametrize("protocol", [2, 4, 5])
def test_read_reference_file_and_apply(fqe, tmp_path, protocol):
    norb = 6
    rng = numpy.random.default_rng(60 + protocol)
    keys = [(6, 0), (6, 2), (5, 1)]
    sectors = {}
    for nele, sz in keys:
        na, nb = (nele + sz) // 2, (nele - sz) // 2
        shape = (math.comb(norb, na), math.comb(norb, nb))
        sectors[(nele, sz)] = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    path = tmp_path / "ref.wfn"
    path.write_bytes(rp.dumps(norb, sectors, protocol, conserved={"n": 6}))
    from fqe_amd.wavefunction import Wavefunction
    wfn = Wavefunction()
    wfn.read("ref.wfn", path=str(tmp_path))
    assert wfn.norb() == norb
    assert sorted(wfn.sectors()) == sorted(keys)
    for key, c in sectors.items():
        assert numpy.array_equal(wfn.get_coeff(key), c)
    h1, h2 = fo.synthetic_hamiltonian(norb, norb, rng)
    out = wfn.apply(fqe.get_restricted_hamiltonian((h1, h2)))
    for (nele, sz), c in sectors.items():
        g = fo.OracleGraph((nele + sz) // 2, (nele - sz) // 2, norb)
        ref = fo.apply_12(c, h1, h2, g)
        got = out.get_coeff((nele, sz))
        assert numpy.linalg.norm(got - ref) < 1e-12 * numpy.linalg.norm(ref)
