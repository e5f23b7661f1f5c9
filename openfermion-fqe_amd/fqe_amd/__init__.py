"""fqe_amd: MI355X engine for OpenFermion-FQE's RestrictedHamiltonian apply
and time evolution.

Drop-in subset of the ``fqe`` namespace for that path (reference
``src/fqe/_fqe_control.py:145-198``, ``:356-371``, ``:559-575``):

    import fqe_amd as fqe
    wfn = fqe.Wavefunction([[nele, sz, norb]])
    h = fqe.get_restricted_hamiltonian((h1, h2), e_0)
    out = fqe.apply(h, wfn);  evolved = fqe.time_evolve(wfn, t, h)
"""

from . import settings
from .fci_graph import FciGraph
from .fqe_data import FqeData
from .fqe_data_set import FqeDataSet
from .hamiltonians import (Hamiltonian, RestrictedHamiltonian, General,
                           GSOHamiltonian, SSOHamiltonian)
from .util import vdot, alpha_beta_electrons
from .wavefunction import Wavefunction as _Wavefunction
from .wavefunction import build_hamiltonian

__version__ = "0.1.0"


def Wavefunction(param, broken=None, dtype=None) -> _Wavefunction:
    """reference ``src/fqe/_fqe_control.py:145-160``"""
    return _Wavefunction(param, broken=broken, dtype=dtype)


def get_wavefunction(nele: int, m_s: int, norb: int,
                     dtype=None) -> _Wavefunction:
    """reference ``src/fqe/_fqe_control.py:163-178``"""
    return _Wavefunction([[nele, m_s, norb]], dtype=dtype)


def get_general_hamiltonian(tensors, e_0: complex = 0. + 0.j) -> General:
    """reference ``src/fqe/_fqe_control.py:526-539``"""
    return General(tensors, e_0=e_0)


def get_gso_hamiltonian(tensors, e_0: complex = 0. + 0.j) -> GSOHamiltonian:
    """reference ``src/fqe/_fqe_control.py:542-556``"""
    return GSOHamiltonian(tensors, e_0=e_0)


def get_sso_hamiltonian(tensors, e_0: complex = 0. + 0.j) -> SSOHamiltonian:
    """reference ``src/fqe/_fqe_control.py:599-613``"""
    return SSOHamiltonian(tensors, e_0=e_0)


def get_restricted_hamiltonian(tensors, e_0: complex = 0. + 0.j
                               ) -> RestrictedHamiltonian:
    """reference ``src/fqe/_fqe_control.py:559-575``"""
    return RestrictedHamiltonian(tensors, e_0=e_0)


def apply(ops, wfn: _Wavefunction) -> _Wavefunction:
    """reference ``src/fqe/_fqe_control.py:356-371``"""
    return wfn.apply(ops)


def time_evolve(wfn: _Wavefunction, time: float, hamil,
                inplace: bool = False) -> _Wavefunction:
    """reference ``src/fqe/_fqe_control.py:180-198``"""
    return wfn.time_evolve(time, hamil, inplace)


def expectationValue(wfn: _Wavefunction, ops, brawfn=None):
    """reference ``src/fqe/_fqe_control.py:374-391``"""
    return wfn.expectationValue(ops, brawfn)


__all__ = [
    "FciGraph", "FqeData", "Hamiltonian", "RestrictedHamiltonian",
    "Wavefunction", "get_wavefunction", "get_restricted_hamiltonian",
    "apply", "time_evolve", "expectationValue", "vdot",
    "alpha_beta_electrons", "build_hamiltonian", "settings",
]
