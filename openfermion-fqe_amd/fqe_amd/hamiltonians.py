"""Hamiltonian containers for the apply path.

Mirrors ``fqe.hamiltonians.hamiltonian.Hamiltonian`` (reference
``src/fqe/hamiltonians/hamiltonian.py:26-135``) and
``fqe.hamiltonians.restricted_hamiltonian.RestrictedHamiltonian``
(``src/fqe/hamiltonians/restricted_hamiltonian.py:23-155``): same
constructor, ``tensors()``, ``iht()``, ``e_0()``, ``dim()``, ``rank()``,
``quadratic()`` and error behaviour.  The device form of the operator
(pair-basis GEMM matrix) is cached on the object.
"""

from __future__ import annotations

from abc import ABCMeta, abstractmethod
from typing import Any, Dict, Tuple

import numpy


class Hamiltonian(metaclass=ABCMeta):
    """Abstract base (reference ``src/fqe/hamiltonians/hamiltonian.py:26``)."""

    def __init__(self, e_0: complex = 0.0 + 0.0j):
        self._conserve_number = True
        self._e_0 = e_0

    @abstractmethod
    def dim(self) -> int:
        return 0

    def calc_diag_transform(self) -> numpy.ndarray:
        return numpy.empty(0)

    @abstractmethod
    def rank(self) -> int:
        return 0

    def quadratic(self) -> bool:
        return False

    def diagonal(self) -> bool:
        return False

    def diagonal_coulomb(self) -> bool:
        return False

    def conserve_number(self) -> bool:
        return self._conserve_number

    def e_0(self):
        return self._e_0

    def iht(self, time: float) -> Any:
        return tuple()

    def tensors(self) -> Tuple[numpy.ndarray, ...]:
        return tuple()

    def diag_values(self) -> numpy.ndarray:
        return numpy.empty(0)

    def transform(self, trans: numpy.ndarray) -> numpy.ndarray:
        return numpy.empty(0)


class RestrictedHamiltonian(Hamiltonian):
    """Spin-restricted Hamiltonian given by spin-free tensors (reference
    ``src/fqe/hamiltonians/restricted_hamiltonian.py:23-155``)."""

    def __init__(self, tensors: Tuple[numpy.ndarray, ...],
                 e_0: complex = 0.0 + 0.0j) -> None:
        super().__init__(e_0=e_0)
        self._tensor: Dict[int, numpy.ndarray] = {}
        # device operators of the sector-by-sector apply, per device
        self._prepared: Dict[Any, Any] = {}
        for rank, matrix in enumerate(tensors):
            if not (isinstance(rank, int) and
                    isinstance(matrix, numpy.ndarray)):
                raise TypeError("tensors should be a tuple of numpy.ndarray")
            if matrix.ndim % 2:
                raise ValueError("input tensor has an odd rank")
            self._tensor[2 * (rank + 1)] = matrix
        assert self._tensor, (
            "No matrix elements passed into the RestrictedHamiltonian.")
        self._quadratic = len(self._tensor) == 1 and 2 in self._tensor
        self._dim = list(self._tensor.values())[0].shape[0]
        self._prepared: Dict[Any, Any] = {}

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, RestrictedHamiltonian):
            return NotImplemented
        if self.e_0() != other.e_0():
            return False
        if self._tensor.keys() != other._tensor.keys():
            return False
        return all(numpy.allclose(self._tensor[k], other._tensor[k])
                   for k in self._tensor)

    __hash__ = None

    def dim(self) -> int:
        return self._dim

    def rank(self) -> int:
        return 2 * len(self._tensor)

    def tensor(self, rank: int) -> numpy.ndarray:
        return self._tensor[rank]

    def tensors(self) -> Tuple[numpy.ndarray, ...]:
        return tuple(self._tensor[2 * (r + 1)] for r in range(len(self._tensor)))

    def quadratic(self) -> bool:
        return self._quadratic

    def iht(self, time: float) -> Tuple[numpy.ndarray, ...]:
        return tuple(-1.0j * time * self._tensor[2 * (r + 1)]
                     for r in range(len(self._tensor)))

    def calc_diag_transform(self) -> numpy.ndarray:
        _, trans = numpy.linalg.eigh(self._tensor[2])
        return trans

    def transform(self, trans: numpy.ndarray) -> numpy.ndarray:
        return trans.conj().T @ self._tensor[2] @ trans


class _SpinOrbitalHamiltonian(Hamiltonian):
    """Tensors over 2 norb spin orbitals (p < norb alpha, p >= norb beta):
    h1[p, q] a+_p a_q + h2[p, q, r, s] a+_p a+_q a_r a_s + ...  Shared body
    of ``General`` and ``GSOHamiltonian`` (reference
    ``src/fqe/hamiltonians/general_hamiltonian.py:23-120``,
    ``src/fqe/hamiltonians/gso_hamiltonian.py:23-120``)."""

    _label = "spin-orbital Hamiltonian"

    def __init__(self, tensors: Tuple[numpy.ndarray, ...],
                 e_0: complex = 0.0 + 0.0j) -> None:
        super().__init__(e_0=e_0)
        self._tensor: Dict[int, numpy.ndarray] = {}
        # device operators of the sector-by-sector apply, per device
        self._prepared: Dict[Any, Any] = {}
        for rank, matrix in enumerate(tensors):
            if not isinstance(matrix, numpy.ndarray):
                raise TypeError(
                    "Arg tensors should be a tuple of numpy.ndarray, but "
                    f"tensors[{rank}] = {type(tensors[rank])}.")
            if matrix.ndim % 2:
                raise ValueError("input tensor has an odd rank")
            self._tensor[2 * (rank + 1)] = matrix
        assert self._tensor, f"No matrix elements passed into the {self._label}."
        self._quadratic = len(self._tensor) == 1 and 2 in self._tensor
        self._dim = list(self._tensor.values())[0].shape[0]

    def __eq__(self, other: object) -> bool:
        if type(other) is not type(self):
            return NotImplemented
        if self.e_0() != other.e_0():
            return False
        if self._tensor.keys() != other._tensor.keys():
            return False
        return all(numpy.allclose(self._tensor[k], other._tensor[k])
                   for k in self._tensor)

    __hash__ = None

    def dim(self) -> int:
        return self._dim

    def rank(self) -> int:
        return 2 * len(self._tensor)

    def tensor(self, rank: int) -> numpy.ndarray:
        return self._tensor[rank]

    def tensors(self) -> Tuple[numpy.ndarray, ...]:
        return tuple(self._tensor[2 * (r + 1)] for r in range(len(self._tensor)))

    def quadratic(self) -> bool:
        return self._quadratic

    def iht(self, time: float) -> Tuple[numpy.ndarray, ...]:
        return tuple(-1.0j * time * self._tensor[2 * (r + 1)]
                     for r in range(len(self._tensor)))


class General(_SpinOrbitalHamiltonian):
    """reference ``src/fqe/hamiltonians/general_hamiltonian.py:23``"""

    _label = "GeneralHamiltonian"


class GSOHamiltonian(_SpinOrbitalHamiltonian):
    """reference ``src/fqe/hamiltonians/gso_hamiltonian.py:23``"""

    _label = "GSOHamiltonian"


class SSOHamiltonian(_SpinOrbitalHamiltonian):
    """Distinct alpha and beta terms, no alpha/beta mixing blocks (reference
    ``src/fqe/hamiltonians/sso_hamiltonian.py:26``)."""

    _label = "SSOHamiltonian"
