"""Writers of files in the reference's ``Wavefunction.save`` layout (a
pickled list [symmetry map, conserved, conserve_spin, conserve_number, norb,
[key, fqe.fqe_data.FqeData], ...], reference src/fqe/wavefunction.py:
748-765), for the tests of the no-execution reader.  The reference's
classes are stood in for by small local classes registered under the
reference's module names only while ``pickle.dump`` runs, so that the
stream names fqe.fqe_data.FqeData and fqe.fci_graph.FciGraph as the
reference's own files do; their attributes follow the reference's
FqeData.__init__ (src/fqe/fqe_data.py:60-107) and FciGraph.__init__
(src/fqe/fci_graph.py:99-154)."""

import contextlib
import math
import pickle
import sys
import types

import numpy


class FciGraph:
    pass


class FqeData:
    pass


@contextlib.contextmanager
def _reference_modules():
    saved = {k: sys.modules.get(k) for k in ("fqe", "fqe.fqe_data",
                                              "fqe.fci_graph")}
    fqe = types.ModuleType("fqe")
    data = types.ModuleType("fqe.fqe_data")
    graph = types.ModuleType("fqe.fci_graph")
    FqeData.__module__, FqeData.__qualname__ = "fqe.fqe_data", "FqeData"
    FciGraph.__module__, FciGraph.__qualname__ = "fqe.fci_graph", "FciGraph"
    data.FqeData = FqeData
    graph.FciGraph = FciGraph
    fqe.fqe_data, fqe.fci_graph = data, graph
    sys.modules.update({"fqe": fqe, "fqe.fqe_data": data,
                        "fqe.fci_graph": graph})
    try:
        yield
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v


def _graph(nalpha, nbeta, norb):
    g = FciGraph()
    g._norb, g._nalpha, g._nbeta = norb, nalpha, nbeta
    g._lena = math.comb(norb, nalpha)
    g._lenb = math.comb(norb, nbeta)
    g._astr = numpy.arange(g._lena, dtype=numpy.uint64)
    g._aind = {int(s): i for i, s in enumerate(g._astr)}
    g._alpha_map = {(0, 1): numpy.array([[0, 1, -1]], dtype=numpy.int32)}
    return g


def sector(nalpha, nbeta, norb, coeff):
    d = FqeData()
    d._core = _graph(nalpha, nbeta, norb)
    d._dtype = coeff.dtype.type
    d._low_thresh = 0.3
    d._nele = nalpha + nbeta
    d._m_s = nalpha - nbeta
    d.coeff = coeff
    return d


def dumps(norb, sectors, protocol, conserve_spin=True, conserve_number=True,
          conserved=None, symmetry_map=None):
    """Bytes of a reference-layout file; ``sectors`` maps (nele, sz) to a
    coefficient array."""
    data = [symmetry_map if symmetry_map is not None else {},
            conserved if conserved is not None else {},
            conserve_spin, conserve_number, norb]
    for (nele, sz), c in sectors.items():
        data.append([(nele, sz), sector((nele + sz) // 2, (nele - sz) // 2,
                                        norb, c)])
    with _reference_modules():
        return pickle.dumps(data, protocol=protocol)
