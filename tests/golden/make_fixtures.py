"""Writes the golden fixtures under tests/golden/ from the reference's own
test data.  Nothing of the reference is imported or executed:

* ``tests/unittest_data/fqe_data/*.npy`` are raw float64 dumps
  (``ndarray.tofile``, reference
  ``tests/unittest_data/fqe_data/generate_data.py:47-95``; read back with
  ``numpy.fromfile`` as ``tests/unittest_data/fqe_data_loader.py:28-31``
  does).  Copied as arrays: the FqeData(2,3,6), (2,1,4) and (1,1,2)
  coefficients, h1, h2 and the FQE outputs of apply((h1,)), apply((0,h2))
  and apply((h1,h2)); and the 1- and 2-RDMs of FqeData(2,1,4).
* literal tables are read from the reference test sources with
  ``ast.literal_eval`` on their literal nodes: strings and addresses of
  ``tests/fci_graph_test.py:33-112``, the n=4 maps of ``:151-190``, and the
  eigenvector / eigenvalue of ``tests/fqe_data_test.py:761-826``.
* the LiH (norb=6, 2a2b) h1e, h2e and ground-state wavefunction literals of
  ``tests/unittest_data/build_lih_data.py:30-1270`` (energy branch), whose
  FCI energy -8.877719570384043 (``tests/davidson_test.py:27``) and HF
  energy -8.857341498221992 (``tests/wavefunction_test.py:537-548``) the
  Davidson tests use.
* ``tests/unittest_data/fqe_data_set/set0202/*.npy`` (raw float64 dumps,
  reference ``tests/unittest_data/fqe_data_set/generate_data.py:25-35``,
  ``:54-106``, ``:108-120``; read back with ``numpy.fromfile`` as
  ``tests/unittest_data/fqe_data_set_loader.py:26-36`` does): the
  FqeDataSet(nele=2, norb=2) sector coefficients, the outputs of
  apply((h1,)), apply((0, h2)) and apply((h1, h2)), and the spin-orbital
  1- and 2-RDMs.  The spin-orbital h1 and h2 those outputs were made with
  are rebuilt from their closed form (reference
  ``tests/unittest_data/build_hamiltonian.py:67-119``, ``:230-286``).

Run from the repository root:  python tests/golden/make_fixtures.py
"""

import ast
import json
import os

import numpy

REF = "/root/reference/tests"
OUT = os.path.dirname(os.path.abspath(__file__))

SECTORS = [(2, 3, 6), (2, 1, 4), (1, 1, 2)]


def raw(name):
    return numpy.fromfile(os.path.join(REF, "unittest_data", "fqe_data", name))


def fqe_data_arrays():
    from math import comb
    arrays = {}
    for na, nb, n in SECTORS:
        tag = f"{na:02d}{nb:02d}{n:02d}"
        shape = (comb(n, na), comb(n, nb))
        arrays[f"c_{tag}"] = (raw(f"cr{tag}.npy") + 1j * raw(f"ci{tag}.npy")
                              ).reshape(shape)
        arrays[f"h1_{tag}"] = raw(f"h1{tag}.npy").reshape(n, n)
        arrays[f"h2_{tag}"] = raw(f"h2{tag}.npy").reshape(n, n, n, n)
        if tag == "020104":
            # FqeData(2,1,4).rdm1() / rdm12() outputs (generate_data.py:114-123)
            for order in (1, 2):
                shape_r = (n,) * (2 * order)
                arrays[f"rdm{order}_{tag}"] = (
                    raw(f"dr{order}{tag}.npy") + 1j * raw(f"di{order}{tag}.npy")
                ).reshape(shape_r)
        for suffix in ("1", "2", "12"):
            arrays[f"out{suffix}_{tag}"] = (
                raw(f"cr{tag}_{suffix}.npy") + 1j * raw(f"ci{tag}_{suffix}.npy")
            ).reshape(shape)
    return arrays


def spin_orbital_h(norb):
    """h1[i,j] = 2i delta_ij + 0.02 (i + j), h2[i,j,k,l] = 0.02 (i + k)
    (j + l) on spatial orbitals, placed on the alpha-alpha, beta-beta and
    mixed blocks of the spin-orbital tensors with unit prefactors."""
    idx = numpy.arange(norb, dtype=float)
    h1 = numpy.diag(2.0 * idx) + 0.02 * (idx[:, None] + idx[None, :])
    ik = idx[:, None] + idx[None, :]
    h2 = 0.02 * numpy.einsum("ik,jl->ijkl", ik, ik)
    m = 2 * norb
    h1s = numpy.zeros((m, m))
    h1s[:norb, :norb] = h1
    h1s[norb:, norb:] = h1
    h2s = numpy.zeros((m,) * 4)
    a, b = slice(0, norb), slice(norb, m)
    for s1, s2 in ((a, a), (a, b), (b, a), (b, b)):
        h2s[s1, s2, s1, s2] = h2
    return h1s, h2s


def fqe_data_set_arrays(nele=2, norb=2):
    from math import comb
    base = os.path.join(REF, "unittest_data", "fqe_data_set",
                        f"set{nele:02d}{norb:02d}")

    def rd(name):
        return numpy.fromfile(os.path.join(base, name))

    arrays = {}
    for nbeta in range(nele + 1):
        nalpha = nele - nbeta
        tag = f"{nalpha:02d}{nbeta:02d}{norb:02d}"
        shape = (comb(norb, nalpha), comb(norb, nbeta))
        arrays[f"c_{tag}"] = (rd(f"cr{tag}.npy") + 1j * rd(f"ci{tag}.npy")
                              ).reshape(shape)
        for suffix in ("1", "2", "12"):
            arrays[f"out{suffix}_{tag}"] = (
                rd(f"cr{tag}_{suffix}.npy") + 1j * rd(f"ci{tag}_{suffix}.npy")
            ).reshape(shape)
    m = 2 * norb
    for order in (1, 2):
        arrays[f"rdm{order}"] = (rd(f"dr{order}.npy") + 1j * rd(f"di{order}.npy")
                                 ).reshape((m,) * (2 * order))
    arrays["h1"], arrays["h2"] = spin_orbital_h(norb)
    return arrays


def _find_assign(tree, func_name, var):
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == func_name:
            for sub in ast.walk(node):
                if (isinstance(sub, ast.Assign) and len(sub.targets) == 1
                        and isinstance(sub.targets[0], ast.Name)
                        and sub.targets[0].id == var):
                    return sub.value
    raise KeyError((func_name, var))


def _literal(node):
    # numpy.array([...], dtype=...) / numpy.asarray(...): take the first arg
    if isinstance(node, ast.Call):
        node = node.args[0]
    return ast.literal_eval(node)


def literal_tables():
    with open(os.path.join(REF, "fci_graph_test.py")) as f:
        tree = ast.parse(f.read())
    refdata = _literal(_find_assign(tree, "test_fci_graph", "refdata"))
    refdict = _literal(_find_assign(tree, "test_fci_graph", "refdict"))
    amap = _literal(_find_assign(tree, "test_fci_graph_maps", "ref_alpha_map"))
    bmap = _literal(_find_assign(tree, "test_fci_graph_maps", "ref_beta_map"))
    with open(os.path.join(REF, "fqe_data_test.py")) as f:
        tree = ast.parse(f.read())
    wfn = _literal(_find_assign(tree, "test_2_body", "wfn"))
    scale = _literal(_find_assign(tree, "test_2_body", "scale"))
    return {
        "strings_n8_k4": refdata,
        "index_n8_k4": {str(k): v for k, v in refdict.items()},
        "maps_n4_k2": {f"{i},{j}": v for (i, j), v in amap.items()},
        "maps_n4_k1": {f"{i},{j}": v for (i, j), v in bmap.items()},
        "eigvec_2_2_4": [[[z.real, z.imag] for z in row] for row in wfn],
        "eigval_2_2_4": scale,
    }


def lih_arrays():
    with open(os.path.join(REF, "unittest_data", "build_lih_data.py")) as f:
        tree = ast.parse(f.read())
    out = {}
    for var in ("wfn", "h1e", "h2e"):
        out[var] = numpy.asarray(_literal(_find_assign(tree, "build_lih_data",
                                                       var)))
    return out


def main():
    numpy.savez(os.path.join(OUT, "fqe_data_golden.npz"), **fqe_data_arrays())
    numpy.savez(os.path.join(OUT, "lih_energy.npz"), **lih_arrays())
    numpy.savez(os.path.join(OUT, "fqe_data_set_golden.npz"),
                **fqe_data_set_arrays())
    with open(os.path.join(OUT, "literal_tables.json"), "w") as f:
        json.dump(literal_tables(), f, indent=1)


if __name__ == "__main__":
    main()
