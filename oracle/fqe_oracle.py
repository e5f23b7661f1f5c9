"""CPU oracle for the determinant-space RestrictedHamiltonian apply.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker.  The product path (``fqe_amd``) never imports it and has no CPU
fallback.

This is a NumPy (complex128 / float64) restatement, written from the
mathematics of the reference's algorithm, of:

* the Knowles-Handy string addressing and excitation tables of
  ``FciGraph`` (reference ``src/fqe/fci_graph.py:36-96``, ``:204-244``,
  ``:301-333``, ``:401-419``; ``src/fqe/lib/fci_graph.c:27-147``;
  ``src/fqe/lib/bitstring.h:23-76``; ``src/fqe/lib/bitstring.c:29-46``);
* the spatial 1+2-body apply in its dvec form
  (``src/fqe/fqe_data.py:646-657``, ``:2209-2234``, ``:2309-2334``) and the
  1-body apply (``src/fqe/fqe_data.py:477-531``);
* ``Wavefunction.apply`` / ``_apply_array`` (``src/fqe/wavefunction.py:314-399``),
  the Taylor branch of ``apply_generated_unitary``
  (``src/fqe/wavefunction.py:509-611``) and the dense branch of
  ``time_evolve`` (``src/fqe/wavefunction.py:961-1054``), including the
  reference's doubled ``exp(-i t e_0)`` phase;
* the Slater-Condon diagonal used as the Davidson preconditioner
  (replacing the one-apply-per-determinant loop of
  ``src/fqe/algorithm/davidson.py:162-180``).

How each function is pinned (tests/test_oracle.py):

* string tables: literal tables of ``tests/fci_graph_test.py:33-112`` and
  ``:151-190``; an independent Knowles-Handy Z-matrix address; the
  reference's compiled C (``oracle/_ref``) for strings, maps and dexc.
* apply: FQE's float64 golden outputs (``tests/unittest_data/fqe_data``),
  the closed-form eigenvector of ``tests/fqe_data_test.py:761-826``, a
  Jordan-Wigner dense Hamiltonian built independently (``oracle/jw_dense.py``)
  and the reference's compiled lm kernels (``oracle/_ref``).
* Taylor evolution: exact ``exp(-iHt)`` from ``eigh`` of the dense matrix.
* Slater-Condon diagonal: the diagonal of the Jordan-Wigner matrix.
"""

from __future__ import annotations

import itertools
import math
from typing import Dict, Tuple

import numpy

Spinmap = Dict[Tuple[int, int], numpy.ndarray]


# ---------------------------------------------------------------------------
# (a) string addressing and excitation tables
# ---------------------------------------------------------------------------

def binom(n: int, k: int) -> int:
    if k < 0 or k > n or n < 0:
        return 0
    return math.comb(n, k)


def build_strings(norb: int, nele: int) -> numpy.ndarray:
    """All ``nele``-electron bitstrings of ``norb`` orbitals in address order.

    The address of a string is the rank of its ascending occupation tuple
    in lexicographic order (``itertools.combinations`` order), which is
    what the reference's Z-matrix address produces
    (``src/fqe/fci_graph.py:301-333``, ``src/fqe/lib/fci_graph.c:123-147``).
    """
    out = []
    for occ in itertools.combinations(range(norb), nele):
        s = 0
        for o in occ:
            s |= 1 << o
        out.append(s)
    return numpy.asarray(out, dtype=numpy.uint64).reshape(-1)


def z_matrix(norb: int, nele: int) -> numpy.ndarray:
    """Knowles-Handy Z matrix (reference ``src/fqe/fci_graph.py:66-96``).

    Z[k-1, l-1] for electron k (1-based) sitting in orbital l (1-based) is
    the number of strings that precede it because of that placement:
    Z[k-1,l-1] = sum_{m=norb-l+1}^{norb-k} [C(m, nele-k) - C(m-1, nele-k-1)]
    for k < nele, and Z[nele-1, l-1] = l - nele.
    """
    z = numpy.zeros((nele, norb), dtype=numpy.int64)
    if z.size == 0:
        return z
    for k in range(1, nele):
        for ll in range(k, norb - nele + k + 1):
            z[k - 1, ll - 1] = sum(
                binom(m, nele - k) - binom(m - 1, nele - k - 1)
                for m in range(norb - ll + 1, norb - k + 1))
    for ll in range(nele, norb + 1):
        z[nele - 1, ll - 1] = ll - nele
    return z


def z_address(string: int, z: numpy.ndarray) -> int:
    """Address of ``string`` as sum_k Z[k, occ_k] (reference
    ``src/fqe/lib/fci_graph.c:123-134``)."""
    addr = 0
    k = 0
    for orb in range(z.shape[1]):
        if (string >> orb) & 1:
            addr += int(z[k, orb])
            k += 1
    return addr


def bits_between(string: int, i: int, j: int) -> int:
    """Number of occupied orbitals strictly between i and j
    (reference ``src/fqe/lib/bitstring.h:61-65``)."""
    lo, hi = min(i, j), max(i, j)
    mask = ((1 << hi) - 1) & ~((1 << (lo + 1)) - 1)
    return bin(string & mask).count("1")


def build_maps(strings: numpy.ndarray, norb: int) -> Spinmap:
    """Excitation maps of a^dagger_i a_j (reference
    ``src/fqe/fci_graph.py:204-244``, ``src/fqe/lib/fci_graph.c:76-121``).

    For each ordered pair (i, j), the list of (source, target, sign) with
    target = a^dagger_i a_j source, in ascending source address, sign =
    (-1)^(occupied orbitals strictly between i and j in source);
    i == j lists (s, s, +1) for every string with i occupied.
    """
    index = {int(s): a for a, s in enumerate(strings)}
    out: Spinmap = {}
    for i in range(norb):
        for j in range(norb):
            rows = []
            for src_addr, s in enumerate(strings):
                s = int(s)
                if (s >> j) & 1 and not (s >> i) & 1:
                    t = (s | (1 << i)) & ~(1 << j)
                    sign = -1 if bits_between(s, i, j) % 2 else 1
                    rows.append((src_addr, index[t], sign))
                elif i == j and (s >> i) & 1:
                    rows.append((src_addr, src_addr, 1))
            out[(i, j)] = numpy.asarray(rows, dtype=numpy.int32).reshape(-1, 3)
    return out


def build_dexc(maps: Spinmap, nstrings: int, norb: int,
               nele: int) -> numpy.ndarray:
    """Per-target de-excitation table (reference
    ``src/fqe/fci_graph.py:36-63``, ``src/fqe/lib/fci_graph.c:57-73``).

    dexc[target] holds nele*(norb-nele+1) rows (source, i*norb+j, sign),
    in ascending i*norb+j.
    """
    nexc = nele * (norb - nele + 1)
    dexc = numpy.zeros((nstrings, nexc, 3), dtype=numpy.int32)
    fill = numpy.zeros(nstrings, dtype=numpy.int64)
    for i in range(norb):
        for j in range(norb):
            for src, tgt, sign in maps[(i, j)]:
                dexc[tgt, fill[tgt]] = (src, i * norb + j, sign)
                fill[tgt] += 1
    assert nstrings == 0 or numpy.all(fill == nexc)
    return dexc


class OracleGraph:
    """Tables of one (nalpha, nbeta, norb) sector, mirroring ``FciGraph``
    (reference ``src/fqe/fci_graph.py:99-154``)."""

    def __init__(self, nalpha: int, nbeta: int, norb: int) -> None:
        self.norb = norb
        self.nalpha = nalpha
        self.nbeta = nbeta
        self.astr = build_strings(norb, nalpha)
        self.bstr = build_strings(norb, nbeta)
        self.lena = self.astr.size
        self.lenb = self.bstr.size
        self.alpha_map = build_maps(self.astr, norb)
        self.beta_map = build_maps(self.bstr, norb)
        self.dexca = build_dexc(self.alpha_map, self.lena, norb, nalpha)
        self.dexcb = build_dexc(self.beta_map, self.lenb, norb, nbeta)


# ---------------------------------------------------------------------------
# (b)-(d) the apply
# ---------------------------------------------------------------------------

def prepare_tensors(h1e: numpy.ndarray, h2e: numpy.ndarray):
    """h1' and V of the dvec form (reference ``src/fqe/fqe_data.py:647-651``).

    V[i,j,k,l] = -h2e[i,k,j,l];  h1'[i,j] = h1e[i,j] - sum_k V[i,k,k,j].
    """
    v = -numpy.moveaxis(numpy.asarray(h2e, dtype=numpy.complex128), 1, 2)
    h1p = numpy.asarray(h1e, dtype=numpy.complex128).copy()
    h1p -= numpy.einsum("ikkj->ij", v)
    return h1p, v


def dvec(coeff: numpy.ndarray, g: OracleGraph) -> numpy.ndarray:
    """D[i,j] = E_ij C (reference ``src/fqe/fqe_data.py:2209-2234``)."""
    n = g.norb
    out = numpy.zeros((n, n, g.lena, g.lenb), dtype=numpy.complex128)
    for i in range(n):
        for j in range(n):
            am = g.alpha_map[(i, j)]
            if am.size:
                out[i, j, am[:, 1], :] += coeff[am[:, 0], :] * am[:, 2:3]
            bm = g.beta_map[(i, j)]
            if bm.size:
                out[i, j, :, bm[:, 1]] += (coeff[:, bm[:, 0]] * bm[:, 2]).T
    return out


def coeff_from_dvec(y: numpy.ndarray, g: OracleGraph) -> numpy.ndarray:
    """out = sum_ij E_ij Y[i,j] (reference ``src/fqe/fqe_data.py:2309-2334``:
    maps taken as (j, i) and used from target to source)."""
    n = g.norb
    out = numpy.zeros((g.lena, g.lenb), dtype=numpy.complex128)
    for i in range(n):
        for j in range(n):
            am = g.alpha_map[(j, i)]
            if am.size:
                numpy.add.at(out, am[:, 0], y[i, j, am[:, 1], :] * am[:, 2:3])
            bm = g.beta_map[(j, i)]
            if bm.size:
                contrib = y[i, j][:, bm[:, 1]] * bm[:, 2]
                for c, s in enumerate(bm[:, 0]):
                    out[:, s] += contrib[:, c]
    return out


def apply_12(coeff: numpy.ndarray, h1e: numpy.ndarray, h2e: numpy.ndarray,
             g: OracleGraph) -> numpy.ndarray:
    """Spatial 1+2-body apply, dvec form (reference
    ``src/fqe/fqe_data.py:646-657``): out = sum h1' D + sum E_ij (V . D)."""
    h1p, v = prepare_tensors(h1e, h2e)
    d = dvec(numpy.asarray(coeff, dtype=numpy.complex128), g)
    out = numpy.einsum("ij,ijkl->kl", h1p, d)
    y = numpy.einsum("ijkl,klmn->ijmn", v, d)
    out += coeff_from_dvec(y, g)
    return out


def apply_1(coeff: numpy.ndarray, h1e: numpy.ndarray,
            g: OracleGraph) -> numpy.ndarray:
    """Spatial 1-body apply: out = sum_ij h1[i,j] E_ij C (reference
    ``src/fqe/fqe_data.py:477-531``)."""
    d = dvec(numpy.asarray(coeff, dtype=numpy.complex128), g)
    return numpy.einsum("ij,ijkl->kl", numpy.asarray(h1e), d)


def apply_tensors(coeff, tensors, g: OracleGraph) -> numpy.ndarray:
    """``FqeData.apply_inplace`` for 1- or 2-tensor spatial tuples
    (reference ``src/fqe/fqe_data.py:421-475``)."""
    if len(tensors) == 1:
        return apply_1(coeff, tensors[0], g)
    if len(tensors) == 2:
        return apply_12(coeff, tensors[0], tensors[1], g)
    raise ValueError("oracle handles 1- and 2-body restricted tensors")


# ---------------------------------------------------------------------------
# spin-orbital tensors on one (nalpha, nbeta) sector
# ---------------------------------------------------------------------------

def dvec_spin(coeff: numpy.ndarray, g: OracleGraph):
    """(Da, Db): Da[i,j] = E^alpha_ij C and Db[i,j] = E^beta_ij C (reference
    ``src/fqe/fqe_data.py:2162-2175``, ``:2236-2285``)."""
    n = g.norb
    da = numpy.zeros((n, n, g.lena, g.lenb), dtype=numpy.complex128)
    db = numpy.zeros_like(da)
    for i in range(n):
        for j in range(n):
            am = g.alpha_map[(i, j)]
            if am.size:
                da[i, j, am[:, 1], :] += coeff[am[:, 0], :] * am[:, 2:3]
            bm = g.beta_map[(i, j)]
            if bm.size:
                db[i, j, :, bm[:, 1]] += (coeff[:, bm[:, 0]] * bm[:, 2]).T
    return da, db


def coeff_from_dvec_spin(ya: numpy.ndarray, yb: numpy.ndarray,
                         g: OracleGraph) -> numpy.ndarray:
    """out = sum_ij E^alpha_ij Ya[i,j] + E^beta_ij Yb[i,j] (reference
    ``src/fqe/fqe_data.py:2355-2383``, _calculate_coeff_spin_with_dvec:
    maps taken as (j, i) and used from target to source, as in the
    spatial case)."""
    n = g.norb
    out = numpy.zeros((g.lena, g.lenb), dtype=numpy.complex128)
    for i in range(n):
        for j in range(n):
            am = g.alpha_map[(j, i)]
            if am.size:
                numpy.add.at(out, am[:, 0], ya[i, j, am[:, 1], :] * am[:, 2:3])
            bm = g.beta_map[(j, i)]
            if bm.size:
                contrib = yb[i, j][:, bm[:, 1]] * bm[:, 2]
                for c, s in enumerate(bm[:, 0]):
                    out[:, s] += contrib[:, c]
    return out


def apply_spin1(coeff: numpy.ndarray, h1e: numpy.ndarray,
                g: OracleGraph) -> numpy.ndarray:
    """One-body spin-orbital tensor (2 norb x 2 norb, alpha first) on the
    sector: only its alpha-alpha and beta-beta blocks act (reference
    ``src/fqe/fqe_data.py:533-583``)."""
    n = g.norb
    h1e = numpy.asarray(h1e)
    da, db = dvec_spin(numpy.asarray(coeff, dtype=numpy.complex128), g)
    return (numpy.einsum("ij,ijkl->kl", h1e[:n, :n], da) +
            numpy.einsum("ij,ijkl->kl", h1e[n:, n:], db))


def apply_spin12(coeff: numpy.ndarray, h1e: numpy.ndarray, h2e: numpy.ndarray,
                 g: OracleGraph) -> numpy.ndarray:
    """One- and two-body spin-orbital tensors on the sector, the dvec form
    with separate alpha and beta dvecs (reference
    ``src/fqe/fqe_data.py:712-741``): V' = -moveaxis(h2e, 1, 2),
    h1' = h1e - sum_k V'[i,k,k,j] over all 2 norb spin orbitals, then only
    the spin-conserving blocks of h1' and V' act."""
    n = g.norb
    v = -numpy.moveaxis(numpy.asarray(h2e, dtype=numpy.complex128), 1, 2)
    h1p = numpy.asarray(h1e, dtype=numpy.complex128).copy()
    h1p -= numpy.einsum("ikkj->ij", v)
    a, b = slice(0, n), slice(n, 2 * n)
    da, db = dvec_spin(numpy.asarray(coeff, dtype=numpy.complex128), g)
    out = (numpy.einsum("ij,ijkl->kl", h1p[a, a], da) +
           numpy.einsum("ij,ijkl->kl", h1p[b, b], db))
    ya = (numpy.einsum("ijkl,klmn->ijmn", v[a, a, a, a], da) +
          numpy.einsum("ijkl,klmn->ijmn", v[a, a, b, b], db))
    yb = (numpy.einsum("ijkl,klmn->ijmn", v[b, b, a, a], da) +
          numpy.einsum("ijkl,klmn->ijmn", v[b, b, b, b], db))
    out += coeff_from_dvec_spin(ya, yb, g)
    return out


def apply_spin_tensors(coeff, tensors, g: OracleGraph) -> numpy.ndarray:
    """``FqeData.apply_inplace`` for 1- or 2-tensor spin-orbital tuples on
    a spin-conserving sector (reference ``src/fqe/fqe_data.py:421-475``)."""
    if len(tensors) == 1:
        return apply_spin1(coeff, tensors[0], g)
    if len(tensors) == 2:
        return apply_spin12(coeff, tensors[0], tensors[1], g)
    raise ValueError("oracle handles 1- and 2-body spin-orbital tensors")


def restricted_as_spin_orbital(h1e: numpy.ndarray, h2e: numpy.ndarray):
    """The spin-orbital tensors of a restricted (h1, h2): h1 on both spin
    blocks and h2so[i rho, j eta, k rho, l eta] = h2[i, j, k, l], the
    convention of h2[i,j,k,l] a+_i(rho) a+_j(eta) a_k(rho) a_l(eta)
    (reference ``src/fqe/wavefunction.py:370-372``)."""
    n = h1e.shape[0]
    h1so = numpy.zeros((2 * n, 2 * n), dtype=numpy.result_type(h1e, h2e))
    h1so[:n, :n] = h1e
    h1so[n:, n:] = h1e
    h2so = numpy.zeros((2 * n,) * 4, dtype=h1so.dtype)
    for rho in range(2):
        for eta in range(2):
            r = slice(rho * n, (rho + 1) * n)
            e = slice(eta * n, (eta + 1) * n)
            h2so[r, e, r, e] = h2e
    return h1so, h2so


def apply_hamiltonian(coeff, tensors, e_0, g: OracleGraph) -> numpy.ndarray:
    """``Wavefunction._apply_array`` for one sector (reference
    ``src/fqe/wavefunction.py:366-399``): out = H C + e_0 C when
    |e_0| > 1e-15."""
    out = apply_tensors(coeff, tensors, g)
    if abs(e_0) > 1.0e-15:
        out = out + e_0 * numpy.asarray(coeff, dtype=numpy.complex128)
    return out


def taylor_evolve(coeff, tensors, e_0, time: float, g: OracleGraph,
                  accuracy: float = 1.0e-15, expansion: int = 30,
                  double_e0_phase: bool = True):
    """``time_evolve`` of a non-quadratic RestrictedHamiltonian (reference
    ``src/fqe/wavefunction.py:1042-1054`` -> ``:509-611``).

    work <- (-i t H_tensors) work ; result += work / order!  for order = 1..
    expansion-1, stopping as soon as ||work|| / order! < accuracy, else
    RuntimeError.  The tensors carry no e_0 (``iht`` builds a new tuple,
    ``src/fqe/hamiltonians/restricted_hamiltonian.py:121-134``); e_0 enters as
    the phase exp(-i t e_0), applied once by ``apply_generated_unitary``
    (``:604-605``) and once more by ``time_evolve`` (``:1051-1052``).

    Returns (evolved coefficients, number of applies performed).
    """
    iht = tuple(-1.0j * time * numpy.asarray(t) for t in tensors)
    result = numpy.asarray(coeff, dtype=numpy.complex128).copy()
    work = result.copy()
    nterms = 0
    for order in range(1, expansion):
        work = apply_tensors(work, iht, g)
        nterms += 1
        c = 1.0 / math.factorial(order)
        result += c * work
        if numpy.linalg.norm(work) * abs(c) < accuracy:
            break
    else:
        raise RuntimeError("maximum taylor expansion limit reached")
    nphase = 2 if double_e0_phase else 1
    if abs(e_0 * time) > 1.0e-15:
        result = result * numpy.exp(-1.0j * time * e_0)
    if nphase == 2 and abs(e_0) > 1.0e-15:
        result = result * numpy.exp(-1.0j * time * e_0)
    return result, nterms


def dense_matrix(h1e, h2e, g: OracleGraph) -> numpy.ndarray:
    """Dense CI matrix of the sector, column by column with ``apply_12``
    (as ``tests/evolution_test.py:554-562`` does with FqeData.apply)."""
    dim = g.lena * g.lenb
    out = numpy.zeros((dim, dim), dtype=numpy.complex128)
    for col in range(dim):
        e = numpy.zeros(dim, dtype=numpy.complex128)
        e[col] = 1.0
        out[:, col] = apply_12(e.reshape(g.lena, g.lenb), h1e, h2e,
                               g).reshape(-1)
    return out


def slater_condon_diagonal(h1e, h2e, g: OracleGraph) -> numpy.ndarray:
    """<I|H|I> for every determinant, from the closed form

    H_dd = sum_{i sigma} h1[i,i] n_{i sigma}
           - sum_{(i rho) != (j eta)} h2[i,j,i,j] n_{i rho} n_{j eta}
           + sum_{i != j, rho} h2[i,j,j,i] n_{i rho} n_{j rho}

    for H = sum h1 E + sum h2[i,j,k,l] a+_{i rho} a+_{j eta} a_{k rho}
    a_{l eta} (reference docstring ``src/fqe/wavefunction.py:370-372``).
    """
    n = g.norb
    h1 = numpy.asarray(h1e)
    h2 = numpy.asarray(h2e)

    def occ(strings):
        return numpy.array([[(int(s) >> k) & 1 for k in range(n)]
                            for s in strings], dtype=numpy.float64).reshape(-1, n)

    na = occ(g.astr)
    nb = occ(g.bstr)
    jmat = numpy.einsum("ijij->ij", h2)  # h2[i,j,i,j]
    kmat = numpy.einsum("ijji->ij", h2)  # h2[i,j,j,i]
    one_a = na @ numpy.diag(h1)
    one_b = nb @ numpy.diag(h1)
    # same-spin: -J + K over i != j ; opposite-spin: -J over all i, j
    offd = 1.0 - numpy.eye(n)
    same = (-jmat + kmat) * offd
    aa = numpy.einsum("ai,ij,aj->a", na, same, na)
    bb = numpy.einsum("bi,ij,bj->b", nb, same, nb)
    ab = -(numpy.einsum("ai,ij->aj", na, jmat) @ nb.T)
    ab = ab - (numpy.einsum("bi,ij->bj", nb, jmat) @ na.T).T
    diag = (one_a[:, None] + one_b[None, :] + aa[:, None] + bb[None, :]
            + ab)
    return diag


# ---------------------------------------------------------------------------
# synthetic workloads (documented RNG convention)
# ---------------------------------------------------------------------------

def physics_to_fqe_h2(eri_chem: numpy.ndarray) -> numpy.ndarray:
    """Chemist (ij|kl) -> FQE restricted h2.

    OpenFermion's physics tensor is V_phys[p,q,r,s] = (ps|qr); the reference's
    ``integrals_to_fqe_restricted`` (``src/fqe/openfermion_utils.py:390-402``)
    returns einsum('ijlk', -0.5 V_phys), i.e. h2[i,j,k,l] = -0.5 (ik|jl).
    """
    v_phys = numpy.einsum("psqr->pqrs", eri_chem)
    return numpy.einsum("ijlk", -0.5 * v_phys)


def synthetic_hamiltonian(norb: int, nfactors: int,
                          rng: numpy.random.Generator):
    """h1 (real symmetric, off-diagonal N(0, 0.1), diagonal -2 + N(0, 0.5))
    and h2 from ERIs (ij|kl) = sum_k L_k[i,j] L_k[k,l] with L_k real symmetric,
    entries N(0, 0.1)."""
    off = rng.normal(0.0, 0.1, size=(norb, norb))
    h1 = numpy.triu(off, 1)
    h1 = h1 + h1.T
    h1[numpy.diag_indices(norb)] = -2.0 + rng.normal(0.0, 0.5, size=norb)
    eri = numpy.zeros((norb,) * 4)
    for _ in range(nfactors):
        lk = rng.normal(0.0, 0.1, size=(norb, norb))
        lk = numpy.triu(lk) + numpy.triu(lk, 1).T
        eri += numpy.einsum("ij,kl->ijkl", lk, lk)
    return h1, physics_to_fqe_h2(eri)


def synthetic_coeff(lena: int, lenb: int,
                    rng: numpy.random.Generator) -> numpy.ndarray:
    """Complex coefficients, real and imaginary parts i.i.d. N(0, 1), then
    normalised."""
    re = rng.normal(size=(lena, lenb))
    im = rng.normal(size=(lena, lenb))
    c = re + 1.0j * im
    return c / numpy.linalg.norm(c)


def synthetic_workload(norb: int, nalpha: int, nbeta: int, seed: int,
                       nfactors: int):
    """Draw order: C real, C imag, h1 off-diagonal, h1 diagonal, factors."""
    rng = numpy.random.default_rng(seed)
    lena, lenb = binom(norb, nalpha), binom(norb, nbeta)
    coeff = synthetic_coeff(lena, lenb, rng)
    h1, h2 = synthetic_hamiltonian(norb, nfactors, rng)
    return coeff, h1, h2

def rdm12(coeff: numpy.ndarray, g: OracleGraph, bra=None):
    """1- and 2-RDMs from the dvec (reference ``src/fqe/fqe_data.py:
    1767-1782``, its non-accelerated branch): rdm1 = (sum_I conj(D2) C)^T
    and rdm2[j, k, i, l] = -sum_I conj(D2[i, j]) D[k, l] + delta_ki
    rdm1[j, l], with D = dvec(ket), D2 = dvec(bra)."""
    d = dvec(coeff, g)
    d2 = d if bra is None else dvec(bra, g)
    out1 = numpy.tensordot(d2.conj(), coeff, axes=((2, 3), (0, 1))).T
    m = numpy.tensordot(d2.conj(), d, axes=((2, 3), (2, 3)))
    out2 = -m.transpose(1, 2, 0, 3)
    for i in range(g.norb):
        out2[:, i, i, :] += out1
    return out1, out2

