"""The no-execution reader of the reference's wavefunction files
(fqe_amd.reference_format): files in the reference's ``save`` layout,
written here with stand-in classes (tests/reference_pickles.py) at pickle
protocols 2-5, read back field for field and bit for bit; and streams that
would run code under pickle.load (os.system, exec, a REDUCE in place of the
coefficients, persistent ids) rejected with ValueError and without side
effects."""

import os
import pickle

import numpy
import pytest

from fqe_amd import reference_format as rf

import reference_pickles as rp


def _coeffs(rng, norb, keys, dtype=numpy.complex128):
    import math
    out = {}
    for nele, sz in keys:
        na, nb = (nele + sz) // 2, (nele - sz) // 2
        shape = (math.comb(norb, na), math.comb(norb, nb))
        out[(nele, sz)] = (rng.normal(size=shape) +
                           1j * rng.normal(size=shape)).astype(dtype)
    return out


@pytest.mark.parametrize("protocol", [2, 3, 4, 5])
@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_reads_reference_layout(protocol, dtype):
    rng = numpy.random.default_rng(protocol)
    norb = 5
    sectors = _coeffs(rng, norb, [(4, 0), (4, 2), (3, 1)], dtype)
    raw = rp.dumps(norb, sectors, protocol, conserved={"n": 4},
                   symmetry_map={(4, 0): (2, 2)})
    got = rf.read_reference_wavefunction(raw)
    assert got["norb"] == norb
    assert got["conserve_spin"] is True and got["conserve_number"] is True
    assert got["conserved"] == {"n": 4}
    assert got["symmetry_map"] == {(4, 0): (2, 2)}
    assert set(got["sectors"]) == set(sectors)
    for key, c in sectors.items():
        assert got["sectors"][key].dtype == c.dtype
        assert numpy.array_equal(got["sectors"][key], c)


def test_fortran_ordered_and_memoised_arrays():
    rng = numpy.random.default_rng(9)
    norb = 4
    c = numpy.asfortranarray(_coeffs(rng, norb, [(4, 0)])[(4, 0)])
    raw = rp.dumps(norb, {(4, 0): c, (3, 1): c[:, :4].copy()}, 4)
    got = rf.read_reference_wavefunction(raw)
    assert numpy.array_equal(got["sectors"][(4, 0)], c)


def _evil(tmp_path, name):
    class Evil:
        def __init__(self, func, arg):
            self.func, self.arg = func, arg

        def __reduce__(self):
            return (self.func, (self.arg,))
    marker = os.path.join(str(tmp_path), name)
    return marker, Evil


def test_code_in_the_stream_is_not_run(tmp_path):
    marker, Evil = _evil(tmp_path, "ran")
    streams = [
        pickle.dumps(Evil(os.system, f"touch {marker}"), protocol=4),
        pickle.dumps(Evil(exec, f"open({marker!r}, 'w').close()"), protocol=2),
        pickle.dumps([{}, {}, True, True, 2,
                      [(2, 0), Evil(os.system, f"touch {marker}")]],
                     protocol=4),
    ]
    for raw in streams:
        with pytest.raises(ValueError):
            rf.read_reference_wavefunction(raw)
    assert not os.path.exists(marker)
    # parse_pickle itself returns inert records for the calls
    rec = rf.parse_pickle(streams[0])
    assert isinstance(rec, rf.Opaque)
    assert rec.func == rf.Global("posix", "system") or \
        rec.func == rf.Global("os", "system") or rec.func.name == "system"
    assert not os.path.exists(marker)


def test_coefficients_from_a_call_are_refused(tmp_path):
    marker, Evil = _evil(tmp_path, "ran2")
    rng = numpy.random.default_rng(2)
    sectors = _coeffs(rng, 3, [(2, 0)])
    d = rp.sector(1, 1, 3, sectors[(2, 0)])
    d.coeff = Evil(os.system, f"touch {marker}")
    with rp._reference_modules():
        raw = pickle.dumps([{}, {}, True, True, 3, [(2, 0), d]], protocol=4)
    with pytest.raises(ValueError, match="coefficient"):
        rf.read_reference_wavefunction(raw)
    assert not os.path.exists(marker)


def test_layout_errors():
    rng = numpy.random.default_rng(3)
    good = _coeffs(rng, 4, [(4, 0)])
    # wrong shape for the sector
    with pytest.raises(ValueError, match="shape"):
        rf.read_reference_wavefunction(rp.dumps(4, {(4, 2): good[(4, 0)]}, 4))
    # not a list
    with pytest.raises(ValueError):
        rf.read_reference_wavefunction(pickle.dumps({"a": 1}))
    # truncated stream
    raw = rp.dumps(4, good, 4)
    with pytest.raises(ValueError):
        rf.read_reference_wavefunction(raw[: len(raw) // 2])
    # persistent ids are refused
    import io

    class P(pickle.Pickler):
        def persistent_id(self, obj):
            return "x" if isinstance(obj, numpy.ndarray) else None
    buf = io.BytesIO()
    P(buf, protocol=4).dump([{}, {}, True, True, 4, good[(4, 0)]])
    with pytest.raises(ValueError, match="PERSID|BINPERSID"):
        rf.parse_pickle(buf.getvalue())


def test_object_and_structured_arrays_are_refused():
    arr = numpy.array([1, "a"], dtype=object)
    with pytest.raises(ValueError):
        rf.parse_pickle(pickle.dumps(arr, protocol=4))
    st = numpy.zeros(3, dtype=[("a", "<f8"), ("b", "<i4")])
    with pytest.raises(ValueError):
        rf.parse_pickle(pickle.dumps(st, protocol=4))
