"""GPU tests of the batched-sector graph (one apply + fixed 10-term Taylor
per sector, fqehip_batch_*) against the oracle, the eager sequential path
and across graph replays; every kernel variant of the autotuner against
the oracle.  Replays agree to rounding, not bit for bit: the row kernels
accumulate in LDS with atomic adds, whose order varies between runs."""

import numpy
import pytest

from oracle import fqe_oracle as fo

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-4


@pytest.fixture(scope="module")
def fqe():
    import torch
    assert torch.cuda.is_available()
    import fqe_amd
    return fqe_amd


def rel(a, b):
    return numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300)


def _oracle(spec, t, nterms):
    g = fo.OracleGraph(spec.nalpha, spec.nbeta, spec.norb)
    c = spec.coeff.astype(numpy.complex128)
    out = fo.apply_12(c, spec.h1, spec.h2, g)
    term = c.copy()
    res = c.copy()
    for k in range(1, nterms + 1):
        term = (-1j * t / k) * fo.apply_12(term, spec.h1, spec.h2, g)
        res = res + term
    return out, res


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_batched_graph_matches_oracle_and_sequential(fqe, dtype):
    import torch
    from fqe_amd import batched
    tdt = torch.complex128 if dtype == "c128" else torch.complex64
    tol = FP64_TOL if dtype == "c128" else MIXED_TOL
    same = 1e-14 if dtype == "c128" else 1e-6    # run-to-run rounding
    specs = batched.random_specs(14, seed=33, norb_range=(4, 8))
    print("shapes", [(s.norb, s.nalpha, s.nbeta) for s in specs], flush=True)
    bt = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=tdt, nstreams=4)
    bt.build()
    assert bt.nodes() > 0
    refs = [_oracle(sp, 0.1, 10) for sp in specs]
    for rep in range(3):
        bt.launch()
        got = bt.results()
        for (ap, ev), (rap, rev), sp in zip(got, refs, specs):
            assert rel(ap, rap) < tol, (rep, sp.norb, sp.nalpha, sp.nbeta)
            assert rel(ev, rev) < tol, (rep, sp.norb, sp.nalpha, sp.nbeta)
        if rep == 0:
            first = got
        else:
            for (a0, e0), (a1, e1) in zip(first, got):
                assert rel(a1, a0) < same and rel(e1, e0) < same
        # eager work between replays
        torch.zeros(1000, device="cuda").sum().item()
    bt.run_sequential()
    seq = bt.results()
    for (a0, e0), (a1, e1) in zip(first, seq):
        assert rel(a1, a0) < same and rel(e1, e0) < same
    bt.close()


def test_grouped_batch_matches_oracle_and_sequential(fqe):
    """complex64 sectors of the rows4/rows8 variants run as one group (one
    row launch and one reduction per Taylor step for all of them) beside
    split and unfused sectors on their own kernels: every sector against
    the oracle, replays against each other, and the group against the
    per-sector kernels of run_sequential."""
    import torch
    from fqe_amd import batched
    specs = batched.random_specs(24, seed=44, norb_range=(4, 9))
    cycle = ["rows8", "rows4", "rows8", "split", "unfused"]
    tiles, want = {}, 0
    for i, sp in enumerate(specs):
        key = f"{sp.norb},{sp.nalpha},{sp.nbeta}"
        if key not in tiles:
            tiles[key] = cycle[i % len(cycle)]
    for sp in specs:
        if tiles[f"{sp.norb},{sp.nalpha},{sp.nbeta}"] in ("rows4", "rows8"):
            want += 1
    print("variants", tiles, flush=True)
    bt = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=torch.complex64,
                               tiles=tiles, nstreams=4, group=True)
    bt.build()
    assert bt.grouped() == want > 0
    refs = [_oracle(sp, 0.1, 10) for sp in specs]
    for rep in range(2):
        bt.launch()
        got = bt.results()
        for (ap, ev), (rap, rev), sp in zip(got, refs, specs):
            assert rel(ap, rap) < MIXED_TOL, (rep, sp.norb, sp.nalpha, sp.nbeta)
            assert rel(ev, rev) < MIXED_TOL, (rep, sp.norb, sp.nalpha, sp.nbeta)
        # the rows4/rows8 row body accumulates beta excitations with LDS
        # atomics: replays round differently, and 10 Taylor terms carry it
        if rep == 0:
            first = got
        else:
            for (a0, e0), (a1, e1) in zip(first, got):
                assert rel(a1, a0) < 1e-5 and rel(e1, e0) < 1e-5
        torch.zeros(1000, device="cuda").sum().item()
    # run_sequential takes each sector's own variant (rows8 sectors on the
    # 8-wave kernel, the group always on the rows4 body): the beta scatter
    # rounds in another order, so the Taylor sums agree to mixed rounding
    bt.run_sequential()
    for (a0, e0), (a1, e1) in zip(first, bt.results()):
        assert rel(a1, a0) < 2e-5 and rel(e1, e0) < 2e-5
    bt.close()


@pytest.mark.parametrize("contraction", ["fp64", "ozaki"])
def test_grouped_batch_complex128(fqe, contraction):
    """complex128 sectors of the rows variant (the FP64 row kernel) run as
    one group beside split (FP64 or Ozaki-digit) and unfused sectors on
    their own kernels: every sector against the oracle to 1e-11 (Ozaki
    split sectors: 1e-9), replays against each other, and the group
    against the per-sector kernels of run_sequential."""
    import torch
    from fqe_amd import batched
    specs = batched.random_specs(24, seed=46, norb_range=(4, 9))
    cycle = ["rows4", "split", "rows4", "unfused"]
    tiles, want = {}, 0
    for i, sp in enumerate(specs):
        key = f"{sp.norb},{sp.nalpha},{sp.nbeta}"
        if key not in tiles:
            tiles[key] = cycle[i % len(cycle)]
    print("variants", tiles, flush=True)
    bt = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=torch.complex128,
                               tiles=tiles, nstreams=4, group=True,
                               contraction=contraction)
    bt.build()
    want = sum(1 for p in bt.sectors if p.variant == "rows4")
    assert bt.grouped() == want > 0
    refs = [_oracle(sp, 0.1, 10) for sp in specs]
    tol = 1e-9 if contraction == "ozaki" else 1e-11
    for rep in range(2):
        bt.launch()
        got = bt.results()
        for (ap, ev), (rap, rev), sp in zip(got, refs, specs):
            assert rel(ap, rap) < tol, (rep, sp.norb, sp.nalpha, sp.nbeta)
            assert rel(ev, rev) < tol, (rep, sp.norb, sp.nalpha, sp.nbeta)
        if rep == 0:
            first = got
        else:
            for (a0, e0), (a1, e1) in zip(first, got):
                assert rel(a1, a0) < 1e-13 and rel(e1, e0) < 1e-13
        torch.zeros(1000, device="cuda").sum().item()
    bt.run_sequential()
    for (a0, e0), (a1, e1) in zip(first, bt.results()):
        assert rel(a1, a0) < 1e-13 and rel(e1, e0) < 1e-13
    bt.close()


def test_grouped_batch_without_apply(fqe):
    """Items with d_apply_out NULL (Taylor series only) inside a group: the
    grouped step-0 launches skip them, their Taylor sums still match the
    oracle, and the items that do ask for H C0 get it."""
    import ctypes
    import torch
    from fqe_amd import _lib, batched
    specs = batched.random_specs(6, seed=55, norb_range=(4, 7))
    tiles = {f"{sp.norb},{sp.nalpha},{sp.nbeta}": "rows4" for sp in specs}
    bt = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=torch.complex64,
                               tiles=tiles, nstreams=2, group=True)
    items = []
    for i, p in enumerate(bt.sectors):
        it = p.item()
        if i % 2:
            it.d_apply_out = None
            p.apply_out.fill_(7.0)
        items.append(it)
    arr = (_lib.BatchItem * len(items))(*items)
    handle = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.call("fqehip_batch_create", bt.prec, len(items), arr, bt.time,
              bt.nterms, bt.nstreams, 1, _lib.stream_ptr(),
              ctypes.byref(handle))
    bt._batch = handle
    assert bt.grouped() == len(specs)
    bt.launch()
    got = bt.results()
    for i, ((ap, ev), sp) in enumerate(zip(got, specs)):
        rap, rev = _oracle(sp, 0.1, 10)
        assert rel(ev, rev) < MIXED_TOL, i
        if i % 2:
            assert numpy.all(ap == 7.0), i
        else:
            assert rel(ap, rap) < MIXED_TOL, i
    bt.close()


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_every_variant_matches_oracle(fqe, dtype):
    """Each kernel variant of each precision (complex128: rows4, split,
    unfused) in a one-sector graph against the oracle."""
    import torch
    from fqe_amd import batched
    tdt = getattr(torch, dtype)
    tol = MIXED_TOL if dtype == "complex64" else 1e-11
    specs = batched.random_specs(3, seed=5, norb_range=(7, 8))
    for sp in specs:
        ref, _ = _oracle(sp, 0.1, 0)
        base = batched._Prepared(sp, tdt, None)
        variants = batched.variants_for(base.op, base.prec)
        if dtype == "complex128" and base.op.pair_symmetric:
            assert "split" in variants
        for var in variants:
            bt = batched.BatchedTaylor([sp], 0.1, nterms=0, dtype=tdt,
                                       tiles={f"{sp.norb},{sp.nalpha},{sp.nbeta}": var},
                                       nstreams=1)
            assert bt.sectors[0].variant == var
            assert bool(bt.sectors[0].s.hbb_f64) == (
                var == "split" and dtype == "complex128")
            bt.build()
            bt.launch()
            ap, ev = bt.results()[0]
            assert rel(ap, ref) < tol, var
            assert rel(ev, sp.coeff) < 1e-7, var
            bt.close()


def test_autotune_picks_a_valid_variant(fqe):
    import torch
    from fqe_amd import batched
    specs = batched.random_specs(6, seed=8, norb_range=(5, 9))
    table = batched.autotune(specs, torch.complex64, reps=2)
    for sp in specs:
        key = f"{sp.norb},{sp.nalpha},{sp.nbeta}"
        assert table[key] in batched.VARIANTS
