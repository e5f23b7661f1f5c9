"""Benchmark of the RestrictedHamiltonian apply on one MI355X.

Workload (the headline configuration): norb=14, 7 alpha + 7 beta electrons
(3,432 x 3,432 = 11,778,624 determinants), complex coefficients with real
and imaginary parts i.i.d. N(0,1) from seed 2, normalised; h1/h2 from 14
Cholesky-like factors (fqe_amd.synthetic.workload).  A step is
one apply H|C> with C resident in HBM.

The engine runs it three ways:
  * mixed precision at the metric's 1e-9 gate (``value``): complex128
    storage, both contractions as exact products of 8-bit digits on the
    int8 matrix cores (Ozaki scheme, int32 accumulation, FP64 assembly);
  * FP64 (complex128 storage, FP64 matrix cores) beside it (``fp64``);
  * the BF16x3 engine (complex64 storage, BF16 hi/lo products with FP32
    accumulation), labelled separately (``bf16x3_c64``): faster, but about
    4e-6 from the FP64 result, outside the gate.
Each result states its relative error against the FP64 GPU result and the
reference's own complex128 apply (its compiled C lm kernels, oracle/_ref).

Per-kernel durations are measured with HIP events on the stream the
kernels run on; ``roofline`` is given for the dominant kernel of the
headline engine.  ``cpu_baseline`` times the reference's compiled C on the
host cores: one warm-up, then the median of five applies, all cores,
OMP_PROC_BIND=close, plus a single-thread record on the norb=12 sector.
``cfg1`` is the production-size plain-method workload (norb=12 6a6b, seed
1): 100 chained applies plus time_evolve(0.05), FP64 on the GPU against
the reference's C on the host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""

import os

# host cores available to this process, counted before any OpenMP runtime
# binds the main thread
HOST_CORES = len(os.sched_getaffinity(0))
# the CPU baseline's OpenMP threads stay on neighbouring cores; set before
# any OpenMP runtime is loaded
os.environ.setdefault("OMP_PROC_BIND", "close")

import argparse
import json
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "openfermion-fqe_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("applies/sec (norb=14,7α7β): mixed vs FP64 on GPU at "
          "≤1e-9 rel err")
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "bf16": 2500.0, "i8": 5000.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--norb", type=int, default=14)
    ap.add_argument("--nalpha", type=int, default=7)
    ap.add_argument("--nbeta", type=int, default=7)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default="fp64,ozaki,bf16x3",
                    help="engines to run (profiling runs use one): fp64, "
                         "ozaki (the headline), bf16x3")
    ap.add_argument("--no-cfg1", action="store_true",
                    help="skip the norb=12 chained-apply workload")
    ap.add_argument("--no-evolve", action="store_true",
                    help="skip the timed time_evolve lines")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="first copy this many bytes device-to-device (the "
                         "calibration kernel of HBM traffic counters)")
    ap.add_argument("--split-tile", type=int, default=None,
                    help="kernel of the complex64 split row pass (0: live "
                         "pair rows, 16: all pair rows; default: the engine "
                         "setting)")
    ap.add_argument("--split-tile-fp64", type=int, default=None,
                    help="kernel of the complex128 split row pass (0: "
                         "column-owned waves, 32: shared D tile)")
    ap.add_argument("--no-split-fp64", action="store_true",
                    help="FP64 engine without the split apply (the FP64 row "
                         "kernel also scatters beta excitations)")
    ap.add_argument("--unfused", action="store_true",
                    help="time the unfused gather / GEMM / scatter kernels")
    ap.add_argument("--complex-op", action="store_true",
                    help="give h1 and h2 imaginary parts (complex operator: "
                         "full pair basis, real 2 x 2 block GEMM)")
    ap.add_argument("--batched", action="store_true",
                    help="run the batched-sector workload instead (256 "
                         "sectors, norb 6..12, seed 3, one apply + 10-term "
                         "Taylor each) and print its own JSON line")
    ap.add_argument("--sectors", type=int, default=256)
    ap.add_argument("--no-group", action="store_true",
                    help="batched: every sector on its own kernels (no "
                         "grouped launch of the rows4/rows8 sectors)")
    ap.add_argument("--davidson", action="store_true",
                    help="run the Davidson workload instead: train the "
                         "learned preconditioner on small synthetic "
                         "Hamiltonians, then solve held-out norb=14 7a7b "
                         "ones (seed 4) to a 1e-8 residual with and without "
                         "it; prints its own JSON line")
    ap.add_argument("--set", action="store_true",
                    help="run the spin-orbital FqeDataSet workload instead "
                         "(every sector of nele electrons on norb orbitals, "
                         "random Hermitian spin-mixing h1/h2, seed 5, "
                         "complex128 1+2-body apply) and print its own JSON "
                         "line")
    ap.add_argument("--set-norb", type=int, default=10,
                    help="set workload: norb (nele = norb)")
    ap.add_argument("--train", type=int, default=120,
                    help="training Hamiltonians (norb 8 and 9)")
    ap.add_argument("--heldout", type=int, default=3,
                    help="held-out norb=14 Hamiltonians")
    return ap.parse_args()


def davidson_main(args):
    """Davidson to a 1e-8 residual (FP64 applies, FP64 Gram-Schmidt and
    Rayleigh-Ritz) with the plain diagonal preconditioner and with the
    learned correction, on held-out norb=14 7a7b Hamiltonians."""
    import numpy
    import torch
    import fqe_amd as fqe
    from fqe_amd import synthetic
    from fqe_amd.algorithm import mlp_precond as mp
    t0 = time.perf_counter()
    xs, ts, ws, its = [], [], [], []
    rng = numpy.random.default_rng(40)
    for j in range(args.train):
        norb = 8 if j % 2 == 0 else 9
        na, nb = norb // 2, norb - norb // 2
        h1, h2 = synthetic.hamiltonian(norb, norb, rng)
        ham = fqe.get_restricted_hamiltonian((h1, h2))
        out = mp.correction_samples(ham, na, nb, per_snapshot=256, seed=j)
        if out is not None:
            xs.append(out[0]); ts.append(out[1]); ws.append(out[2])
            its.append(out[3])
    t_data = time.perf_counter() - t0
    model = mp.MLPPreconditioner(seed=0)
    x, t, w = torch.cat(xs), torch.cat(ts), torch.cat(ws)
    t1 = time.perf_counter()
    losses = model.train(x, t, w, epochs=20, batch=4096, lr=1e-3)
    t_train = time.perf_counter() - t1
    rng = numpy.random.default_rng(4)
    plain, learned, tp, tl, gap = [], [], 0.0, 0.0, 0.0
    mixed, tm, gap_m, switches = [], 0.0, 0.0, []
    ozaki_its, tz, gap_z = [], 0.0, 0.0
    for j in range(args.heldout):
        h1, h2 = synthetic.hamiltonian(14, 14, rng)
        ham = fqe.get_restricted_hamiltonian((h1, h2))
        r = mp.solve_both(ham, 7, 7, model, tol=1e-8, max_iter=600)
        plain.append(r["plain"].iterations)
        learned.append(r["learned"].iterations)
        tp += r["t_plain"]
        tl += r["t_learned"]
        gap = max(gap, abs(r["plain"].eigenvalues[0] -
                           r["learned"].eigenvalues[0]))
        rm = mp.solve_mixed(ham, 7, 7, tol=1e-8, max_iter=600)
        mixed.append(rm.iterations)
        tm += rm.seconds
        gap_m = max(gap_m, abs(rm.eigenvalues[0] -
                               r["plain"].eigenvalues[0]))
        print(f"held-out {j}: plain {plain[-1]} its {r['t_plain']:.1f} s, "
              f"learned {learned[-1]} its {r['t_learned']:.1f} s, "
              f"mixed-then-FP64 {mixed[-1]} its {rm.seconds:.1f} s "
              f"(switch at {rm.switch_iteration}, converged "
              f"{rm.converged})", file=sys.stderr, flush=True)
        switches.append(rm.switch_iteration)
        # every apply on the Ozaki-digit engine (complex128, <= 1e-9 of the
        # FP64 apply, far below the 1e-8 residual)
        from fqe_amd import settings
        from fqe_amd.algorithm import davidson as dvm
        saved = settings.contraction
        settings.contraction = "ozaki"
        try:
            t3 = time.perf_counter()
            ro = dvm.davidson(ham, 7, 7, tol=1e-8, max_iter=600,
                              dtype=torch.complex128)
            to_ = time.perf_counter() - t3
        finally:
            settings.contraction = saved
        ozaki_its.append(ro.iterations)
        tz += to_
        gap_z = max(gap_z, abs(ro.eigenvalues[0] - r["plain"].eigenvalues[0]))
        print(f"held-out {j}: Ozaki applies {ro.iterations} its {to_:.1f} s "
              f"(converged {ro.converged})", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "Davidson iterations to a 1e-8 residual, norb=14 7a7b "
                  "(11,778,624 determinants), learned vs plain diagonal "
                  "preconditioner",
        "plain_iterations": plain, "learned_iterations": learned,
        "mean_plain": float(numpy.mean(plain)),
        "mean_learned": float(numpy.mean(learned)),
        "iteration_reduction": float(1.0 - numpy.mean(learned) /
                                     numpy.mean(plain)),
        "seconds_plain": round(tp, 2), "seconds_learned": round(tl, 2),
        "max_eigenvalue_gap": gap,
        "mixed_then_fp64_iterations": mixed,
        "mixed_phase_iterations": switches,
        "seconds_mixed_then_fp64": round(tm, 2),
        "speedup_mixed_then_fp64_vs_fp64": round(tp / max(tm, 1e-9), 3),
        "max_eigenvalue_gap_mixed_vs_fp64": gap_m,
        "ozaki_iterations": ozaki_its,
        "seconds_ozaki": round(tz, 2),
        "speedup_ozaki_vs_fp64": round(tp / max(tz, 1e-9), 3),
        "max_eigenvalue_gap_ozaki_vs_fp64": gap_z,
        "train": {"hamiltonians": len(xs), "samples": int(x.shape[0]),
                  "mean_plain_iterations": float(numpy.mean(its)),
                  "loss_first_last": [losses[0], losses[-1]],
                  "data_s": round(t_data, 1), "train_s": round(t_train, 1)},
        "applies": "FP64 engine", "n_gpus": 1, "data": "synthetic",
    }))


def _time_graph(graph, steps, warmup):
    import torch
    for _ in range(warmup):
        graph.launch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        graph.launch()
    torch.cuda.synchronize()
    return (time.perf_counter() - t1) / steps


def batched_main(args):
    """Batched uneven sectors, one graph launch per batch: the autotuned
    complex128 graph with Ozaki-digit split contractions (the mixed engine
    at <= 1e-9; value), the autotuned complex128 FP64 graph, the complex64
    BF16x3 graph (outside the gate), and the per-sector sequential FP64
    path, on the same sectors."""
    import numpy
    import torch
    from fqe_amd import batched
    specs = batched.random_specs(args.sectors, seed=3, norb_range=(6, 12))
    fp64 = batched.BatchedTaylor(specs, 0.1, nterms=10,
                                 dtype=torch.complex128, contraction="fp64")
    fp64.run_sequential()
    torch.cuda.synchronize()
    reps = max(1, min(args.steps, 5))
    t2 = time.perf_counter()
    for _ in range(reps):
        fp64.run_sequential()
    torch.cuda.synchronize()
    wall_seq = (time.perf_counter() - t2) / reps
    res_seq = fp64.results()
    fp64.close()
    del fp64

    def errs(res):
        ea = max(float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))
                 for (a, _), (b, _) in zip(res, res_seq))
        ee = max(float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))
                 for (_, a), (_, b) in zip(res, res_seq))
        return ea, ee

    runs = {}
    for name, dtype, contraction in (("ozaki", torch.complex128, "ozaki"),
                                     ("fp64", torch.complex128, "fp64"),
                                     ("bf16x3", torch.complex64, None)):
        t0 = time.perf_counter()
        tiles = batched.autotune(specs, dtype, reps=2,
                                 contraction=contraction)
        t_tune = time.perf_counter() - t0
        g = batched.BatchedTaylor(specs, 0.1, nterms=10, dtype=dtype,
                                  tiles=tiles, group=not args.no_group,
                                  contraction=contraction)
        t_build = g.build()
        wall = _time_graph(g, args.steps, args.warmup)
        counts = {}
        for v in tiles.values():
            counts[v] = counts.get(v, 0) + 1
        runs[name] = {"wall": wall, "errs": errs(g.results()),
                      "variants": counts, "grouped": g.grouped(),
                      "tune": t_tune, "build": t_build,
                      "ndet": g.determinants()}
        print(f"{name} graph: {1e3 * wall:.2f} ms per batch, errors vs FP64 "
              f"sequential {runs[name]['errs']}", file=sys.stderr, flush=True)
        g.close()
        del g
        torch.cuda.empty_cache()
    applies = 11 * len(specs)
    oz, fp, bf = runs["ozaki"], runs["fp64"], runs["bf16x3"]
    print(json.dumps({
        "metric": "batched sector applies/s (256 uneven sectors, 1 apply + "
                  "10-term Taylor each, one graph launch): mixed at <= 1e-9 "
                  "vs FP64",
        "value": round(applies / oz["wall"], 1), "unit": "sector-applies/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * oz["wall"], 3), "higher_is_better": True,
        "dtype": ("c128 storage; split sectors with Ozaki int8 digit "
                  "products, other sectors FP64"),
        "data": "synthetic",
        "config": {"workload": f"{len(specs)} sectors, norb 6..12, nalpha and "
                               f"nbeta uniform in [1, norb-1], seed 3, "
                               f"{oz['ndet']} determinants in total",
                   "variants": oz["variants"]},
        "max_rel_err_apply_vs_fp64": oz["errs"][0],
        "max_rel_err_taylor_vs_fp64": oz["errs"][1],
        "gate_1e-9_met": max(oz["errs"]) <= 1e-9,
        "fp64_graph": {"value": round(applies / fp["wall"], 1),
                       "unit": "sector-applies/s",
                       "ms_per_batch": round(1e3 * fp["wall"], 3),
                       "variants": fp["variants"],
                       "max_rel_err_vs_sequential": max(fp["errs"])},
        "speedup_mixed_vs_fp64_graph": round(fp["wall"] / oz["wall"], 3),
        "fp64_sequential": {"value": round(applies / wall_seq, 1),
                            "unit": "sector-applies/s",
                            "ms_per_batch": round(1e3 * wall_seq, 3)},
        "speedup_mixed_vs_fp64_sequential": round(wall_seq / oz["wall"], 3),
        "bf16x3_c64_graph": {
            "note": "complex64 storage, BF16 hi/lo products: outside the gate",
            "value": round(applies / bf["wall"], 1),
            "unit": "sector-applies/s",
            "ms_per_batch": round(1e3 * bf["wall"], 3),
            "variants": bf["variants"], "grouped_sectors": bf["grouped"],
            "max_rel_err_apply_vs_fp64": bf["errs"][0],
            "max_rel_err_taylor_vs_fp64": bf["errs"][1]},
        "autotune_s": {k: round(v["tune"], 2) for k, v in runs.items()},
        "graph_build_s": {k: round(v["build"], 3) for k, v in runs.items()},
    }))


class EngineRun:
    """Device buffers and the kernel launches of one apply (fused path when
    the engine supports the sector, else gather / GEMM / scatter)."""

    def __init__(self, graph, op, coeff_np, torch_dtype, fused=True,
                 contraction="fp64"):
        import ctypes
        import torch
        from fqe_amd import _lib, engine, settings
        self.torch = torch
        self._lib = _lib
        self.graph = graph
        self.op = op
        dev = _lib.device()
        self.c = torch.from_numpy(coeff_np).to(dev).to(torch_dtype)
        self.out = torch.empty_like(self.c)
        saved = settings.fused, settings.contraction
        settings.fused = fused
        settings.contraction = contraction
        self.prec = engine.precision_of(torch_dtype)
        self.s, self.keep = engine.sector_struct(graph, op, self.prec)
        settings.fused, settings.contraction = saved
        self.opc = op.c_struct()
        self.sref = ctypes.byref(self.s)
        self.oref = ctypes.byref(self.opc)
        self.fused = bool(_lib.lib().fqehip_fused_supported(
            self.prec, self.sref, self.oref))
        self.ws = engine.workspace(self.s, self.prec, dev)
        self.n = self.c.numel()
        esz = 8 if self.prec == _lib.C128 else 4
        self.y_ptr = self.ws.data_ptr() + op.npair * 2 * self.n * esz
        self.split = bool(self.s.hbb_hi) or bool(self.s.hbb_f64)
        self.rowacc_ptr = self.ws.data_ptr() + self.s.nslot * 2 * self.n * esz
        self.names = (["fused_rows", "slot_reduce"] if self.fused else
                      ["dvec_gather", "pair_gemm", "scatter"])

    def launch(self, events=None):
        L, s, op = self._lib, self.s, self.op
        st = L.stream_ptr()
        if events:
            events[0].record()
        if self.fused:
            L.call("fqehip_apply_fused_rows", self.prec, self.sref, self.oref,
                   L.ptr(self.c), L.ptr(self.ws), None, st)
            if events:
                events[1].record()
            L.call("fqehip_apply_slot_reduce", self.prec, self.sref,
                   L.ptr(self.ws), 1.0, 0.0, 0.0, 0.0, None, L.ptr(self.out),
                   None, None, None, st)
            if events:
                events[2].record()
            return
        L.call("fqehip_dvec_gather", self.prec, op.npair, s.lena, s.lenb,
               s.ptab_a, s.ptab_b, L.ptr(self.c), L.ptr(self.ws), st)
        if events:
            events[1].record()
        if op.complex_op:
            L.call("fqehip_pair_gemm", self.prec, 2 * (op.npair + 1),
                   2 * op.npair, L.ptr(op.a), self.n, L.ptr(self.ws),
                   self.y_ptr, st)
        else:
            L.call("fqehip_pair_gemm", self.prec, op.npair + 1, op.npair,
                   L.ptr(op.a), 2 * self.n, L.ptr(self.ws), self.y_ptr, st)
        if events:
            events[2].record()
        L.call("fqehip_scatter", self.prec, op.npair, s.lena, s.lenb,
               s.ptab_a, s.ptab_b, self.y_ptr, 1.0, 0.0, 0.0, 0.0, None,
               L.ptr(self.out), None, None, None, st)
        if events:
            events[3].record()

    def run(self, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            self.launch()
        torch.cuda.synchronize()
        nev = len(self.names) + 1
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(nev)]
               for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            self.launch(evs[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = {nm: sum(e[i].elapsed_time(e[i + 1]) for e in evs) / steps
               for i, nm in enumerate(self.names)}
        if self.split:
            # fused_rows = beta-beta GEMM + alpha row kernel; the GEMM is
            # timed alone (same stream, same buffers) to split the two
            L, s = self._lib, self.s
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            work = self.rowacc_ptr + 2 * self.n * 8   # row maxima, digits
            for _ in range(steps):
                if self.prec == L.C128 and s.hbb_oz:
                    L.call("fqehip_beta_gemm_digits", s.lena, s.lenb,
                           s.hbb_oz, L.ptr(self.c), work, self.rowacc_ptr,
                           None, L.stream_ptr())
                elif self.prec == L.C128:
                    L.call("fqehip_beta_gemm_f64", s.lena, s.lenb, s.hbb_f64,
                           L.ptr(self.c), self.rowacc_ptr, None,
                           L.stream_ptr())
                else:
                    L.call("fqehip_beta_gemm", s.lena, s.lenb, s.hbb_hi,
                           s.hbb_lo, s.ldh, L.ptr(self.c), self.rowacc_ptr,
                           None, L.stream_ptr())
            ev[1].record()
            torch.cuda.synchronize()
            tb = ev[0].elapsed_time(ev[1]) / steps
            rows = per.pop("fused_rows")
            per = {"beta_gemm": tb, "alpha_rows": rows - tb, **per}
        return wall, per

    def result(self):
        return self.out.to(self.torch.complex128).cpu().numpy()

    def free(self):
        del self.c, self.out, self.ws
        self.torch.cuda.empty_cache()


def kernel_work(op, graph, n, esz, mixed, ozaki=False):
    """Algorithmic bytes and flops of one launch of each kernel.  GEMM
    flops are those of the real (npair+1) x npair pair matrix on 2n real
    columns; the mixed engine executes three BF16 products per flop.
    Split apply: the alpha row kernel multiplies the rows of the pair
    matrix that are scattered, nalpha * (norb - nalpha) excitations to
    other strings plus the folded diagonal row (all npair + 1 rows with the
    all-rows kernel), by npair + 1 columns (extra h1' column); the
    beta-beta GEMM is lena x lenb x lenb complex times real (bfloat16
    hi/lo copies of Hbb for the mixed engine, FP64 columns for FP64).
    The Ozaki-digit engine moves the FP64 engine's bytes except for the
    slot rows, which it writes and the reduction reads in 48 bits per
    double (12 bytes per complex value).  Its algorithmic int8 operations
    ("ops") are the scheme's 20 digit products per product of the FP64
    contraction: over the nslot + 1 live pair rows and npair + 1 gathered
    rows per column (alpha rows) and 2 lena x lenb x lenb (beta-beta GEMM);
    "executed_ops" are those of the padded tiles the kernels run (64 rows x
    128 gathered rows; K = lenb rounded up to 64)."""
    from fqe_amd import settings
    p = op.npair
    nslot = graph.nalpha() * (graph.norb() - graph.nalpha())
    # a complex operator runs on its real 2 x 2 block form: m = 2(p + 1),
    # k = 2p on n columns, twice the flops of the real (p + 1) x p on 2n
    flops = 2.0 * (p + 1) * p * 2 * n * (2 if op.complex_op else 1)
    rows = p + 1 if (mixed and settings.split_tile) else nslot + 1
    lenb = graph.lenb()
    ldh = ((lenb + 127) // 128) * 128
    ldk = ((lenb + 63) // 64) * 64
    if ozaki:
        lena = graph.lena()
        return {
            "beta_gemm": {"bytes": 2 * 2 * n * esz + 5 * lenb * ldk,
                          "ops": 20 * 2.0 * (2 * lena) * lenb * lenb,
                          "executed_ops": 20 * 2.0 * (2 * lena) * ldk * lenb},
            "alpha_rows": {"bytes": nslot * 12 * n + 3 * 2 * n * esz,
                           "ops": 20 * 2.0 * (nslot + 1) * (p + 1) * 2 * n,
                           "executed_ops": 20 * 2.0 * 64 * 128 * 2 * n},
            "slot_reduce": {"bytes": nslot * 12 * n + 2 * 2 * n * esz},
        }
    return {
        "beta_gemm": {"bytes": 2 * 2 * n * esz + (2 * ldh * ldh * 2 if mixed
                                                  else lenb * lenb * 8),
                      "flops": 2.0 * 2 * n * lenb},
        "alpha_rows": {"bytes": (nslot + 3) * 2 * n * esz,
                       "flops": 2.0 * rows * (p + 1) * 2 * n},
        "dvec_gather": {"bytes": (p + 1) * 2 * n * esz},
        "pair_gemm": {"bytes": (2 * p + 1) * 2 * n * esz, "flops": flops},
        "scatter": {"bytes": (p + 1 + 1) * 2 * n * esz},
        "fused_rows": {"bytes": (nslot + 2) * 2 * n * esz, "flops": flops},
        "slot_reduce": {"bytes": (nslot + 2) * 2 * n * esz},
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "apply_norb14_pmc_traffic.md")
SET_TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                                "set_apply_norb10_pmc_traffic.md")


def _traffic_block(path):
    if not os.path.exists(path):
        return {}
    text = open(path).read()
    i = text.find("```json")
    if i < 0:
        return {}
    j = text.find("```", i + 7)
    return json.loads(text[i + 7:j])["traffic_bytes_per_launch"]


def measured_traffic(norb, na, nb):
    """HBM bytes per launch of each kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (norb=14 7a7b only), else {}."""
    if (norb, na, nb) != (14, 7, 7) or not os.path.exists(TRAFFIC_FILE):
        return {}
    text = open(TRAFFIC_FILE).read()
    i = text.find("```json")
    if i < 0:
        return {}
    j = text.find("```", i + 7)
    return json.loads(text[i + 7:j])["traffic_bytes_per_launch"]


def roofline_entry(name, per_ms, work, mode, traffic=None):
    """Roofline of one kernel: the resource whose time at peak is longer
    bounds it (HBM bytes / 8 TB/s vs MFMA flops / peak).  ``traffic`` is
    the measured HBM bytes per launch (rocprofv3 counters), if known."""
    t = per_ms[name] * 1e-3
    w = work[name]
    if mode == "ozaki":
        # the scheme's int8 digit products (algorithmic operations, no tile
        # padding) against the int8 MFMA peak, or its algorithmic bytes
        # against HBM
        t_hbm = w["bytes"] / (HBM_PEAK_GBS * 1e9)
        t_mfma = w.get("ops", 0.0) / (PEAK_TFLOPS["i8"] * 1e12)
        if t_mfma > t_hbm:
            ach = w["ops"] / t / 1e12
            return {"kernel": name, "bound": "mfma", "achieved": round(ach, 1),
                    "peak": PEAK_TFLOPS["i8"], "unit": "TOP/s (int8)",
                    "frac": round(ach / PEAK_TFLOPS["i8"], 4),
                    "traffic": traffic, "ops_per_launch": w["ops"],
                    "executed_ops_per_launch": w.get("executed_ops"),
                    "algorithmic_bytes_per_launch": w["bytes"]}
        ach = w["bytes"] / t / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": w["bytes"],
                "int8_ops_per_launch": w.get("ops")}
    peak_fl = PEAK_TFLOPS["f64"] if mode == "fp64" else PEAK_TFLOPS["bf16"]
    mult = 1.0 if mode == "fp64" else 3.0
    t_hbm = w["bytes"] / (HBM_PEAK_GBS * 1e9)
    t_mfma = mult * w.get("flops", 0.0) / (peak_fl * 1e12)
    if t_mfma > t_hbm:
        ach = mult * w["flops"] / t / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": round(ach, 2),
                "peak": peak_fl, "unit": "TFLOP/s",
                "frac": round(ach / peak_fl, 4), "traffic": traffic,
                "flops_per_launch": mult * w["flops"],
                "algorithmic_bytes_per_launch": w["bytes"]}
    ach = w["bytes"] / t / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": w["bytes"]}


def api_rates(coeff, h1, h2, norb, na, nb, reps=10):
    """Applies/s of the drop-in Python call ``Wavefunction.apply`` with the
    coefficients resident in HBM (a new Wavefunction per call, as in the
    reference; the operator's device form cached on the Hamiltonian),
    Ozaki-digit and FP64 contractions."""
    import numpy
    import torch
    import fqe_amd as fqe
    from fqe_amd import settings
    key = (na + nb, na - nb)
    saved = settings.contraction
    out = {}
    try:
        ham = fqe.get_restricted_hamiltonian((h1, h2))
        wfn = fqe.Wavefunction([[na + nb, na - nb, norb]],
                               dtype=numpy.complex128)
        wfn.set_wfn(strategy="from_data",
                    raw_data={key: torch.from_numpy(coeff)})
        for mode in ("ozaki", "fp64"):
            settings.contraction = mode
            res = wfn.apply(ham)                  # warm-up
            del res
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                res = wfn.apply(ham)
                del res
            torch.cuda.synchronize()
            out[mode] = round(reps / (time.perf_counter() - t0), 2)
    finally:
        settings.contraction = saved
    del wfn, ham
    torch.cuda.empty_cache()
    return out


def host_boundary_rate(coeff, h1, h2, norb, na, nb, reps=3):
    """Headline-engine applies/s when the caller hands over and takes back
    host arrays (Wavefunction.set_wfn from numpy, apply, get_coeff): the
    PCIe-inclusive rate of the Python boundary."""
    import numpy
    import torch
    import fqe_amd as fqe
    from fqe_amd import settings
    key = (na + nb, na - nb)
    saved = settings.contraction
    settings.contraction = "ozaki"
    try:
        ham = fqe.get_restricted_hamiltonian((h1, h2))
        wfn = fqe.Wavefunction([[na + nb, na - nb, norb]],
                               dtype=numpy.complex128)
        times = []
        for r in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            wfn.set_wfn(strategy="from_data", raw_data={key: coeff})
            wfn.apply(ham).get_coeff(key)
            torch.cuda.synchronize()
            if r:
                times.append(time.perf_counter() - t0)
    finally:
        settings.contraction = saved
    del wfn, ham
    torch.cuda.empty_cache()
    return reps / sum(times)


def set_main(args):
    """FqeDataSet 1+2-body apply (complex128) through fqehip_set_apply, with
    the three kernels timed on their own; the oracle's apply of the same
    workload is the CPU baseline and the parity check."""
    import math
    import numpy
    import torch
    from fqe_amd import _lib as L
    norb = args.set_norb
    nele = norb
    m = 2 * norb
    m2 = m * m
    n = int(L.call("fqehip_set_length", norb, nele))
    assert n == math.comb(2 * norb, nele)
    rng = numpy.random.default_rng(5)
    h1 = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    h2 = rng.standard_normal((m,) * 4) + 1j * rng.standard_normal((m,) * 4)
    h1 = h1 + h1.conj().T
    h2 = h2 + h2.conj().transpose(3, 2, 1, 0)
    coeff = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    coeff /= numpy.linalg.norm(coeff)
    v = -numpy.moveaxis(h2, 1, 2)
    h1p = h1 - numpy.einsum("ikkj->ij", v)
    dev = L.device()
    if args.copy_probe:
        # calibration copy for the FETCH_SIZE / WRITE_SIZE counter passes
        src = torch.empty(args.copy_probe, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    tc = torch.from_numpy(coeff).to(dev)
    th1 = torch.from_numpy(numpy.ascontiguousarray(h1p)).to(dev)
    tv = torch.from_numpy(numpy.ascontiguousarray(v.reshape(m2, m2))).to(dev)
    out = torch.empty_like(tc)
    ws = torch.empty(int(L.call("fqehip_set_workspace_bytes", norb, nele)),
                     dtype=torch.uint8, device=dev)
    d_ptr = ws.data_ptr()
    st = L.stream_ptr()

    def apply():
        L.call("fqehip_set_apply", norb, nele, L.ptr(th1), L.ptr(tv),
               L.ptr(tc), L.ptr(out), d_ptr, st)

    def timed(fn, steps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(steps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / steps

    for _ in range(args.warmup):
        apply()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        apply()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_apply = timed(apply, args.steps)
    t_dvec = timed(lambda: L.call("fqehip_set_dvec", norb, nele, L.ptr(tc),
                                  d_ptr, st), args.steps)
    t_vd = timed(lambda: L.call("fqehip_set_contract", norb, nele, L.ptr(tv),
                                d_ptr, st), args.steps)
    apply()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    flops_vd = 8.0 * m2 * m2 * n
    achieved = flops_vd / (t_vd * 1e-3) / 1e12
    line = {
        "metric": "FqeDataSet 1+2-body applies/s (spin-orbital, complex128)",
        "value": round(args.steps / wall, 3), "unit": "applies/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic",
        "config": {"workload": (f"FqeDataSet nele={nele} norb={norb}: every "
                                f"sector, {n} determinants, random Hermitian "
                                f"spin-mixing h1/h2, seed 5")},
        "kernels_ms": {"k_set_dvec": round(t_dvec, 4),
                       "contract (k_set_expand_v + k_set_gemm)": round(t_vd, 4),
                       "k_set_out": round(t_apply - t_dvec - t_vd, 4),
                       "apply": round(t_apply, 4)},
        "roofline": {"kernel": "k_set_gemm", "bound": "mfma",
                     "achieved": round(achieved, 3), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(achieved / 78.6, 4),
                     "traffic": (_traffic_block(SET_TRAFFIC_FILE).get(
                         "k_set_gemm") if norb == 10 else None),
                     "note": ("8 M^4 L algorithmic FP64 flops per "
                              "contraction (real block GEMM, m = k = 2 M^2, "
                              "L columns); time includes the V expansion")},
    }
    if not args.no_cpu_baseline:
        from oracle import fqe_set_oracle as so
        cores = HOST_CORES
        sg = so.SetGraph(nele, norb)
        coeffs, off = {}, 0
        for k in sg.keys:
            sz = sg.shape(k)[0] * sg.shape(k)[1]
            coeffs[k] = coeff[off:off + sz].reshape(sg.shape(k))
            off += sz
        t1 = time.perf_counter()
        want = so.set_apply(coeffs, (h1, h2), sg)
        dt = time.perf_counter() - t1
        want = numpy.concatenate([want[k].ravel() for k in sg.keys])
        line["oracle_rel_err"] = float(numpy.linalg.norm(got - want) /
                                       numpy.linalg.norm(want))
        line["cpu_baseline"] = {
            "value": round(1.0 / dt, 5), "unit": "applies/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
            "kind": "port",
            "sample": (f"1 apply of the same workload by the NumPy oracle "
                       f"(dense string matrices, BLAS threads), {dt:.2f} s")}
    print(json.dumps(line), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(coeff, h1, h2, norb, na, nb, repeats=5):
    """The reference's compiled lm kernels (complex128) on the host cores:
    one warm-up apply, then the median of ``repeats`` applies with every
    core (OMP_PROC_BIND=close); plus one single-thread apply of the norb=12
    6a6b sector (seed 1) as a per-core record."""
    from oracle import ref_lm
    if not ref_lm.available():
        return None, None
    import platform
    import statistics
    import numpy
    import threadpoolctl
    from fqe_amd import synthetic
    threads = int(os.environ.get("OMP_NUM_THREADS", HOST_CORES))
    graph = ref_lm.RefGraph(na, nb, norb)
    out = ref_lm.apply_12_lm(coeff, h1, h2, graph)            # warm-up
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = ref_lm.apply_12_lm(coeff, h1, h2, graph)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    c12, h1s, h2s = synthetic.workload(12, 6, 6, 1, 12)
    g12 = ref_lm.RefGraph(6, 6, 12)
    with threadpoolctl.threadpool_limits(1):
        ref_lm.apply_12_lm(c12, h1s, h2s, g12)
        t0 = time.perf_counter()
        ref_lm.apply_12_lm(c12, h1s, h2s, g12)
        t1 = time.perf_counter() - t0
    return {"value": round(1.0 / med, 5), "unit": "applies/s",
            "cores": threads, "kind": "reference",
            "sample": (f"median of {repeats} applies (after 1 warm-up) of the "
                       f"same workload (norb={norb}, {na}a{nb}b) by the "
                       f"reference's compiled C lm kernels in complex128 "
                       f"(OpenMP {threads} threads, OMP_PROC_BIND="
                       f"{os.environ.get('OMP_PROC_BIND')}, scipy BLAS "
                       f"zaxpy, {ref_lm.ISA}); runs "
                       f"{', '.join(f'{t:.2f}' for t in times)} s; host CPU "
                       f"{_cpu_model()} ({platform.machine()})"),
            "single_thread": {"value": round(1.0 / t1, 4),
                              "unit": "applies/s",
                              "sample": ("1 apply of norb=12 6a6b (seed 1) "
                                         "on one thread")}}, out


def _ref_taylor(coeff, h1, h2, graph, time_, accuracy=1e-15, expansion=30):
    """The reference's Taylor loop (src/fqe/wavefunction.py:556-568) over its
    compiled apply, with the tensors of iht(t) = -i t (h1, h2)."""
    import math
    import numpy
    from oracle import ref_lm
    a1, a2 = -1j * time_ * h1, -1j * time_ * h2
    work = coeff.copy()
    result = coeff.copy()
    for order in range(1, expansion):
        work = ref_lm.apply_12_lm(work, a1, a2, graph)
        c = 1.0 / math.factorial(order)
        result += c * work
        if numpy.linalg.norm(work) * c < accuracy:
            return result, order
    raise RuntimeError("maximum taylor expansion limit reached")


def cfg1_line(with_cpu, chain=100, cpu_chain=10):
    """norb=12 6a6b (853,776 determinants), seed 1, 12 factors: ``chain``
    chained applies C <- H C / ||H C|| plus time_evolve(0.05) of the start
    vector, FP64 on the GPU (Wavefunction API, C resident in HBM); the
    reference's compiled C on the host times the first ``cpu_chain`` links
    (median per apply, scaled to ``chain``) and the whole Taylor series."""
    import statistics
    import numpy
    import torch
    import fqe_amd as fqe
    from fqe_amd import synthetic
    c, h1, h2 = synthetic.workload(12, 6, 6, 1, 12)
    key = (12, 0)
    ham = fqe.get_restricted_hamiltonian((h1, h2))
    wfn = fqe.Wavefunction([[12, 0, 12]])
    wfn.set_wfn(strategy="from_data", raw_data={key: c})
    x = wfn.apply(ham)            # warm-up: tables, operator, kernels
    x.normalize()
    wfn.time_evolve(0.05, ham)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = wfn
    links = []
    for _ in range(chain):
        x = x.apply(ham)
        x.normalize()
        if len(links) < cpu_chain:
            links.append(x)
    torch.cuda.synchronize()
    t_chain = time.perf_counter() - t0
    t0 = time.perf_counter()
    ev = wfn.time_evolve(0.05, ham)
    torch.cuda.synchronize()
    t_evolve = time.perf_counter() - t0
    line = {"workload": ("norb=12 6a6b (853,776 determinants), seed 1, 12 "
                         f"factors: {chain} chained applies (normalised) + "
                         "time_evolve(0.05), Taylor, accuracy 1e-15"),
            "gpu_fp64_s": round(t_chain + t_evolve, 4),
            "gpu_fp64_applies_per_s": round(chain / t_chain, 2),
            "gpu_fp64_time_evolve_s": round(t_evolve, 4),
            "taylor_terms": int(ev._taylor_terms)}
    if not with_cpu:
        return line
    from oracle import ref_lm
    if not ref_lm.available():
        return line
    g = ref_lm.RefGraph(6, 6, 12)
    y = c.copy()
    per = []
    errs = []
    for k in range(cpu_chain):
        t1 = time.perf_counter()
        y = ref_lm.apply_12_lm(y, h1, h2, g)
        y /= numpy.linalg.norm(y)
        if k:                        # the first link warms the caches
            per.append(time.perf_counter() - t1)
        errs.append(float(numpy.linalg.norm(links[k].get_coeff(key) - y)))
    t1 = time.perf_counter()
    ref_ev, nterms = _ref_taylor(c, h1, h2, g, 0.05)
    t_cpu_ev = time.perf_counter() - t1
    cpu_s = statistics.median(per) * chain + t_cpu_ev
    gev = ev.get_coeff(key)
    line.update({
        "cpu_reference_s": round(cpu_s, 3),
        "cpu_reference_time_evolve_s": round(t_cpu_ev, 3),
        "cpu_sample": (f"{cpu_chain - 1} chained applies timed (median per "
                       f"apply x {chain}) + the whole Taylor series, "
                       f"{HOST_CORES} host cores"),
        "speedup_gpu_vs_cpu": round(cpu_s / (t_chain + t_evolve), 1),
        "max_rel_err_chain_links": max(errs),
        "rel_err_time_evolve": float(numpy.linalg.norm(gev - ref_ev) /
                                     numpy.linalg.norm(ref_ev)),
        "reference_taylor_terms": int(nterms),
        "unitarity": float(abs(numpy.linalg.norm(gev) - 1.0))})
    return line


def _timed_evolve(wfn, ham, t, repeats=3):
    """Median wall time of wfn.time_evolve(t, ham) over ``repeats`` calls
    after one untimed call (tables, operator, graph recording), and the
    last result."""
    import statistics
    import torch
    ev = wfn.time_evolve(t, ham)
    torch.cuda.synchronize()
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ev = wfn.time_evolve(t, ham)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), ev


def time_evolve_lines(with_cpu):
    """Timed time_evolve (Taylor, accuracy 1e-15, expansion 30) on the
    Wavefunction API with C resident in HBM, per engine, eager and with the
    series recorded in one HIP graph:
    - norb=8 4a4b (4,900 determinants), seed 0, 8 factors: one apply plus
      time_evolve(0.1), against the reference's compiled C on the host;
    - norb=14 7a7b, seed 2, 14 factors: time_evolve(0.05) (the reference's
      C would take about a minute per series here, so no CPU leg)."""
    import numpy
    import torch
    import fqe_amd as fqe
    from fqe_amd import settings, synthetic
    saved = settings.contraction, settings.taylor_graph
    out = {}
    try:
        for name, (norb, nel, seed, t) in {"norb8_4a4b": (8, 4, 0, 0.1),
                                          "norb14_7a7b": (14, 7, 2, 0.05)
                                          }.items():
            c, h1, h2 = synthetic.workload(norb, nel, nel, seed, norb)
            key = (2 * nel, 0)
            ham = fqe.get_restricted_hamiltonian((h1, h2))
            wfn = fqe.Wavefunction([[2 * nel, 0, norb]])
            wfn.set_wfn(strategy="from_data", raw_data={key: c})
            line = {"workload": (f"norb={norb} {nel}a{nel}b ({c.size} "
                                 f"determinants), seed {seed}, {norb} "
                                 f"factors: time_evolve({t}), Taylor, "
                                 "accuracy 1e-15, expansion 30"),
                    "timing": "median of 3 calls after one untimed call"}
            res = {}
            for eng, graph in (("fp64", False), ("fp64", True),
                               ("ozaki", False), ("ozaki", True)):
                settings.contraction, settings.taylor_graph = eng, graph
                try:
                    sec, ev = _timed_evolve(wfn, ham, t)
                except RuntimeError as err:
                    line[f"{eng}{'_graph' if graph else ''}"] = str(err)
                    continue
                tag = f"{eng}{'_graph' if graph else ''}"
                line[f"gpu_{tag}_s"] = round(sec, 5)
                line["taylor_terms"] = int(ev._taylor_terms)
                res[tag] = ev.get_coeff(key)
            if "fp64" in res and "ozaki" in res:
                line["rel_err_ozaki_vs_fp64"] = float(
                    numpy.linalg.norm(res["ozaki"] - res["fp64"]) /
                    numpy.linalg.norm(res["fp64"]))
            settings.contraction, settings.taylor_graph = "fp64", False
            if norb == 8:
                t0 = time.perf_counter()
                wfn.apply(ham).get_coeff(key)
                line["gpu_fp64_apply_host_arrays_s"] = round(
                    time.perf_counter() - t0, 5)
            if with_cpu and norb == 8:
                from oracle import ref_lm
                if ref_lm.available():
                    g = ref_lm.RefGraph(nel, nel, norb)
                    t0 = time.perf_counter()
                    ref_lm.apply_12_lm(c, h1, h2, g)
                    ref_ev, nterms = _ref_taylor(c, h1, h2, g, t)
                    line["cpu_reference_apply_plus_evolve_s"] = round(
                        time.perf_counter() - t0, 4)
                    line["reference_taylor_terms"] = int(nterms)
                    line["cpu_sample"] = (f"one apply + the whole series, "
                                          f"{HOST_CORES} host cores")
                    for tag, v in res.items():
                        line[f"rel_err_{tag}_vs_reference"] = float(
                            numpy.linalg.norm(v - ref_ev) /
                            numpy.linalg.norm(ref_ev))
            out[name] = line
            del wfn
            torch.cuda.empty_cache()
    finally:
        settings.contraction, settings.taylor_graph = saved
    return out


def main():
    args = parse()
    if args.batched and args.gpus == 1:
        batched_main(args)
        return
    if args.davidson and args.gpus == 1:
        davidson_main(args)
        return
    if args.set and args.gpus == 1:
        set_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not "
              f"supported (the engine has no multi-GPU sharding)",
              file=sys.stderr)
        sys.exit(2)
    import numpy
    import torch
    import fqe_amd as fqe
    from fqe_amd import engine, _lib, settings
    from fqe_amd import synthetic

    if args.split_tile is not None:
        settings.split_tile = args.split_tile
    if args.no_split_fp64:
        settings.split_fp64 = False
    if args.split_tile_fp64 is not None:
        settings.split_tile_fp64 = args.split_tile_fp64
    if args.copy_probe:
        src = torch.empty(args.copy_probe, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}),
              file=sys.stderr, flush=True)
        del src, dst
    norb, na, nb = args.norb, args.nalpha, args.nbeta
    coeff, h1, h2 = synthetic.workload(norb, na, nb, args.seed, norb)
    if args.complex_op:
        rng = numpy.random.default_rng(args.seed + 1000)
        h1 = h1 + 0.05j * rng.normal(size=h1.shape)
        h1 = 0.5 * (h1 + h1.conj().T)
        h2 = h2 + 0.01j * rng.normal(size=h2.shape)
    graph = fqe.FciGraph(na, nb, norb)
    dev = _lib.device()
    op = engine.prepare_operator((h1, h2), norb, dev)
    n = coeff.size

    engines = {"fp64": (torch.complex128, "fp64"),
               "ozaki": (torch.complex128, "ozaki"),
               "bf16x3": (torch.complex64, "fp64")}
    modes = [m for m in engines if m in args.modes.split(",")]
    if args.unfused:
        # the Ozaki-digit contractions exist only in the split kernels
        modes = [m for m in modes if m != "ozaki"]
    results, outs = {}, {}
    for mode in modes:
        dt, contraction = engines[mode]
        run = EngineRun(graph, op, coeff, dt, fused=not args.unfused,
                        contraction=contraction)
        wall, per = run.run(args.steps, args.warmup)
        outs[mode] = run.result()
        results[mode] = {"wall": wall, "per": per, "fused": run.fused,
                         "tile": int(run.s.tile), "digits_beta":
                         bool(run.s.hbb_oz)}
        print(f"{mode}: {args.steps / wall:.2f} applies/s "
              f"{ {k: round(v, 3) for k, v in per.items()} }",
              file=sys.stderr, flush=True)
        run.free()
        del run
    if set(modes) != set(engines):
        for mode in modes:
            line = {"mode": mode,
                    "applies_per_s": args.steps / results[mode]["wall"],
                    "kernels_ms": results[mode]["per"]}
            if not results[mode]["fused"]:
                # the (b) gather, (c) pair GEMM and (d) scatter against
                # their own rooflines
                esz = 8 if mode == "fp64" else 4
                work = kernel_work(op, graph, n, esz, mode != "fp64")
                line["rooflines"] = [
                    roofline_entry(k, results[mode]["per"], work,
                                   "fp64" if mode == "fp64" else "mixed")
                    for k in ("dvec_gather", "pair_gemm", "scatter")]
            print(json.dumps(line))
        return

    def rel(a, b):
        return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))

    cpu, ref_out = (None, None)
    if not args.no_cpu_baseline:
        cpu, ref_out = cpu_baseline(coeff, h1, h2, norb, na, nb)
    cfg1 = None
    if not args.no_cfg1 and (norb, na, nb) == (14, 7, 7):
        cfg1 = cfg1_line(not args.no_cpu_baseline)
    host_rate = host_boundary_rate(coeff, h1, h2, norb, na, nb)
    api = api_rates(coeff, h1, h2, norb, na, nb)
    evolve = None
    if not args.no_evolve and (norb, na, nb) == (14, 7, 7):
        evolve = time_evolve_lines(not args.no_cpu_baseline)

    oz, fp, bf = results["ozaki"], results["fp64"], results["bf16x3"]
    value = args.steps / oz["wall"]
    fp64_value = args.steps / fp["wall"]
    bf_value = args.steps / bf["wall"]
    work_o = kernel_work(op, graph, n, 8, False, ozaki=True)
    work_d = kernel_work(op, graph, n, 8, False)
    work_m = kernel_work(op, graph, n, 4, True)
    dom_o = max(oz["per"], key=oz["per"].get)
    dom_d = max(fp["per"], key=fp["per"].get)
    dom_m = max(bf["per"], key=bf["per"].get)
    traffic = measured_traffic(norb, na, nb)
    err_oz = rel(outs["ozaki"], outs["fp64"])
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * oz["wall"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c128 storage; int8 digit products on MFMA (Ozaki: A' 5 "
                  "digits, D' and C 48-bit integers, 20 digit pairs), exact "
                  "int32 accumulation, FP64 assembly"),
        "data": "synthetic",
        "config": {"workload": (f"RestrictedHamiltonian apply, norb={norb} "
                                f"{na}a{nb}b ({n} determinants), seed "
                                f"{args.seed}, {norb} factors"),
                   "pair_basis": "compressed" if op.compressed else "ordered",
                   "npair": op.npair,
                   "kernels": "split" if oz["fused"] else "unfused",
                   "digit_beta_gemm": oz["digits_beta"]},
        "rel_err_mixed_vs_fp64_gpu": err_oz,
        "gate_1e-9_met": err_oz <= 1e-9,
        "speedup_mixed_vs_fp64": round(value / fp64_value, 3),
        "kernels_ms": {k: round(v, 4) for k, v in oz["per"].items()},
        "roofline": roofline_entry(dom_o, oz["per"], work_o, "ozaki",
                                   traffic.get(dom_o + "_ozaki")),
        "fp64": {"value": round(fp64_value, 3), "unit": "applies/s",
                 "ms_per_step": round(1e3 * fp["wall"] / args.steps, 3),
                 "kernels_ms": {k: round(v, 4) for k, v in fp["per"].items()},
                 "roofline": roofline_entry(
                     dom_d, fp["per"], work_d, "fp64",
                     traffic.get(dom_d + "_fp64"))},
        "bf16x3_c64": {
            "note": ("complex64 storage, BF16 hi/lo products with FP32 "
                     "accumulation: outside the 1e-9 gate"),
            "value": round(bf_value, 3), "unit": "applies/s",
            "ms_per_step": round(1e3 * bf["wall"] / args.steps, 3),
            "rel_err_vs_fp64_gpu": rel(outs["bf16x3"], outs["fp64"]),
            "kernels_ms": {k: round(v, 4) for k, v in bf["per"].items()},
            "roofline": roofline_entry(dom_m, bf["per"], work_m, "mixed",
                                       traffic.get(dom_m))},
        "host_boundary_applies_per_s": round(host_rate, 3),
        "wavefunction_apply_applies_per_s": api,
        "cpu_baseline": cpu,
    }
    if ref_out is not None:
        line["rel_err_mixed_vs_reference"] = rel(outs["ozaki"], ref_out)
        line["rel_err_fp64_vs_reference"] = rel(outs["fp64"], ref_out)
        line["bf16x3_c64"]["rel_err_vs_reference"] = rel(outs["bf16x3"],
                                                         ref_out)
    if cfg1 is not None:
        line["cfg1"] = cfg1
    if evolve is not None:
        line["time_evolve"] = evolve
    print(json.dumps(line))


if __name__ == "__main__":
    main()
