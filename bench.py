#!/usr/bin/env python3
"""Benchmark: condensed + recovered cells/s of the HDG hybrid path on one MI355X.

A step is one pass of the hot path over the whole mesh, with every input
resident in HBM: fused cell integration + static condensation of every cell
(hdg_condense), Dirichlet lifting + deterministic skeleton assembly
(hdg_assemble), and backward recovery of (q, u) from a given trace
(hdg_recover).  The global skeleton solve is outside the step (it stays a host
sparse direct solve; at 64^3 it is impractical), so recovery takes the
manufactured trace interpolated at the free trace nodes.

Workload (default): 3D Darcy HDG, 64^3 hexes, k=2 -- 262,144 cells, ni=108,
nb=54 -- kappa lognormal(0,1) per cell from the seed.  The 2D Darcy HDG
512^2 k=2 case is measured in the same run as a secondary line item.

Prints ONE JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "gridaphybrid.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "condensed+recovered cells/s (mixed vs FP64) at 1e-10 rel err; p50/p99 latency"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
SEED = 20261202


def f_cond(ni, nb):
    """LAPACK-nominal flops of one cell's condensation (getrf + getrs + gemm +
    getrs + gemv, src/StaticCondensationMap.jl:179-192)."""
    return 2 * ni ** 3 / 3 + 2 * ni * ni * nb + 2 * ni * nb * nb + 2 * ni * ni + 2 * ni * nb


def f_rec(ni, nb):
    return 2 * ni * nb + 2 * ni * ni


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n3d", type=int, default=64, help="cells per direction of the 3D mesh")
    ap.add_argument("--n2d", type=int, default=512, help="cells per direction of the 2D mesh")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-2d", action="store_true", help="skip the secondary 2D line item")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


class Case:
    """One mesh + operator, with device buffers for the recovery input/output."""

    def __init__(self, gh, ctx, D, n, k):
        import torch
        from gridaphybrid import _lib
        self.gh, self.ctx, self.D, self.n, self.k = gh, ctx, D, n, k
        nc = n ** D
        rng = np.random.default_rng(SEED)
        self.kappa = np.exp(rng.standard_normal(nc))
        model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
        X = gh.HDGSpaces.darcy_baseline(model, k)
        wf = gh.DarcyHDG(tau=1.0, kappa=self.kappa, source=_lib.SOURCE_DARCY_SINPROD)
        t0 = time.perf_counter()
        self.op = gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx)
        self.setup_s = time.perf_counter() - t0
        info = self.op.info
        self.info = info
        lam = np.zeros(info.nfree)
        _lib.check(_lib.lib().hdg_mesh_interpolate_trace(self.op.mesh, lam.ctypes.data), ctx.h,
                   "interpolate")
        self.lam_host = lam
        with ctx.stream():
            self.lam = torch.as_tensor(lam, device="cuda")
            self.out = torch.empty(info.nfree_total, dtype=torch.float64, device="cuda")
        ctx.synchronize()

    def step(self, ev=None):
        from gridaphybrid import _lib
        L = _lib.lib()
        h = self.ctx.h
        if ev:
            ev[0].record()
        _lib.check(L.hdg_condense(self.op.mesh, 0), h, "condense")
        if ev:
            ev[1].record()
        _lib.check(L.hdg_assemble(self.op.mesh), h, "assemble")
        if ev:
            ev[2].record()
        _lib.check(L.hdg_recover(self.op.mesh, self.lam.data_ptr(), self.out.data_ptr()), h,
                   "recover")
        if ev:
            ev[3].record()

    def run(self, steps, warmup):
        import torch
        ctx = self.ctx
        with ctx.stream():
            for _ in range(warmup):
                self.step()
            ctx.synchronize()
            evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ctx.synchronize()
            t0 = time.perf_counter()
            start.record()
            for s in range(steps):
                self.step(evs[s])
            stop.record()
            ctx.synchronize()
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        total_ms = start.elapsed_time(stop)
        cond = np.mean([e[0].elapsed_time(e[1]) for e in evs])
        asm = np.mean([e[1].elapsed_time(e[2]) for e in evs])
        rec = np.mean([e[2].elapsed_time(e[3]) for e in evs])
        nbad = C.c_int64(0)
        from gridaphybrid import _lib
        _lib.check(_lib.lib().hdg_mesh_singular_count(self.op.mesh, C.byref(nbad)), ctx.h, "count")
        return dict(ms_per_step=total_ms / steps, wall_s=wall, condense_ms=cond, assemble_ms=asm,
                    recover_ms=rec, singular_cells=int(nbad.value))

    def summary(self, r):
        info = self.info
        nc = info.ncells
        ni, nb = info.ni, info.nb
        cells_s = nc / (r["ms_per_step"] * 1e-3)
        fl = f_cond(ni, nb) * nc
        achieved = fl / (r["condense_ms"] * 1e-3) / 1e12
        return dict(cells=nc, ni=ni, nb=nb, nfree=info.nfree, nnz=info.nnz,
                    cells_per_s=cells_s, ms_per_step=r["ms_per_step"],
                    condense_ms=r["condense_ms"], assemble_ms=r["assemble_ms"],
                    recover_ms=r["recover_ms"], condense_tflops=achieved,
                    condense_frac_fp64_peak=achieved / PEAK_FP64_TFLOPS,
                    singular_cells=r["singular_cells"], setup_s=self.setup_s)


def cpu_baseline(case, seconds):
    """The oracle (C, OpenMP) on a bounded sample of the same workload:
    per-cell integration + condensation + recovery (with the reference's
    re-factorisation), on the host cores.  Also returns the GPU-vs-oracle
    error per field on the sample."""
    from oracle import oracle as O
    import torch
    from gridaphybrid import _lib
    D, n, k = case.D, case.n, case.k
    prob = O.Problem.darcy(D, k, mms=O.MMS_SINPROD)
    mesh = O.cartesian(D, tuple([n] * D))
    s = prob.sizes()
    ni, nb, na = s["ni"], s["nb"], s["n_flux"]
    cd, nfree, ndir, fd = O.trace_dofs(mesh, s["n_tf"])
    lam_dir = O.dirichlet_values(prob, mesh, fd, ndir)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def run(cells):
        A, b = O.cell_systems(prob, mesh, case.kappa, cells=cells)
        S, g, info = O.condense(A, b, ni, nb)
        xb = O.gather_lambda(cd[cells], case.lam_host, lam_dir)
        xi, info2 = O.recover(A, b, xb, ni, nb)
        return S, g, xi

    # calibrate, then time a sample sized to about `seconds` of CPU work
    cal = np.arange(min(256, mesh.ncells))
    t0 = time.perf_counter()
    run(cal)
    per = (time.perf_counter() - t0) / len(cal)
    ns = int(min(mesh.ncells, max(256, seconds / max(per, 1e-9))))
    cells = np.arange(ns)
    t0 = time.perf_counter()
    S, g, xi = run(cells)
    dt = time.perf_counter() - t0
    # GPU results for the same cells
    Sd, gd, cdd = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L = _lib.lib()
    _lib.check(L.hdg_cell_blocks(case.op.mesh, C.byref(Sd), C.byref(gd), C.byref(cdd)), case.ctx.h,
               "blocks")
    Sg = np.empty((ns, nb, nb))
    gg = np.empty((ns, nb))
    L.hdg_memcpy_d2h(case.ctx.h, Sg.ctypes.data, Sd, Sg.nbytes)
    L.hdg_memcpy_d2h(case.ctx.h, gg.ctypes.data, gd, gg.nbytes)
    fv = case.out.cpu().numpy()
    nc = mesh.ncells
    qg = fv[: ns * na].reshape(ns, na)
    ug = fv[nc * na: nc * na + ns * (ni - na)].reshape(ns, ni - na)

    def relerr(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    err = dict(S=relerr(Sg.transpose(0, 2, 1), S), g=relerr(gg, g), q=relerr(qg, xi[:, :na]),
               u=relerr(ug, xi[:, na:]))
    # same inputs: the oracle's condensation + recovery applied to the GPU's own
    # cell systems isolates the arithmetic from the conditioning of A_ii
    # (cond up to ~1e7 at h = 1/64 with lognormal kappa)
    nsame = min(ns, 2048)
    nt = ni + nb
    Ad = torch.empty((nsame, nt, nt), dtype=torch.float64, device="cuda")
    bd = torch.empty((nsame, nt), dtype=torch.float64, device="cuda")
    _lib.check(L.hdg_cell_systems(case.op.mesh, 0, nsame, Ad.data_ptr(), bd.data_ptr()),
               case.ctx.h, "cell_systems")
    case.ctx.synchronize()
    Ag = Ad.transpose(1, 2).cpu().numpy()
    bg = bd.cpu().numpy()
    S2, g2, _ = O.condense(Ag, bg, ni, nb)
    xi2, _ = O.recover(Ag, bg, O.gather_lambda(cd[:nsame], case.lam_host, lam_dir), ni, nb)
    err_same = dict(S=relerr(Sg[:nsame].transpose(0, 2, 1), S2), g=relerr(gg[:nsame], g2),
                    q=relerr(qg[:nsame], xi2[:, :na]), u=relerr(ug[:nsame], xi2[:, na:]),
                    cells=nsame)
    err = dict(same_cell_systems=err_same, oracle_integration=err)
    _ = torch
    return dict(value=ns / dt, unit="cells/s", cores=threads, kind="port",
                sample=f"first {ns} cells of the {n}^{D} mesh (k={k}): oracle cell integration + "
                       f"StaticCondensationMap + BackwardStaticCondensationMap (re-factorising "
                       f"A_ii as the reference does), OpenMP over cells, FP64; {dt:.1f} s"), err


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the per-cell path shards by cell ranges, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import gridaphybrid as gh
    ctx = gh.Context(0)
    main_case = Case(gh, ctx, 3, args.n3d, args.order)
    r = main_case.run(args.steps, args.warmup)
    s3 = main_case.summary(r)
    cpu, err = (None, None)
    if not args.no_cpu:
        cpu, err = cpu_baseline(main_case, args.cpu_seconds)
    ni, nb, nc = s3["ni"], s3["nb"], s3["cells"]
    roof = dict(bound="mfma", achieved=s3["condense_tflops"], peak=PEAK_FP64_TFLOPS,
                unit="TFLOP/s", frac=s3["condense_tflops"] / PEAK_FP64_TFLOPS, traffic=None,
                kernel="condense_kernel (fused cell integration + static condensation, FP64)",
                per_cell_flops=f_cond(ni, nb))
    main_case.op.close()
    del main_case
    torch.cuda.empty_cache()
    s2 = None
    if not args.no_2d:
        c2 = Case(gh, ctx, 2, args.n2d, args.order)
        s2 = c2.summary(c2.run(args.steps, args.warmup))
        c2.op.close()
    line = {
        "metric": METRIC,
        "value": s3["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s3["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal kappa, manufactured solution)",
        "config": {
            "workload": f"3D Darcy HDG {args.n3d}^3 hex k={args.order} Q_k/Q_k/Q_k, "
                        f"{nc} cells, ni={ni}, nb={nb}; FP64 GPU path; step = condense + "
                        f"assemble + recover",
            "precision": "fp64",
            "cells": nc,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": {"condense": s3["condense_ms"], "assemble": s3["assemble_ms"],
                      "recover": s3["recover_ms"]},
        "err_vs_oracle_sample": err,
        "case_3d": s3,
        "case_2d": s2,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
