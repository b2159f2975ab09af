#!/usr/bin/env python3
"""Benchmark of the CCATT chemistry step on one MI355X.

Workload (N=1): one 120 s chemistry step over a 200x200x40 grid (1.6M cells,
25 000 blocks of 64), SZA varying smoothly in x (half day, half night), a
Gaussian NOx/VOC plume (sigma 20 cells, x50), seed 0, rtol 1e-3,
atol 1e-12 ppb -- the "plain GPU form" workload.  Every step restores the
same initial state in HBM (a device-to-device copy) and integrates it, so
each timed step does identical work.

value: cell-steps/s of the GPU path, inputs resident in HBM.  Until the
surrogate+fallback path exists, the GPU path is the plain FP64 RODAS3 solver
and the line says so in config.path.

roofline: the dominant kernel (the wave64 RODAS3 kernel over all 64-cell
blocks) against the FP64 peak; algorithmic FLOP per launch from the generated
code's exact counts times the device-counted sub-steps and attempts of every
cell; duration from HIP events on the launch stream.

cpu_baseline: the oracle (C restatement of the reference step) on a bounded,
representative sample of the same workload (every 5th column in i and j, all
40 levels: 64 000 cells), all host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6


def flop_model():
    """Algorithmic FP64 FLOP of the reference algorithm, per cell."""
    with open(os.path.join(ROOT, "brams_amd", "csrc", "generated", "relacs_tuv_flops.json")) as fh:
        f = json.load(fh)
    ns = 47
    # per attempt: diagonal update, elimination, 4 solves, f(y3), f(y4), stage
    # vectors (rhs2 2n, y3 2n, rhs3 3n, y4 3n, rhs4 4n, y1 4n, tol 3n, err 3n)
    # and the norm (2 per cell)
    vec = (2 + 2 + 3 + 3 + 4 + 4 + 3 + 3) * ns + 2
    f_att = ns + f["elimination"] + 4 * f["solve"] + 2 * f["fun"] + vec
    # per accepted sub-step: Jacobian (without the diagonal shift) and f(y0)
    f_sub = f["jacobian_assembly"] - ns + f["fun"]
    # per call: rate coefficients (transcendentals counted as one each) and
    # the unit conversions of the gather/scatter
    tr = sum(f["kinetic_transcendentals"].values())
    f_call = f["kinetic_arith"] + tr + 12 + 2 * ns + 2 * ns
    return f_call, f_sub, f_att


def sample_workload(wl, stride=5):
    """Every stride-th column of a workload, as a smaller grid of its own."""
    import numpy as np
    from brams_amd.workload import Grid, alloc_index
    g = wl["grid"]
    ii = np.arange(g.ia - 1, g.iz, stride)
    jj = np.arange(g.ja - 1, g.jz, stride)
    sg = Grid(len(ii), len(jj), g.nz, g.dtlt)

    def cut(a):
        out = np.zeros(a.shape[:-3] + (sg.m1, sg.m2, sg.m3), dtype=a.dtype, order="F")
        out[..., :, 1:-1, 1:-1] = a[..., ii, :][..., jj]
        return np.asfortranarray(out)

    s = dict(grid=sg, seed=wl["seed"])
    for k in ("pp", "pi0", "theta", "rv", "jphoto"):
        s[k] = cut(wl[k])
    s["sc_p"] = [cut(x) for x in wl["sc_p"]]
    s["block_end"] = alloc_index(sg)
    s["last_accepted_dt"] = np.full(len(s["block_end"]), g.dtlt)
    return s


def cpu_baseline(wl, atol, rtol):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    sub = sample_workload(wl)
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    hs, st, mask, rc = oracle_lib.run_oracle(sub, atol, rtol, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": sub["grid"].ncell / dt, "unit": "cell-steps/s", "cores": cores,
            "kind": "port",
            "sample": "oracle (C restatement, -O2) on every 5th column of the workload in i and j: "
                      "%d cells x 1 step, %.1f s wall, %d attempts" % (sub["grid"].ncell, dt, st.attempts)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--tol", default="baseline", choices=["baseline", "reference"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the chemistry step has no cross-GPU exchange "
              "on one grid; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    from brams_amd.workload import Grid, make_workload, tolerances

    grid = Grid(args.nx, args.ny, args.nz, dtlt=120.0)
    wl = make_workload(grid, seed=0, sza="split", plume_sigma=20.0, plume_factor=50.0)
    atol, rtol = tolerances(args.tol)
    step = DeviceStep(wl, atol, rtol, device=0)
    print("bench: workload %dx%dx%d = %d cells, %d blocks" % (args.nx, args.ny, args.nz, grid.ncell,
                                                               len(wl["block_end"])), file=sys.stderr, flush=True)
    for i in range(args.warmup):
        step.reset()
        step.step()
        print("bench: warmup %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    main_ms, stats = [], None
    t0 = time.perf_counter()
    for i in range(args.steps):
        step.reset()
        rc = step.step()
        s = abi.stats()
        main_ms.append(s.main_ms)
        stats = s
        print("bench: step %d rc=%d kernel %.1f ms (block-1 %.1f ms, main %.1f ms)"
              % (i, rc, s.kernel_ms, s.block1_ms, s.main_ms), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = grid.ncell * args.steps / wall

    f_call, f_sub, f_att = flop_model()
    flops_all = f_call * stats.cells + f_sub * stats.cell_substeps + f_att * stats.cell_attempts
    frac_main = stats.main_cells / stats.cells
    flops_main = flops_all * frac_main
    avg_main_ms = sum(main_ms) / len(main_ms)
    achieved = flops_main / (avg_main_ms * 1e-3) / 1e12
    line = {
        "metric": "chem cell-steps/s, surrogate+fallback vs FP64 Rosenbrock, at <=1% rel err",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox generator, brams_amd/workload.py)",
        "config": {"workload": "RELACS_TUV RODAS3 step, %dx%dx%d cells (1.6M), dt 120 s, SZA split "
                               "day/night, NOx/VOC plume sigma 20 x50, seed 0" % (args.nx, args.ny, args.nz),
                   "path": "plain FP64 RODAS3 (GPU)", "tolerances": args.tol,
                   "cells": grid.ncell, "blocks": len(wl["block_end"]),
                   "attempts_per_cell": stats.cell_attempts / stats.cells,
                   "substeps_per_cell": stats.cell_substeps / stats.cells,
                   "failed_blocks": stats.failed_blocks, "kernel_ms": stats.kernel_ms,
                   "block1_ms": stats.block1_ms},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "rodas3_blocks<64,false>", "avg_ms": avg_main_ms,
                     "flop_per_launch": flops_main,
                     "flop_model": {"per_cell_call": f_call, "per_substep": f_sub, "per_attempt": f_att},
                     "peak_note": "FP64 vector (the kernel issues no MFMA; MI355X FP64 vector "
                                  "and matrix peaks are both 78.6 TFLOP/s)"},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(wl, atol, rtol)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
