#!/usr/bin/env python3
"""Benchmark of the CCATT chemistry step on one MI355X.

Workload: a 200x200x40 grid (1.6M cells, 25 000 blocks of 64), dt 120 s,
SZA varying smoothly in x (half day, half night) and advancing 0.5 deg per
step, a Gaussian NOx/VOC plume (sigma 20 cells, x50), seed 0, rtol 1e-3,
atol 1e-12 ppb.  The seeded state is first advanced by --spinup FP64 RODAS3
steps (untimed; the first of them, the cold step out of the synthetic state,
is also reported on its own).  From that warm state:

  leg A (value): --steps consecutive surrogate+fallback steps
                 (brams_sur_step), bracketed by synchronisations;
  leg B:         the same --steps steps by the plain FP64 RODAS3 solver, the
                 accuracy reference (per-species median / p99 relative error
                 of leg A's states after every step) and the speed reference.

Inputs stay resident in HBM; each leg A step also copies the state to a
snapshot buffer for the accuracy check (a 322 MB device copy, ~0.1 ms).
--path fp64 times leg B alone as the headline (no surrogate model needed).
--path ros2 times the ROS2 solver (CHEMISTRY=3) the same way, from its own
spin-up; --path qssa the QSSA solver (CHEMISTRY=1); --path tuv the FAST-TUV
j-value mapping into jphoto; --path nested the three nested grids of one
model step.

roofline: the kernel with the largest share of leg A's time, against its
peak (FP64 for the RODAS3 fallback kernel, BF16 MFMA for the fused MLP);
the plain FP64 kernel of leg B and the other kernels are listed too.
Algorithmic FLOP: the generated code's exact counts times the device-counted
sub-steps and attempts of every cell (RODAS3), 1.356 MFLOP per cell for the
3-member 4x256 ensemble (MLP).  Durations from HIP events on the launch
stream.

cpu_baseline: the oracle (C restatement of the reference step) on block 1
and every 4th block of the warm state (6 251 blocks, 400 064 cells, one step
with each block's carried last_accepted_dt), all host cores; the cold step
from the seeded state on block 1 and every 40th block is reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6


def mech_flops():
    """The generator's operation counts (generated/<mech>_flops.json) and the
    species count of the process's mechanism (BRAMS_MECH, RELACS_TUV by
    default)."""
    from brams_amd.mechanism import default_name, load_mechanism
    with open(os.path.join(ROOT, "brams_amd", "csrc", "generated", "%s_flops.json" % default_name())) as fh:
        f = json.load(fh)
    return f, load_mechanism().nspecies


def flop_model():
    """Algorithmic FP64 FLOP of the reference algorithm, per cell."""
    f, ns = mech_flops()
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


def ros2_flop_model():
    """Algorithmic FP64 FLOP of the ROS2 step (chem_spack_ros_dyndt.f90), per
    cell: per attempt the diagonal update, the elimination, 2 solves, f(y+k1/g)
    and the stage vectors (y+k1/g 2n, rhs2 3n, y_new 4n, error term 4n, tol
    3n, err 3n) with the norm; per accepted sub-step and per call as RODAS3."""
    f, ns = mech_flops()
    vec = (2 + 3 + 4 + 4 + 3 + 3) * ns + 2
    f_att = ns + f["elimination"] + 2 * f["solve"] + f["fun"] + vec
    f_call, f_sub, _ = flop_model()
    return f_call, f_sub, f_att


def qssa_flop_model(itchim):
    """Algorithmic FP64 FLOP of one QSSA call per cell (chem_spack_qssa.f90):
    the rate coefficients and conversions (as RODAS3's per-call term), the
    P/L split, per species the setup (ybid, P/ybid, the exponent) and the
    tendency (10), and per species and sub-step one subtraction and one FMA."""
    f, ns = mech_flops()
    f_call, _, _ = flop_model()
    return f_call + f["prod_loss"] + 10 * ns + 3 * ns * itchim


# PMC-measured HBM bytes per cell of qssa_cells (read + written),
# profiles/qssa_kernel_stats_and_traffic.md
QSSA_TRAFFIC_PER_CELL = 881 + 289
# the same for tuv_map_interior, bytes per interior grid point at kj 80, kz
# 125 with 17 reactions mapped (profiles/tuv_jphoto_mapping.md)
TUV_TRAFFIC_PER_POINT = 254


MLP_FLOP_PER_CELL = 3 * 2 * (68 * 256 + 3 * 256 * 256 + 256 * 47)
BF16_PEAK_TFLOPS = 2500.0
# HBM traffic measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by
# the counters' reading of a 1 GiB device copy: x2.0 and x1.0) on this
# workload, profiles/rodas3_hbm_traffic_pmc.md (RODAS3) and
# profiles/ros2_kernel_stats_and_traffic.md (ROS2).  Per cell-attempt (read +
# written bytes), the same in every launch of a kernel: the spill traffic of
# the per-lane LU and stage vectors, which depends on the register budget
# (waves per SIMD).  The fused MLP: bytes per cell of one launch.  A kernel
# without a measurement reports traffic null.
BYTES_PER_CELL_ATTEMPT = {
    "rodas3_blocks<64,false,false,4>": 49000 + 14200,
    "rodas3_blocks<64,false,false,2>": 39200 + 10800,
    "rodas3_blocks<64,false,false,1>": 29700 + 8300,
    "rodas3_blocks<64,false,true,2>": 39000 + 11100,
    "rodas3_blocks<64,false,true,1>": 29300 + 8500,
    "ros2_blocks<64,4>": 31300 + 9350,          # profiles/ros2_kernel_stats_and_traffic.md
}
# the same for the lean library of another mechanism, by (mechanism, kernel):
# profiles/other_mechanisms_solvers.md
BYTES_PER_CELL_ATTEMPT_MECH = {
    ("RELACS_MX", "rodas3_blocks<64,false,false,1>"): 29900 + 8400,
    ("RELACS_MX", "rodas3_blocks<64,false,false,4>"): 49100 + 14100,
    ("CB07_TUV", "rodas3_blocks<64,false,false,1>"): 17300 + 5000,
    ("CB07_TUV", "rodas3_blocks<64,false,false,4>"): 30800 + 9700,
    ("RACM_TUV", "rodas3_blocks<64,false,false,1>"): 78700 + 19000,
    ("RACM_TUV", "rodas3_blocks<64,false,false,4>"): 111700 + 31300,
}
MLP_BYTES_PER_CELL = 1825
FORM_WPS = {3: 1, 4: 2, 5: 4}


def spill_traffic(kernel, cell_attempts):
    b = BYTES_PER_CELL_ATTEMPT.get(kernel)
    return None if b is None else b * cell_attempts


def cpu_baseline(warm_wl, cold_wl, atol, rtol, warm_every=1, cold_every=4, solver="rodas3"):
    """The oracle (C restatement of the reference step, gcc -O2, pthreads over
    blocks) on every warm_every-th block of the warm state (every block: the
    whole warm step), with each block's carried last_accepted_dt (the
    reference's semantics for those blocks), and on every cold_every-th block
    of the seeded cold state; block 1 runs first in both, as it fixes the
    mask (RODAS3).  solver "ros2": chem_ros_dyndt's restatement, no
    calibration run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib
    cores = len(os.sched_getaffinity(0))
    out = {}
    for tag, wl, every in (("warm", warm_wl, warm_every), ("cold", cold_wl, cold_every)):
        nob = len(wl["block_end"])
        blocks = np.arange(max(every, 1), nob, every, dtype=np.int32)
        cells = int(wl["block_end"][0]) + int(np.sum(np.asarray(wl["block_end"])[blocks]))
        t0 = time.perf_counter()
        hs, st, mask, rc = oracle_lib.run_oracle_blocks(wl, atol, rtol, blocks, nthreads=cores, solver=solver)
        dt = time.perf_counter() - t0
        out[tag] = {"value": cells / dt, "cells": cells, "blocks": len(blocks) + 1, "wall_s": dt,
                    "cell_attempts": int(st.cell_attempts)}
    w, c = out["warm"], out["cold"]
    calib = calibrate_oracle(warm_wl, atol, rtol, cores) if solver == "rodas3" else None
    name = {"rodas3": "chem_rodas3_dyndt", "ros2": "chem_ros_dyndt"}[solver]
    return {"value": w["value"], "unit": "cell-steps/s", "cores": cores, "kind": "port",
            "calibration": calib,
            "sample": "oracle (C restatement of " + name + " with the reference's kernels' "
                      "arithmetic, gcc -O2 -ffp-contract=off, pthreads over blocks) on %s of the warm state "
                      "(%d blocks, %d cells x 1 step, carried last_accepted_dt, %.1f s wall, "
                      "%d cell-attempts)" % (_which(warm_every), w["blocks"], w["cells"], w["wall_s"],
                                             w["cell_attempts"]),
            "cold_value": c["value"],
            "cold_sample": "%s of the seeded cold state (%d cells, %.1f s wall, %d cell-attempts)"
                           % (_which(cold_every), c["cells"], c["wall_s"], c["cell_attempts"])}


def _which(every):
    return "every block" if every <= 1 else "block 1 and every %dth block" % every


def calibrate_oracle(wl, atol, rtol, cores, every=200):
    """The oracle against the reference's own compiled pieces: the same block
    sample (block 1 and every `every`-th block of the warm state) once with the
    oracle's LU and once with its LU routed through the reference's compiled
    Sparse1.3 (spOrderAndFactor / spFactor / spSolve on the value-ordered
    matrix, oracle/_ref).  Equal outputs and a comparable time show that the
    port is not a slowed-down restatement."""
    import numpy as np
    import oracle_lib
    nob = len(wl["block_end"])
    blocks = np.arange(every, nob, every, dtype=np.int32)
    res = {}
    try:
        for tag, ext in (("oracle_lu", False), ("reference_sparse", True)):
            oracle_lib.use_reference_sparse(ext)
            t0 = time.perf_counter()
            hs, st, mask, rc = oracle_lib.run_oracle_blocks(wl, atol, rtol, blocks, nthreads=cores)
            res[tag] = (time.perf_counter() - t0, [np.asarray(x).copy() for x in hs.sc_p], rc)
    except OSError as e:                                # oracle/_ref not built on this machine
        return {"error": "reference pieces not available: %s" % e}
    finally:
        oracle_lib.use_reference_sparse(False)
    same = all(np.array_equal(a, b) for a, b in zip(res["oracle_lu"][1], res["reference_sparse"][1]))
    return {"sample": "block 1 and every %dth block of the warm state (%d blocks)" % (every, len(blocks) + 1),
            "oracle_lu_s": res["oracle_lu"][0], "reference_sparse_s": res["reference_sparse"][0],
            "outputs_identical": bool(same)}


def rodas3_flops(cells, cell_substeps, cell_attempts):
    f_call, f_sub, f_att = flop_model()
    return f_call * cells + f_sub * cell_substeps + f_att * cell_attempts


def bench_nested(args):
    """Three nested grids per step from the same seeded states (cold steps):
    first one chem_rodas3_dyndt call per grid, as the reference drives nested
    grids (chemistry.f90:313-320), then all three in one pass
    (brams_chem_rodas3_dyndt_device_grids).  value: cell-steps/s of the
    one-pass form over the 2.08M cells; both forms' times are in the line."""
    import ctypes as C
    import torch
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    from brams_amd.workload import Grid, make_workload, tolerances
    atol, rtol = tolerances(args.tol)
    grids = [Grid(200, 200, 40, dtlt=120.0), Grid(98, 98, 40, dtlt=60.0), Grid(50, 50, 40, dtlt=30.0)]
    steps = [DeviceStep(make_workload(g, seed=0, sza="split", plume_sigma=20.0 * g.nx / 200.0,
                                      plume_factor=50.0), atol, rtol) for g in grids]
    ncell = sum(g.ncell for g in grids)
    L = abi.lib()
    L.brams_chem_rodas3_dyndt_device_grids.argtypes = [C.POINTER(abi.ChemDevArgs), C.c_int32]
    arr = (abi.ChemDevArgs * len(steps))(*[st.args for st in steps])
    print("bench: nested grids, %d cells" % ncell, file=sys.stderr, flush=True)

    def calls():
        out = []
        for st in steps:
            st.reset()
            st.step()
            s = abi.stats()
            out.append((s.kernel_ms, s.cell_attempts, s.cells))
        return out

    def one_pass():
        for st in steps:
            st.reset()
        rc = L.brams_chem_rodas3_dyndt_device_grids(arr, len(steps))
        if rc < 0:
            raise RuntimeError(L.brams_chem_last_error().decode())
        s = abi.stats()
        return [(s.kernel_ms, s.cell_attempts, s.cells)]

    timed = {}
    for name, fn in (("calls", calls), ("pass", one_pass)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        per = []
        for i in range(args.steps):
            per.append(fn())
            print("bench: nested %s step %d: %s ms" % (name, i, ", ".join("%.1f" % x[0] for x in per[-1])),
                  file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        timed[name] = (time.perf_counter() - t0, per)
    wall, per = timed["pass"]
    wall_calls, per_calls = timed["calls"]
    line = {
        "metric": "chem cell-steps/s, surrogate+fallback vs FP64 Rosenbrock, at <=1% rel err",
        "value": ncell * args.steps / wall, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox generator, brams_amd/workload.py)",
        "config": {"workload": "RELACS_TUV RODAS3 cold step on three nested grids: 200x200x40 (dt 120 s), "
                               "98x98x40 (dt 60 s), 50x50x40 (dt 30 s), SZA split, plume, seed 0",
                   "path": "plain FP64 RODAS3 (GPU), the three grids in one pass", "tolerances": args.tol,
                   "cells": ncell, "one_pass_ms_per_step": 1e3 * wall / args.steps,
                   "one_call_per_grid_ms_per_step": 1e3 * wall_calls / args.steps,
                   "kernel_ms_per_grid_one_call_each": [sum(p[g][0] for p in per_calls) / args.steps
                                                        for g in range(3)],
                   "attempts_per_cell_per_grid": [per_calls[-1][g][1] / per_calls[-1][g][2] for g in range(3)]},
        "roofline": None,
    }
    print(json.dumps(line), flush=True)


def calibration_copy(mib, dev, log):
    """A device-to-device copy of known size (mib MiB read and written), the
    reference reading that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    if not mib:
        return
    import torch
    src = torch.ones(mib << 20, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("calibration copy: %d MiB read + %d MiB written" % (mib, mib))
    del src, dst


def bench_plain(args, solver="ros2"):
    """A plain FP64 solver on the headline workload: the ROS2 solver
    (CHEMISTRY=3, brams_chem_ros_dyndt_device), or RODAS3 (CHEMISTRY=4) for
    a mechanism other than RELACS_TUV (whose RODAS3 line is main()'s):
    --spinup steps from the seeded state (the first is the cold step), then
    --steps timed steps, bracketed by synchronisations.  value: cell-steps/s
    of the timed steps."""
    import numpy as np
    import torch
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    from brams_amd.mechanism import load_mechanism
    from brams_amd.workload import Grid, jphoto_from_sza, make_workload, tolerances

    mech = load_mechanism()
    uname = {"ros2": "ROS2", "rodas3": "RODAS3"}[solver]

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    grid = Grid(args.nx, args.ny, args.nz, dtlt=120.0)
    wl = make_workload(grid, seed=0, sza="split", plume_sigma=20.0, plume_factor=50.0)
    atol, rtol = tolerances(args.tol)
    K, S = args.steps, args.spinup
    ds = DeviceStep(wl, atol, rtol, device=0, solver=solver)
    jph = [torch.from_numpy(np.ascontiguousarray(jphoto_from_sza(grid, wl["sza_col"] + 0.5 * t)
                                                 .ravel(order="F"))).to(ds.dev) for t in range(S + K)]
    log("%s %s: workload %dx%dx%d = %d cells, %d blocks" % (mech.name, solver, args.nx, args.ny, args.nz,
                                                            grid.ncell, len(wl["block_end"])))
    calibration_copy(args.copy_calib_mb, ds.dev, log)
    cold = None
    for t in range(S):
        ds.args.jphoto = jph[t].data_ptr()
        ds.step()
        st = abi.stats()
        if t == 0:
            cold = {"kernel_ms": st.kernel_ms, "attempts_per_cell": st.cell_attempts / st.cells,
                    "substeps_per_cell": st.cell_substeps / st.cells,
                    "cell_steps_per_s": grid.ncell / (st.kernel_ms * 1e-3)}
        log("%s spin-up step %d: %.1f ms (%d cell-attempts)" % (solver, t, st.kernel_ms, st.cell_attempts))
    warm_sc = ds.sc_p.clone()
    warm_dt = ds.last_dt.clone()
    f_call, f_sub, f_att = ros2_flop_model() if solver == "ros2" else flop_model()
    wall, main_ms, flops_main, att_main = 0.0, 0.0, 0.0, 0.0
    for t in range(K):
        ds.args.jphoto = jph[S + t].data_ptr()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ds.step()
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        st = abi.stats()
        main_ms += st.main_ms
        share = st.main_cells / st.cells
        flops_main += (f_call * st.cells + f_sub * st.cell_substeps + f_att * st.cell_attempts) * share
        att_main += st.cell_attempts * share
        log("%s step %d: %.1f ms (main kernel %.1f ms, %d cell-attempts)"
            % (solver, t, st.kernel_ms, st.main_ms, st.cell_attempts))
    lean = mech.name != "RELACS_TUV"        # a lean library (traffic measured where listed above)
    if solver == "ros2":
        kernel = "ros2_blocks<64,%d>" % FORM_WPS.get(abi.stats().kernel_form, 0)   # the autotuned form
    else:
        kernel = "rodas3_blocks<64,false,false,%d>" % FORM_WPS.get(abi.tuning()[0], 0)
    achieved = flops_main / (main_ms * 1e-3) / 1e12
    line = {
        "metric": "chem cell-steps/s, FP64 %s (CHEMISTRY=%d) warm rollout" % (uname, 3 if solver == "ros2" else 4),
        "value": grid.ncell * K / wall, "unit": "cell-steps/s", "n_gpus": 1, "steps": K,
        "warmup": S, "ms_per_step": 1e3 * wall / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox generator, brams_amd/workload.py)",
        "config": {"workload": "%s chemistry rollout, %dx%dx%d cells, %d x 120 s %s steps after "
                               "%d %s spin-up steps, SZA split day/night +0.5 deg/step, NOx/VOC plume "
                               "sigma 20 x50, seed 0" % (mech.name, args.nx, args.ny, args.nz, K, uname, S, uname),
                   "mechanism": mech.name,
                   "path": "plain FP64 %s (GPU), warm rollout" % uname, "tolerances": args.tol,
                   "cells": grid.ncell, "blocks": len(wl["block_end"]), "fp64_cold_step": cold,
                   "flop_model_" + solver: {"per_cell_call": f_call, "per_substep": f_sub, "per_attempt": f_att}},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     # spill traffic per cell-attempt, where it was measured
                     "traffic": (BYTES_PER_CELL_ATTEMPT_MECH.get((mech.name, kernel)) * att_main / K
                                 if (mech.name, kernel) in BYTES_PER_CELL_ATTEMPT_MECH
                                 else None if lean else spill_traffic(kernel, att_main / K)),
                     "kernel": kernel, "avg_ms": main_ms / K,
                     "algorithmic_bytes": (2 * mech.nspecies * 4 + 4 * 4 + mech.nr_photo * 8) * grid.ncell},
    }
    if not args.no_cpu:
        warm_wl = dict(wl)
        ds.sc_p.copy_(warm_sc)
        warm_wl["sc_p"] = ds.species_host()
        warm_wl["jphoto"] = jphoto_from_sza(grid, wl["sza_col"] + 0.5 * S)
        warm_wl["last_accepted_dt"] = warm_dt.cpu().numpy().copy()
        log("cpu baseline: %s oracle on block samples of the warm and the cold state" % uname)
        line["cpu_baseline"] = cpu_baseline(warm_wl, wl, atol, rtol, warm_every=4, cold_every=16,
                                            solver=solver)
    print(json.dumps(line), flush=True)


def bench_qssa(args):
    """The QSSA solver (CHEMISTRY=1, brams_chem_qssa_device, itchim 75) on the
    headline grid with every species in ppbm: --steps timed calls with J
    advancing as in the other paths, bracketed by synchronisations; the
    tendencies accumulate in sc_t as the reference's do between dynamics
    steps.  value: cell-steps/s."""
    import numpy as np
    import torch
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    from brams_amd.mechanism import load_mechanism
    from brams_amd.workload import Grid, jphoto_from_sza, make_workload

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    mech = load_mechanism()
    ns = mech.nspecies
    lean = mech.name != "RELACS_TUV"
    qbytes = 3 * ns * 4 + 4 * 4 + mech.nr_photo * 8     # sc_p in, sc_t in and out, met, J
    itchim = 75
    grid = Grid(args.nx, args.ny, args.nz, dtlt=120.0)
    wl = make_workload(grid, seed=0, sza="split", plume_sigma=20.0, plume_factor=50.0, all_ppbm=True)
    K = args.steps
    ds = DeviceStep(wl, np.zeros(ns), np.zeros(ns), device=0)
    jph = [torch.from_numpy(np.ascontiguousarray(jphoto_from_sza(grid, wl["sza_col"] + 0.5 * t)
                                                 .ravel(order="F"))).to(ds.dev) for t in range(K)]
    log("qssa: workload %dx%dx%d = %d cells" % (args.nx, args.ny, args.nz, grid.ncell))
    calibration_copy(args.copy_calib_mb, ds.dev, log)
    for i in range(args.warmup):
        ds.args.jphoto = jph[0].data_ptr()
        ds.qssa(itchim)
    ds.sc_t.zero_()
    torch.cuda.synchronize()
    wall, kms = 0.0, 0.0
    for t in range(K):
        ds.args.jphoto = jph[t].data_ptr()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = ds.qssa(itchim)
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        st = abi.stats()
        kms += st.kernel_ms
        log("qssa step %d: %.3f ms (%d cells at 1e20)" % (t, st.kernel_ms, rc))
    flop = qssa_flop_model(itchim)
    achieved = flop * grid.ncell * K / (kms * 1e-3) / 1e12
    line = {
        "metric": "chem cell-steps/s, QSSA (CHEMISTRY=1, itchim 75)",
        "value": grid.ncell * K / wall, "unit": "cell-steps/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Philox generator, brams_amd/workload.py), every species in ppbm",
        "config": {"workload": "%s chemistry, %dx%dx%d cells, %d QSSA calls of 120 s with 75 "
                               "sub-steps, SZA split day/night +0.5 deg/call, NOx/VOC plume sigma 20 x50, "
                               "seed 0" % (mech.name, args.nx, args.ny, args.nz, K),
                   "mechanism": mech.name,
                   "path": "QSSA (GPU)", "cells": grid.ncell, "itchim": itchim,
                   "flop_per_cell": flop, "algorithmic_bytes_per_cell": qbytes},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     # traffic measured for RELACS_TUV's kernel only
                     "traffic": None if lean else QSSA_TRAFFIC_PER_CELL * grid.ncell,
                     "kernel": "qssa_cells",
                     "avg_ms": kms / K, "algorithmic_bytes": qbytes * grid.ncell,
                     "hbm_frac": qbytes * grid.ncell / (kms / K * 1e-3) / 8e12},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        cores = len(os.sched_getaffinity(0))
        wl1 = dict(wl)
        wl1["jphoto"] = jphoto_from_sza(grid, wl["sza_col"])
        log("cpu baseline: the oracle's QSSA on the whole grid, %d threads" % cores)
        t0 = time.perf_counter()
        hs, orc = oracle_lib.run_oracle_qssa(wl1, itchim, nthreads=cores)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": grid.ncell / dt, "unit": "cell-steps/s", "cores": cores,
                                "kind": "port",
                                "sample": "oracle (C restatement of chem_qssa, gcc -O2 -ffp-contract=off, "
                                          "pthreads over cells) on the whole %d-cell grid, one call "
                                          "(%.1f s wall)" % (grid.ncell, dt)}
    print(json.dumps(line), flush=True)


def bench_tuv(args):
    """The FAST-TUV output mapping and RELACS J relations
    (brams_chem_tuv_jphoto_device) on the headline grid with TUV's own
    sizes (kj = 80 reactions, kz = 125 levels, 17 of them mapped): --steps
    timed calls on resident device arrays.  value: grid points per second;
    the roofline is HBM against the bytes the mapping must move."""
    import numpy as np
    import torch
    from brams_amd import abi
    from brams_amd.chem import tuv_jphoto_device
    from brams_amd.workload import Grid

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    grid = Grid(args.nx, args.ny, args.nz)
    kj, kz = 80, 125
    ni, nj = grid.iz - grid.ia + 1, grid.jz - grid.ja + 1
    xref = np.zeros(kj, np.int32)
    xref[np.arange(2, 2 + 4 * 17, 4)] = np.arange(1, 18)     # 17 TUV reactions, one per rate
    g = torch.Generator(device="cuda").manual_seed(0)
    valj = torch.rand(ni * nj * kj * kz, dtype=torch.float64, device="cuda", generator=g) * 1e-2
    n3d = grid.m1 * grid.m2 * grid.m3
    jph = torch.rand(17 * n3d, dtype=torch.float64, device="cuda", generator=g) * 1e-3
    calibration_copy(args.copy_calib_mb, torch.device("cuda", 0), log)
    for i in range(args.warmup):
        tuv_jphoto_device(grid, valj, kj, kz, xref, jph)
    torch.cuda.synchronize()
    wall, kms = 0.0, 0.0
    for t in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tuv_jphoto_device(grid, valj, kj, kz, xref, jph)
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        kms += abi.stats().kernel_ms
    K = args.steps
    npts = grid.m1 * ni * nj
    # mapped slices read (levels 2..m1 from valj), jphoto written at every
    # interior point, J14/J16 read-modified-written elsewhere (3 read + 2 written)
    nbound = n3d - npts
    algo = 17 * ni * nj * (grid.m1 - 1) * 8 + 17 * npts * 8 + nbound * 5 * 8
    achieved = algo / (kms / K * 1e-3) / 1e9
    line = {
        "metric": "photolysis mapping grid points/s (FAST-TUV output -> jphoto + RELACS J relations)",
        "value": npts * K / wall, "unit": "points/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (torch.rand j-values)",
        "config": {"workload": "TUV j-values of a %dx%dx%d grid, kj 80, kz 125, 17 reactions mapped"
                               % (args.nx, args.ny, args.nz), "points": npts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": TUV_TRAFFIC_PER_POINT * npts,
                     "kernel": "tuv_map_interior",
                     "avg_ms": kms / K, "algorithmic_bytes": algo},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spinup", type=int, default=6)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--tol", default="baseline", choices=["baseline", "reference"])
    ap.add_argument("--path", default="surrogate", choices=["surrogate", "fp64", "nested", "ros2", "qssa", "tuv"],
                    help="nested: the plain FP64 step on three nested grids (200x200x40 at dt 120 s, "
                         "98x98x40 at 60 s, 50x50x40 at 30 s), one call per grid and in one pass; "
                         "ros2: the ROS2 solver (CHEMISTRY=3) rollout on the 200x200x40 workload; "
                         "qssa: the QSSA solver (CHEMISTRY=1) on it; tuv: the FAST-TUV j-value "
                         "mapping into jphoto")
    ap.add_argument("--model", default=os.path.join(ROOT, "brams_amd", "models", "relacs_tuv_surrogate.npz"))
    ap.add_argument("--spread-tol", type=float, default=None, help="override the model's spread check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="first copy this many MiB device to device (a kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the chemistry step has no cross-GPU exchange "
              "on one grid; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from brams_amd.mechanism import default_name
    if default_name() != "relacs_tuv":
        # another mechanism (BRAMS_MECH): the solvers' plain FP64 paths; the
        # surrogate model and the TUV line are RELACS_TUV's
        if args.path in ("surrogate", "tuv"):
            print("bench.py --path %s measures RELACS_TUV; with BRAMS_MECH=%s use --path fp64, ros2, qssa "
                  "or nested" % (args.path, default_name()), file=sys.stderr)
            sys.exit(2)
        if args.path == "fp64":
            return bench_plain(args, "rodas3")
    if args.path == "nested":
        return bench_nested(args)
    if args.path == "ros2":
        return bench_plain(args, "ros2")
    if args.path == "qssa":
        return bench_qssa(args)
    if args.path == "tuv":
        return bench_tuv(args)

    import numpy as np
    import torch
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    from brams_amd.surrogate import (NIN, SurrogateModel, SurrogateStep, physical_cells,
                                     species_errors, trusted_cells)
    from brams_amd.workload import Grid, jphoto_from_sza, make_workload, tolerances

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    grid = Grid(args.nx, args.ny, args.nz, dtlt=120.0)
    wl = make_workload(grid, seed=0, sza="split", plume_sigma=20.0, plume_factor=50.0)
    atol, rtol = tolerances(args.tol)
    K, S = args.steps, args.spinup
    ref = DeviceStep(wl, atol, rtol, device=0)
    dev = ref.dev
    jph = [torch.from_numpy(np.ascontiguousarray(jphoto_from_sza(grid, wl["sza_col"] + 0.5 * t)
                                                 .ravel(order="F"))).to(dev) for t in range(S + K)]
    fbuf = torch.empty((grid.ncell, NIN), dtype=torch.float32, device=dev)
    valid = torch.ones(grid.ncell, dtype=torch.bool, device=dev)
    log("workload %dx%dx%d = %d cells, %d blocks" % (args.nx, args.ny, args.nz, grid.ncell, len(wl["block_end"])))
    calibration_copy(args.copy_calib_mb, dev, log)

    def block_spread(tag):
        att, _, _, _ = abi.block_status(len(wl["block_end"]))
        q = np.quantile(att, [0.5, 0.9, 0.99])
        log("%s: attempts per block mean %.1f, p50 %.0f, p90 %.0f, p99 %.0f, max %d"
            % (tag, att.mean(), q[0], q[1], q[2], att.max()))

    # ---- spin-up (untimed); step 0 is the cold step out of the seeded state
    cold = None
    for t in range(S):
        ref.args.jphoto = jph[t].data_ptr()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref.step()
        torch.cuda.synchronize()
        st = abi.stats()
        if t == 0:
            # the first call of a grid also times the kernel forms (autotuner):
            # the cold step's rate is taken over its own kernels' event time
            form, tune_ms = abi.tuning()
            cold = {"kernel_ms": st.kernel_ms, "call_ms_with_autotune": 1e3 * (time.perf_counter() - t0),
                    "attempts_per_cell": st.cell_attempts / st.cells,
                    "substeps_per_cell": st.cell_substeps / st.cells,
                    "cell_steps_per_s": grid.ncell / (st.kernel_ms * 1e-3),
                    "autotune": {"form": form, "kernel": "rodas3_blocks<64,false,false,%d>" % FORM_WPS.get(form, 0),
                                 "ms_per_form": {"wps%d" % FORM_WPS[f]: float(tune_ms[f]) for f in FORM_WPS}}}
        valid &= trusted_cells(wl["block_end"], dev) & physical_cells(ref, fbuf)
        log("spin-up step %d: %.1f ms (%d cell-attempts)" % (t, st.kernel_ms, st.cell_attempts))
        if t == 0:
            block_spread("cold step")
    warm_sc = ref.sc_p.clone()
    warm_dt = ref.last_dt.clone()

    sur = None
    if args.path == "surrogate":
        model = SurrogateModel.load(args.model)
        if args.spread_tol is not None:
            model.spread_tol = args.spread_tol
        sds = DeviceStep(wl, atol, rtol, device=0)
        sur = SurrogateStep(model, sds)
        for i in range(args.warmup):
            sds.sc_p.copy_(warm_sc)
            sds.last_dt.copy_(warm_dt)
            sds.args.jphoto = jph[S].data_ptr()
            sur.step()
            log("warmup %d done" % i)
        sds.sc_p.copy_(warm_sc)
        sds.last_dt.copy_(warm_dt)
        sur.reset()
        snaps = torch.empty((K,) + tuple(sds.sc_p.shape), dtype=sds.sc_p.dtype, device=dev)
        sstats = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(K):
            sds.args.jphoto = jph[S + t].data_ptr()
            s = sur.step()
            snaps[t].copy_(sds.sc_p)
            sstats.append(s)
            log("surrogate step %d: %.1f ms (mlp %.2f, fallback %.1f ms, %d fallback cells, %d cell-attempts)"
                % (t, s.total_ms, s.mlp_ms, s.fallback_ms, s.fallback_cells, s.fallback_cell_attempts))
        torch.cuda.synchronize()
        wall_a = time.perf_counter() - t0

    # ---- leg B: FP64 rollout from the same warm state
    ref.sc_p.copy_(warm_sc)
    ref.last_dt.copy_(warm_dt)
    wall_b, b_main_ms, b_flops_main, b_kernel_ms, b_att_main = 0.0, 0.0, 0.0, 0.0, 0.0
    worst = {"median": 0.0, "p99": 0.0}
    per_step = []
    for t in range(K):
        ref.args.jphoto = jph[S + t].data_ptr()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref.step()
        torch.cuda.synchronize()
        wall_b += time.perf_counter() - t0
        st = abi.stats()
        b_main_ms += st.main_ms
        b_kernel_ms += st.kernel_ms
        b_flops_main += rodas3_flops(st.cells, st.cell_substeps, st.cell_attempts) * st.main_cells / st.cells
        b_att_main += st.cell_attempts * st.main_cells / st.cells
        valid &= trusted_cells(wl["block_end"], dev) & physical_cells(ref, fbuf)
        if sur is not None:
            err = species_errors(snaps[t], ref.sc_p, grid, valid=valid)
            med = max(err.items(), key=lambda kv: kv[1][0])
            p99 = max(err.items(), key=lambda kv: kv[1][1])
            worst["median"] = max(worst["median"], med[1][0])
            worst["p99"] = max(worst["p99"], p99[1][1])
            per_step.append([round(med[1][0], 5), med[0], round(p99[1][1], 5), p99[0]])
        log("fp64 step %d: %.1f ms (main kernel %.1f ms, %d cell-attempts)"
            % (t, st.kernel_ms, st.main_ms, st.cell_attempts))
        if t == K - 1:
            block_spread("last fp64 step")
    fp64_value = grid.ncell * K / wall_b

    pk = "rodas3_blocks<64,false,false,%d>" % FORM_WPS.get(abi.stats().kernel_form, 0)
    plain = {"kernel": pk, "bound": "fp64", "unit": "TFLOP/s",
             "avg_ms": b_main_ms / K, "achieved": b_flops_main / (b_main_ms * 1e-3) / 1e12,
             "peak": FP64_PEAK_TFLOPS, "traffic": spill_traffic(pk, b_att_main / K),
             "algorithmic_bytes": 528 * grid.ncell}
    plain["frac"] = plain["achieved"] / plain["peak"]
    if sur is None:
        value, wall, path = fp64_value, wall_b, "plain FP64 RODAS3 (GPU), warm rollout"
        dominant, others = plain, []
        extra = {}
    else:
        value, wall, path = grid.ncell * K / wall_a, wall_a, "surrogate+fallback (GPU), warm rollout"
        fb_ms = sum(s.fallback_ms for s in sstats)
        mlp_ms = sum(s.mlp_ms for s in sstats)
        fb_flops = sum(rodas3_flops(s.fallback_cells, s.fallback_cell_substeps, s.fallback_cell_attempts)
                       for s in sstats)
        fb_att = sum(s.fallback_cell_attempts for s in sstats)
        fk = "rodas3_blocks<64,false,true,%d>" % sstats[-1].fallback_wps
        fallback = {"kernel": fk, "bound": "fp64", "unit": "TFLOP/s",
                    "avg_ms": fb_ms / K, "achieved": fb_flops / max(fb_ms * 1e-3, 1e-9) / 1e12,
                    "peak": FP64_PEAK_TFLOPS, "traffic": spill_traffic(fk, fb_att / K),
                    "algorithmic_bytes": 528 * sum(s.fallback_cells for s in sstats) / K}
        fallback["frac"] = fallback["achieved"] / fallback["peak"]
        mlp = {"kernel": "mlp_infer<256,4,8,2>", "bound": "mfma", "unit": "TFLOP/s",
               "avg_ms": mlp_ms / K, "traffic": MLP_BYTES_PER_CELL * grid.ncell,
               "algorithmic_bytes": (256 + 3 * 256) * grid.ncell,
               "achieved": MLP_FLOP_PER_CELL * grid.ncell * K / (mlp_ms * 1e-3) / 1e12,
               "peak": BF16_PEAK_TFLOPS}
        mlp["frac"] = mlp["achieved"] / mlp["peak"]
        if fb_ms >= mlp_ms:
            dominant, others = fallback, [mlp, plain]
        else:
            dominant, others = mlp, [fallback, plain]
        final = species_errors(snaps[K - 1], ref.sc_p, grid, valid=valid)
        final_all = species_errors(snaps[K - 1], ref.sc_p, grid, valid=None)
        med_w = max(v[0] for v in final.values())
        p99_w = max(v[1] for v in final.values())
        extra = {
            "accuracy_condition": "worst-species median <= 0.01 and p99 <= 0.05 after the last step",
            "accuracy_met": bool(med_w <= 0.01 and p99_w <= 0.05),
            "fp64_cell_steps_per_s": fp64_value,
            "speedup_vs_fp64_rollout": value / fp64_value,
            "speedup_vs_fp64_cold_step": value / cold["cell_steps_per_s"],
            "fallback_frac_mean": float(np.mean([s.fallback_cells / s.cells for s in sstats])),
            "fallback_frac_last": sstats[-1].fallback_cells / sstats[-1].cells,
            "fallback_reasons_total": dict(zip(["nonfinite", "n_budget", "c_budget", "spread"],
                                               [int(sum(s.reasons[k] for s in sstats)) for k in range(4)])),
            "ms_split": {"features": sum(s.features_ms for s in sstats) / K, "mlp": mlp_ms / K,
                         "apply": sum(s.apply_ms for s in sstats) / K, "fallback": fb_ms / K},
            "error_final_worst_median": med_w,
            "error_final_worst_p99": p99_w,
            "error_final_worst_median_all_cells": max(v[0] for v in final_all.values()),
            "error_final_worst_p99_all_cells": max(v[1] for v in final_all.values()),
            "error_worst_median_over_steps": worst["median"],
            "error_worst_p99_over_steps": worst["p99"],
            "error_per_step": per_step,
            "error_untrusted_fp64_frac": 1.0 - valid.float().mean().item(),
            "spread_tol": model.spread_tol, "n_tol": model.n_tol, "c_tol": model.c_tol,
        }
    roofline = {"bound": dominant["bound"], "achieved": dominant["achieved"], "peak": dominant["peak"],
                "unit": dominant["unit"], "frac": dominant["frac"], "traffic": dominant["traffic"],
                "algorithmic_bytes": dominant["algorithmic_bytes"],
                "traffic_note": "HBM bytes per launch: PMC-measured bytes per cell-attempt (RODAS3) or per "
                                "cell (MLP) x this run's counts, profiles/rodas3_hbm_traffic_pmc.md; "
                                "null where that kernel form was not measured",
                "kernel": dominant["kernel"], "avg_ms": dominant["avg_ms"],
                "peak_note": "FP64 vector for the RODAS3 kernels (no MFMA issued; MI355X FP64 vector "
                             "and matrix peaks are both 78.6 TFLOP/s), dense BF16 MFMA for mlp_infer",
                "other_kernels": others}
    f_call, f_sub, f_att = flop_model()
    line = {
        "metric": "chem cell-steps/s, surrogate+fallback vs FP64 Rosenbrock, at <=1% rel err",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if sur is None else "bf16+f64",
        "data": "synthetic (seeded Philox generator, brams_amd/workload.py)",
        "config": dict({"workload": "RELACS_TUV chemistry rollout, %dx%dx%d cells, %d x 120 s steps after "
                                    "%d FP64 spin-up steps, SZA split day/night +0.5 deg/step, NOx/VOC "
                                    "plume sigma 20 x50, seed 0" % (args.nx, args.ny, args.nz, K, S),
                        "path": path, "tolerances": args.tol, "cells": grid.ncell,
                        "blocks": len(wl["block_end"]), "spinup_steps": S,
                        "fp64_cold_step": cold,
                        "flop_model_rodas3": {"per_cell_call": f_call, "per_substep": f_sub,
                                              "per_attempt": f_att},
                        "mlp_flop_per_cell": MLP_FLOP_PER_CELL}, **extra),
        "roofline": roofline,
    }
    if not args.no_cpu:
        warm_wl = dict(wl)
        ref.sc_p.copy_(warm_sc)
        warm_wl["sc_p"] = ref.species_host()
        warm_wl["jphoto"] = jphoto_from_sza(grid, wl["sza_col"] + 0.5 * S)
        warm_wl["last_accepted_dt"] = warm_dt.cpu().numpy().copy()
        log("cpu baseline: oracle on block samples of the warm and the cold state")
        line["cpu_baseline"] = cpu_baseline(warm_wl, wl, atol, rtol)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
