#!/usr/bin/env python3
"""Train the chemistry surrogate on FP64 RODAS3 rollouts and evaluate it.

  python -m brams_amd.train_surrogate --out brams_amd/models/relacs_tuv_surrogate.npz

1. dataset: 200x200x40 workloads (seeds --seeds), --steps chemistry steps each
   with SZA advancing 0.5 deg per step, integrated by the GPU FP64 RODAS3
   kernels; a random --fraction of the cells of every step is kept;
2. --members networks (4 x 256 GELU) trained by the library's BF16 kernels;
3. the nitrogen/carbon budget thresholds of the checks are calibrated on the
   FP64 data (the larger of 0.1% and the 99.9th percentile of the FP64
   relative budget change);
4. the model is saved (.npz) and, with --eval-steps > 0, rolled out on seed 0
   next to the FP64 solver, reporting per-species median/p99 relative error,
   fallback fraction and timings.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def log(msg):
    print(msg, flush=True)


def calibrate_budgets(model, feat, targ, nsample=400000):
    """Thresholds of the N and C budget checks from FP64 cell-steps."""
    import torch
    from brams_amd.surrogate import atom_vectors
    n = feat.shape[0]
    idx = torch.randperm(n, device=feat.device)[:min(n, nsample)]
    f = feat.index_select(0, idx)[:, :47].double().cpu().numpy()
    t = targ.index_select(0, idx).double().cpu().numpy()
    c0 = 10.0 ** f
    c1 = 10.0 ** (f + t)
    na, ca = atom_vectors()
    out = []
    for atoms in (na, ca):
        b0 = c0 @ atoms
        b1 = c1 @ atoms
        rel = np.abs(b1 - b0) / np.maximum(b0, 1e-30)
        out.append(max(1e-3, float(np.quantile(rel, 0.999))))
    model.n_tol, model.c_tol = out
    return out


def rollout(model, grid, steps, atol, rtol, seed=0, sza_step=0.5, spinup=0):
    """Surrogate+fallback and FP64 rollouts from the same initial state: the
    seeded workload after `spinup` FP64 steps (untimed, shared by both)."""
    import torch
    from brams_amd.chem import DeviceStep
    from brams_amd.surrogate import NIN, SurrogateStep, physical_cells, species_errors, trusted_cells
    from brams_amd.workload import jphoto_from_sza, make_workload
    wl = make_workload(grid, seed=seed, sza="split", plume_sigma=20.0, plume_factor=50.0)
    ref = DeviceStep(wl, atol, rtol)
    sur_state = DeviceStep(wl, atol, rtol)
    sur = SurrogateStep(model, sur_state)
    jph = [torch.from_numpy(np.ascontiguousarray(jphoto_from_sza(grid, wl["sza_col"] + sza_step * t)
                                                 .ravel(order="F"))).to(ref.dev) for t in range(spinup + steps)]
    for t in range(spinup):
        ref.args.jphoto = jph[t].data_ptr()
        ref.step()
    sur_state.sc_p.copy_(ref.sc_p)
    sur_state.last_dt.copy_(ref.last_dt)
    jph = jph[spinup:]
    torch.cuda.synchronize()
    report = {"steps": [], "spinup": spinup}
    t_ref = t_sur = 0.0
    valid = torch.ones(grid.ncell, dtype=torch.bool, device=ref.dev)
    fbuf = torch.empty((grid.ncell, NIN), dtype=torch.float32, device=ref.dev)
    if spinup:
        valid &= trusted_cells(wl["block_end"], ref.dev) & physical_cells(ref, fbuf)
    for t in range(steps):
        ref.args.jphoto = jph[t].data_ptr()
        sur_state.args.jphoto = jph[t].data_ptr()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        valid &= trusted_cells(wl["block_end"], ref.dev) & physical_cells(ref, fbuf)
        s = sur.step()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_ref += t1 - t0
        t_sur += t2 - t1
        err = species_errors(sur_state.sc_p, ref.sc_p, grid, valid=valid)
        worst_med = max(v[0] for v in err.values())
        worst_p99 = max(v[1] for v in err.values())
        d = {"step": t, "ref_s": t1 - t0, "sur_s": t2 - t1,
             "fallback_frac": s.fallback_cells / s.cells, "sur": s.as_dict(),
             "untrusted_fp64_frac": 1.0 - valid.float().mean().item(),
             "worst_median": worst_med, "worst_p99": worst_p99,
             "worst_median_species": max(err, key=lambda k: err[k][0]),
             "worst_p99_species": max(err, key=lambda k: err[k][1])}
        report["steps"].append(d)
        log("rollout step %d: fp64 %.3f s, surrogate %.3f s (fallback %.4f), worst median %.3g (%s), "
            "worst p99 %.3g (%s)" % (t, d["ref_s"], d["sur_s"], d["fallback_frac"], worst_med,
                                     d["worst_median_species"], worst_p99, d["worst_p99_species"]))
    report["final_errors"] = err
    report["fp64_cell_steps_per_s"] = grid.ncell * steps / t_ref
    report["surrogate_cell_steps_per_s"] = grid.ncell * steps / t_sur
    report["speedup"] = t_ref / t_sur
    return report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", default="1,2,3,4,5,6,7,8,9,10")
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--fraction", type=float, default=0.10)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--members", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(ROOT, "brams_amd", "models", "relacs_tuv_surrogate.npz"))
    ap.add_argument("--skip", type=int, default=2)
    ap.add_argument("--spinup", type=int, default=6)
    ap.add_argument("--eval-steps", type=int, default=24)
    ap.add_argument("--report", default=None)
    ap.add_argument("--finetune-from", default=None,
                    help="fine-tune this model on its FP64 dataset plus cell-steps from its own rollouts")
    ap.add_argument("--rollout-steps", type=int, default=24)
    ap.add_argument("--rollout-fraction", type=float, default=0.05)
    ap.add_argument("--finetune-epochs", type=int, default=50)
    ap.add_argument("--finetune-lr", type=float, default=3e-4)
    args = ap.parse_args()

    import torch
    from brams_amd.surrogate import make_dataset, train_ensemble
    from brams_amd.workload import Grid, tolerances
    grid = Grid(args.nx, args.ny, args.nz, dtlt=120.0)
    atol, rtol = tolerances("baseline")
    seeds = [int(s) for s in args.seeds.split(",") if s]
    t0 = time.time()
    feat, targ = make_dataset(grid, seeds, args.steps, args.fraction, atol, rtol, log=log, skip=args.skip)
    log("dataset: %d cell-steps in %.1f s" % (feat.shape[0], time.time() - t0))
    if args.finetune_from:
        from brams_amd.surrogate import SurrogateModel, finetune_ensemble, make_rollout_dataset
        model = SurrogateModel.load(args.finetune_from)
        t1 = time.time()
        rf, rt = make_rollout_dataset(model, grid, seeds, args.rollout_steps, args.rollout_fraction, atol, rtol,
                                      spinup=args.spinup, log=log)
        log("rollout dataset: %d cell-steps in %.1f s" % (rf.shape[0], time.time() - t1))
        feat = torch.cat([feat, rf])
        targ = torch.cat([targ, rt])
        del rf, rt
        model = finetune_ensemble(model, feat, targ, epochs=args.finetune_epochs, batch=args.batch,
                                  lr0=args.finetune_lr, log=log)
    else:
        model = train_ensemble(feat, targ, nmembers=args.members, epochs=args.epochs, batch=args.batch,
                               lr0=args.lr, log=log)
    ntol, ctol = calibrate_budgets(model, feat, targ)
    log("checks: N budget tol %.3g, C budget tol %.3g, spread tol %.3g" % (ntol, ctol, model.spread_tol))
    model.meta.update({"seeds": seeds, "steps": args.steps, "skip": args.skip, "fraction": args.fraction,
                       "grid": [args.nx, args.ny, args.nz]})
    del feat, targ
    torch.cuda.empty_cache()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    model.save(args.out)
    log("saved %s" % args.out)
    if args.eval_steps > 0:
        rep = rollout(model, grid, args.eval_steps, atol, rtol, spinup=args.spinup)
        rep["model"] = model.meta
        log("rollout: fp64 %.4g cell-steps/s, surrogate %.4g cell-steps/s, speedup %.2f"
            % (rep["fp64_cell_steps_per_s"], rep["surrogate_cell_steps_per_s"], rep["speedup"]))
        if args.report:
            with open(args.report, "w") as fh:
                json.dump(rep, fh, indent=1)


if __name__ == "__main__":
    main()
