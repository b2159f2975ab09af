"""GPU checks of the surrogate chemistry step (brams_sur_step) end to end.

* every cell accepted (checks disabled): the written state equals a numpy
  restatement of the step -- features from the Fortran arrays, z-scores,
  bf16 rounding, the CPU MLP oracle (oracle/mlp_oracle.py) for every member,
  ensemble mean, 10**(log10 c + increment), unit conversion back to ppbm;
* every cell rejected (spread tolerance 0): the cells are integrated by the
  RODAS3 fallback, which must agree with the oracle's FP64 step to the
  solver tolerance (the fallback uses the full structural Jacobian, the
  oracle the value-based mask of cell 1, so agreement is to rtol, not bits).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import mlp_oracle as MO  # noqa: E402

pytestmark = pytest.mark.gpu

NH, NL = 64, 2


def _model(wl, seed, targ_sd):
    """A small untrained ensemble with normalisation fitted to the workload."""
    from brams_amd.mlp import MLP
    from brams_amd.surrogate import SurrogateModel
    feats = _features(wl)
    m = SurrogateModel(NH, NL, nmembers=3)
    m.feat_mu = feats.mean(0).astype(np.float32)
    sd = feats.std(0)
    m.feat_sd = np.where(sd > 1e-12, sd, 1.0).astype(np.float32)
    m.targ_mu = np.zeros(MO.NOUT, np.float32)
    m.targ_sd = np.full(MO.NOUT, targ_sd, np.float32)
    m.constant = np.zeros(MO.NOUT, bool)
    m.constant[[12, 13, 15]] = True          # CO2, N2, H2O never change
    for k in range(3):
        net = MLP(NH, NL, max_batch=64, seed=seed + k)
        m.params[k] = net.get_params()
        net.close()
    return m


def _cells(wl):
    """Per interior cell (j -> i -> k order): T, p, volmol, concentrations."""
    from brams_amd.mechanism import load_mechanism
    g = wl["grid"]
    sl = (slice(1, g.m1 - 1), slice(g.ia - 1, g.iz), slice(g.ja - 1, g.jz))

    def flat(a):
        return np.asarray(a)[sl].transpose(2, 1, 0).ravel()   # j outer, k inner

    cp, cpor, p00 = np.float32(1004.0), 1004.0 / 287.0, 1.0e5
    exner = ((flat(wl["pp"]) + flat(wl["pi0"])) / cp).astype(np.float32)
    press = exner.astype(np.float64)
    temp = flat(wl["theta"]).astype(np.float64) * press
    press = press ** np.float64(np.float32(cpor)) * p00
    volmol = (6.02e23 * 1e-15 * 28.96) * press / (8.314 * temp)
    mech = load_mechanism()
    conc = []
    for s, sp in enumerate(mech.species):
        v = flat(wl["sc_p"][s]).astype(np.float64)
        if sp.transported:
            v = np.maximum(v * volmol / np.float64(np.float32(sp.weight)), 0.0)
        conc.append(v)
    jph = np.asarray(wl["jphoto"])[:, 1:g.m1 - 1, g.ia - 1:g.iz, g.ja - 1:g.jz]
    jph = jph.transpose(0, 3, 2, 1).reshape(jph.shape[0], -1)
    rv = flat(wl["rv"]).astype(np.float64)
    return temp, press, volmol, np.stack(conc, 1), jph.T, rv


def _features(wl):
    from brams_amd.surrogate import FLOOR_C
    temp, press, _, conc, jph, rv = _cells(wl)
    g = wl["grid"]
    f = np.zeros((conc.shape[0], MO.NIN), np.float64)
    f[:, :47] = np.log10(np.maximum(conc, FLOOR_C))
    f[:, 47] = temp
    f[:, 48] = np.log(press)
    f[:, 49] = np.log10(np.maximum(rv, 1e-12))
    f[:, 50:67] = jph
    f[:, 67] = np.log10(np.float32(g.dtlt))
    return f.astype(np.float32)


def _setup(targ_sd, spread_tol, atol_kind="reference"):
    from brams_amd.chem import DeviceStep
    from brams_amd.surrogate import SurrogateStep
    from brams_amd.workload import Grid, make_workload, tolerances
    grid = Grid(16, 16, 4)
    wl = make_workload(grid, seed=11, sza="split")
    atol, rtol = tolerances(atol_kind)
    model = _model(wl, 5, targ_sd)
    model.spread_tol = spread_tol
    model.n_tol = model.c_tol = 1e30
    ds = DeviceStep(wl, atol, rtol)
    return wl, model, ds, SurrogateStep(model, ds), atol, rtol


def test_accepted_cells_match_numpy_pipeline():
    from brams_amd.mechanism import load_mechanism
    from brams_amd.surrogate import FLOOR_C
    print("case sur_step all accepted, 16x16x4 seed 11", flush=True)
    wl, model, ds, sur, _, _ = _setup(targ_sd=1e-3, spread_tol=1e30)
    s = sur.step()
    assert s.fallback_cells == 0 and s.cells == wl["grid"].ncell
    # numpy restatement
    feats = _features(wl)
    X = MO.bf16((feats - model.feat_mu) / model.feat_sd).astype(np.float64)
    Xp = np.zeros((X.shape[0], MO.NIN_PAD))
    Xp[:, :MO.NIN] = X
    incs = []
    for p in model.params:
        out, _ = MO.forward(MO.unpack(p.astype(np.float64), NH, NL), Xp)
        incs.append(out[:, :MO.NOUT] * model.targ_sd + model.targ_mu)
    mean = np.mean(incs, 0)
    _, _, volmol, conc, _, _ = _cells(wl)
    l0 = np.log10(np.maximum(conc, FLOOR_C))
    y1 = np.where(model.constant[None, :], conc, 10.0 ** (l0 + mean))
    mech = load_mechanism()
    g = wl["grid"]
    got_all = ds.species_host()
    sl = (slice(1, g.m1 - 1), slice(g.ia - 1, g.iz), slice(g.ja - 1, g.jz))
    worst = 0.0
    for si, sp in enumerate(mech.species):
        got = got_all[si][sl].transpose(2, 1, 0).ravel().astype(np.float64)
        exp = y1[:, si] * np.float64(np.float32(sp.weight)) / volmol if sp.transported else y1[:, si]
        exp = np.maximum(exp.astype(np.float32), 0.0).astype(np.float64)
        rel = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-30)
        worst = max(worst, float(rel.max()))
    print("worst rel %.3g" % worst)
    # bf16 products differ from the FP64 oracle by ~1e-2 of the normalised
    # output; times targ_sd 1e-3 decades that is < 1e-4 relative
    assert worst < 2e-4


def test_rejected_cells_follow_fp64_rosenbrock():
    import oracle_lib as O
    print("case sur_step all rejected, 16x16x4 seed 11, after 3 FP64 steps", flush=True)
    wl, model, ds, sur, atol, rtol = _setup(targ_sd=1.0, spread_tol=0.0)
    # a warm state: the first steps out of the synthetic state end in
    # acceptances forced at dt_min, where two valid integrations may differ
    for _ in range(3):
        ds.step()
    warm = dict(wl)
    warm["sc_p"] = ds.species_host()
    warm["last_accepted_dt"] = ds.last_dt.cpu().numpy().copy()
    s = sur.step()
    assert s.fallback_cells == s.cells == wl["grid"].ncell
    assert s.reasons[3] + s.reasons[0] == s.cells
    hs, st, _, rc = O.run_oracle(warm, atol, rtol)
    print("oracle rc %d, forced blocks %d" % (rc, st.forced_blocks))
    assert rc == 0
    g = wl["grid"]
    sl = (slice(1, g.m1 - 1), slice(g.ia - 1, g.iz), slice(g.ja - 1, g.jz))
    got = np.stack([a[sl] for a in ds.species_host()]).astype(np.float64)
    ref = np.stack([np.asarray(a)[sl] for a in hs.sc_p]).astype(np.float64)
    floor = 1e-6 * np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    rel = np.abs(got - ref) / (np.abs(ref) + floor)
    print("fallback vs oracle: median %.3g, p99 %.3g, max %.3g"
          % (np.median(rel), np.quantile(rel, 0.99), rel.max()))
    # both are RODAS3 at rtol 1e-3; they differ by the Jacobian pattern
    # (full structural vs cell 1's value-based mask)
    assert np.median(rel) < 1e-3
    assert np.quantile(rel, 0.99) < 5e-2


def test_fallback_is_deterministic():
    """The fallback (with its per-cell step memory) gives bit-identical
    results when the same state is integrated twice after a reset."""
    print("case sur_step fallback twice, 16x16x4 seed 11", flush=True)
    wl, model, ds, sur, atol, rtol = _setup(targ_sd=1.0, spread_tol=0.0)
    for _ in range(2):
        ds.step()
    start = ds.sc_p.clone()
    start_dt = ds.last_dt.clone()
    out = []
    for _ in range(2):
        ds.sc_p.copy_(start)
        ds.last_dt.copy_(start_dt)
        sur.reset()
        sur.step()                                      # fills the per-cell memory
        sur.step()                                      # starts from that memory
        out.append(ds.sc_p.cpu().numpy().copy())
    assert np.array_equal(out[0], out[1])


def test_fallback_failure_is_reported():
    """A cell whose temperature is not finite fails the surrogate's checks
    (non-finite result), goes to the RODAS3 fallback, and the fallback block
    ends at once on its non-finite error norm (the bounded failure path of
    the kernel).  brams_sur_step must report it: fallback_failed_blocks and
    a positive return value, not a silent success."""
    import torch
    print("case sur_step fallback failure, 16x16x4 seed 11, one NaN temperature", flush=True)
    wl, model, ds, sur, atol, rtol = _setup(targ_sd=1.0, spread_tol=1e30)
    g = wl["grid"]
    o = 2 + g.m1 * (g.ia + 3) + g.m1 * g.m2 * (g.ja + 2)      # an interior (k,i,j)
    ds.theta[o] = float("nan")
    torch.cuda.synchronize()
    s = sur.step()
    print("fallback cells %d, failed blocks %d" % (s.fallback_cells, s.fallback_failed_blocks))
    assert s.fallback_cells >= 1
    assert s.reasons[0] >= 1                       # flagged as non-finite
    assert s.fallback_failed_blocks >= 1
    from brams_amd.surrogate import _bind
    import ctypes as C
    s2 = type(s)()
    ds.sc_p.copy_(ds.sc_p0)
    rc = _bind().brams_sur_step(C.byref(ds.args), C.byref(sur.cmodel), C.byref(s2))
    assert rc == s2.fallback_failed_blocks >= 1


def test_grids_called_in_turn_keep_their_fallback_memory():
    """Two grids called alternately, as nested grids are (one call per
    grid), keep their own device state: grid A's second surrogate step,
    whose fallback starts from the per-cell step memory of its first, is
    bit-identical whether or not grid B ran in between."""
    from brams_amd.chem import DeviceStep
    from brams_amd.surrogate import SurrogateStep
    from brams_amd.workload import Grid, make_workload, tolerances
    print("case two grids in turn, 16x16x4 and 12x12x4 (dt 60 s), all cells to the fallback", flush=True)
    atol, rtol = tolerances("reference")
    wa = make_workload(Grid(16, 16, 4), seed=11, sza="split")
    wb = make_workload(Grid(12, 12, 4, dtlt=60.0), seed=12, sza="split")
    ma, mb = _model(wa, 5, 1.0), _model(wb, 6, 1.0)
    ma.spread_tol = mb.spread_tol = 0.0
    ma.n_tol = ma.c_tol = mb.n_tol = mb.c_tol = 1e30
    da, db = DeviceStep(wa, atol, rtol), DeviceStep(wb, atol, rtol)
    sa, sb = SurrogateStep(ma, da), SurrogateStep(mb, db)
    sa.reset()
    sa.step()
    sa.step()
    alone = da.sc_p.cpu().numpy().copy()
    da.reset()
    db.reset()
    sa.reset()
    sa.step()
    sb.step()
    sa.step()
    assert np.array_equal(da.sc_p.cpu().numpy(), alone)
