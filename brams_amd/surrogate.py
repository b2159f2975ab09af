"""Learned surrogate of the chemistry step, with FP64 RODAS3 fallback.

No counterpart in the reference.  The pieces:

  make_dataset  cell-steps of FP64 RODAS3 rollouts on the GPU (the plain
                kernels of this package) -> raw features / log10 increments
  fit_normaliser column statistics of the dataset (device kernels)
  train_ensemble members trained by the library's BF16 MFMA forward/backward
                and Adam kernels
  SurrogateModel parameters + normalisation, saved/loaded as .npz
  SurrogateStep  brams_sur_step on device-resident state: features, ensemble
                inference, checks, write-back, RODAS3 on rejected cells

torch tensors are used only as device memory; all arithmetic on the device
runs in libbrams_chem.so.
"""
import ctypes as C
import os
import time

import numpy as np

from . import abi
from .mechanism import load_mechanism
from .mlp import MLP, NIN, NIN_PAD, NOUT, NOUT_PAD

FLOOR_C = 1.0e-5          # molec/cm^3 floor of the log10 features / targets (below the
                          # 1e-12 ppb error floor, 6e-3 molec/cm^3 and up)
# Blocks whose step accepted an error above this at dt_min (the reference's
# forced acceptance, chem_spack_rodas3_dyndt.f90:671-676) or that failed are
# not trustworthy FP64 results: their cells are left out of datasets and
# error statistics from then on.  Forced acceptances with errors of 1e2-1e5
# are routine in the first step from an unbalanced initial state at
# atol = 1e-12 ppb and give physical states; the blocks that end with values
# of 1e20-1e36 molec/cm^3 accept errors above 1e6.
FORCED_LIMIT = 1.0e6
# log10 molec/cm^3 above which a species value is unphysical (1e16 is
# 400 ppm at the surface, above every species of the mechanism).
LOG10_C_MAX = 16.0

# Nitrogen and carbon atoms per molecule used by the budget checks.  N counts
# reactive nitrogen (N2 is inert in this mechanism and left out).  C counts
# are assigned from the lumped species' molar weights in chem1_list.f90
# (builder-defined; the mechanism itself does not conserve carbon, so the
# check threshold is calibrated on FP64 RODAS3 data).
N_ATOMS = {"NO": 1, "NO2": 1, "NO3": 1, "N2O5": 2, "HONO": 1, "HNO3": 1, "HNO4": 1,
           "ONIT": 1, "PAN": 1, "OLN": 1}
C_ATOMS = {"CO": 1, "CH4": 1, "ETH": 2, "ALKA": 4, "ALKE": 3, "BIO": 5, "ARO": 7, "HCHO": 1,
           "ALD": 2, "KET": 4, "CRBO": 4, "ONIT": 4, "PAN": 2, "OP1": 1, "OP2": 2, "ORA1": 1,
           "ORA2": 2, "MO2": 1, "AKAP": 4, "AKEP": 4, "BIOP": 5, "PHO": 7, "ADD": 7, "AROP": 7,
           "CBOP": 4, "OLN": 3}


class SurModel(C.Structure):
    """brams_sur_model"""
    _fields_ = [("members", C.c_void_p * 4), ("nmembers", C.c_int32),
                ("feat_mu", C.c_void_p), ("feat_sd", C.c_void_p),
                ("targ_mu", C.c_void_p), ("targ_sd", C.c_void_p),
                ("natoms", C.c_void_p), ("catoms", C.c_void_p),
                ("floor_c", C.c_double), ("sig_c", C.c_double),
                ("spread_tol", C.c_float), ("n_tol", C.c_float), ("c_tol", C.c_float)]


class SurStats(C.Structure):
    """brams_sur_stats"""
    _fields_ = [("cells", C.c_int64), ("fallback_cells", C.c_int64), ("reasons", C.c_int32 * 4),
                ("fallback_block_attempts", C.c_int64), ("fallback_cell_attempts", C.c_int64),
                ("fallback_cell_substeps", C.c_int64), ("features_ms", C.c_float),
                ("mlp_ms", C.c_float), ("apply_ms", C.c_float), ("fallback_ms", C.c_float),
                ("total_ms", C.c_float), ("fallback_failed_blocks", C.c_int64),
                ("fallback_wps", C.c_int32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "reasons"}
        d["reasons"] = dict(zip(["nonfinite", "n_budget", "c_budget", "spread"], list(self.reasons)))
        return d


def _bind():
    L = abi.lib()
    if getattr(L, "_sur_bound", False):
        return L
    dev = C.POINTER(abi.ChemDevArgs)
    L.brams_sur_features_raw.argtypes = [dev, C.c_void_p, C.c_double]
    L.brams_sur_targets_raw.argtypes = [dev, C.c_void_p, C.c_void_p, C.c_double]
    L.brams_sur_moments.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
    L.brams_sur_normalize.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int, C.c_int]
    L.brams_sur_step.argtypes = [dev, C.POINTER(SurModel), C.POINTER(SurStats)]
    L.brams_sur_reset.argtypes = []
    L.brams_sur_reset.restype = None
    L._sur_bound = True
    return L


def atom_vectors():
    mech = load_mechanism()
    n = np.array([N_ATOMS.get(s.name, 0) for s in mech.species], np.float32)
    c = np.array([C_ATOMS.get(s.name, 0) for s in mech.species], np.float32)
    return n, c


# ------------------------------------------------------------------ dataset
def make_dataset(grid, seeds, steps, fraction, atol, rtol, sza_step=0.5, log=print, rng_seed=0, skip=0):
    """Raw (features, targets) of FP64 RODAS3 rollouts on the GPU.

    For every seed a workload on `grid` is integrated for `steps` chemistry
    steps, SZA advancing `sza_step` degrees per step; before each step the
    features of every cell are taken, after it the log10 increments, and a
    random `fraction` of the cells is kept (none of the first `skip` steps,
    the transient out of the synthetic initial state).  Returns device tensors
    (feat [N][68] f32, targ [N][47] f32)."""
    import torch
    from .chem import DeviceStep
    from .workload import jphoto_from_sza, make_workload
    L = _bind()
    feats, targs = [], []
    ncell = grid.ncell
    gen = torch.Generator(device="cpu").manual_seed(rng_seed)
    for seed in seeds:
        t0 = time.time()
        wl = make_workload(grid, seed=seed, sza="split", plume_sigma=20.0, plume_factor=50.0)
        st = DeviceStep(wl, atol, rtol)
        fbuf = torch.empty((ncell, NIN), dtype=torch.float32, device=st.dev)
        tbuf = torch.empty((ncell, NOUT), dtype=torch.float32, device=st.dev)
        valid = torch.ones(ncell, dtype=torch.bool, device=st.dev)
        for t in range(steps):
            jph = torch.from_numpy(np.ascontiguousarray(
                jphoto_from_sza(grid, wl["sza_col"] + sza_step * t).ravel(order="F"))).to(st.dev)
            st.args.jphoto = jph.data_ptr()
            before = st.sc_p.clone()
            abi.check(L.brams_sur_features_raw(C.byref(st.args), fbuf.data_ptr(), FLOOR_C), "features")
            st.step()
            abi.check(L.brams_sur_targets_raw(C.byref(st.args), before.data_ptr(), tbuf.data_ptr(),
                                              FLOOR_C), "targets")
            valid &= trusted_cells(wl["block_end"], st.dev)
            valid &= torch.isfinite(fbuf).all(1) & torch.isfinite(tbuf).all(1)
            valid &= ((fbuf[:, :NOUT] < LOG10_C_MAX) & (fbuf[:, :NOUT] + tbuf < LOG10_C_MAX)).all(1)
            if t >= skip:
                keep = torch.randperm(ncell, generator=gen)[:int(fraction * ncell)].to(st.dev)
                keep = keep[valid[keep]]
                feats.append(fbuf.index_select(0, keep))
                targs.append(tbuf.index_select(0, keep))
            del jph, before
        log("dataset: seed %d, %d steps, %.1f s, %.5f of the cells left out"
            % (seed, steps, time.time() - t0, 1.0 - valid.float().mean().item()))
        del st
    return torch.cat(feats), torch.cat(targs)


def make_rollout_dataset(model, grid, seeds, steps, fraction, atol, rtol, spinup=6, sza_step=0.5,
                         log=print, rng_seed=1):
    """Cell-steps on the surrogate's own trajectories (dataset aggregation).

    For every seed the seeded state is advanced by `spinup` FP64 steps and
    then by `steps` surrogate+fallback steps; before each of those, the FP64
    RODAS3 step of the state the surrogate has reached gives the target, so
    the network learns from the states its own errors lead to.  Returns raw
    device tensors like make_dataset."""
    import torch
    from .chem import DeviceStep
    from .workload import jphoto_from_sza, make_workload
    L = _bind()
    feats, targs = [], []
    ncell = grid.ncell
    gen = torch.Generator(device="cpu").manual_seed(rng_seed)
    for seed in seeds:
        t0 = time.time()
        wl = make_workload(grid, seed=seed, sza="split", plume_sigma=20.0, plume_factor=50.0)
        sds = DeviceStep(wl, atol, rtol)
        ref = DeviceStep(wl, atol, rtol)
        sur = SurrogateStep(model, sds)
        fbuf = torch.empty((ncell, NIN), dtype=torch.float32, device=sds.dev)
        tbuf = torch.empty((ncell, NOUT), dtype=torch.float32, device=sds.dev)
        valid = torch.ones(ncell, dtype=torch.bool, device=sds.dev)
        fb = 0.0
        for t in range(spinup + steps):
            jph = torch.from_numpy(np.ascontiguousarray(
                jphoto_from_sza(grid, wl["sza_col"] + sza_step * t).ravel(order="F"))).to(sds.dev)
            sds.args.jphoto = jph.data_ptr()
            if t < spinup:
                sds.step()
                valid &= trusted_cells(wl["block_end"], sds.dev)
                continue
            ref.args.jphoto = jph.data_ptr()
            ref.sc_p.copy_(sds.sc_p)
            ref.last_dt.copy_(sds.last_dt)
            before = ref.sc_p.clone()
            abi.check(L.brams_sur_features_raw(C.byref(ref.args), fbuf.data_ptr(), FLOOR_C), "features")
            ref.step()
            abi.check(L.brams_sur_targets_raw(C.byref(ref.args), before.data_ptr(), tbuf.data_ptr(),
                                              FLOOR_C), "targets")
            ok = valid & trusted_cells(wl["block_end"], sds.dev)
            ok &= torch.isfinite(fbuf).all(1) & torch.isfinite(tbuf).all(1)
            ok &= ((fbuf[:, :NOUT] < LOG10_C_MAX) & (fbuf[:, :NOUT] + tbuf < LOG10_C_MAX)).all(1)
            keep = torch.randperm(ncell, generator=gen)[:int(fraction * ncell)].to(sds.dev)
            keep = keep[ok[keep]]
            feats.append(fbuf.index_select(0, keep))
            targs.append(tbuf.index_select(0, keep))
            s = sur.step()
            fb += s.fallback_cells / s.cells
            del jph, before
        log("rollout dataset: seed %d, %d surrogate steps, %.1f s, mean fallback %.3f"
            % (seed, steps, time.time() - t0, fb / max(steps, 1)))
        del sds, ref, sur
    return torch.cat(feats), torch.cat(targs)


def finetune_ensemble(model, feat, targ, epochs=50, batch=8192, lr0=3e-4, log=print, seed=7, zmax=30.0):
    """Continue training every member on a raw dataset, keeping the model's
    normalisation (the inference features and targets are unchanged).
    Samples with a z-score beyond zmax are left out: on the surrogate's own
    trajectories a species driven to the floor can regrow by many decades in
    the FP64 step, and a few such squared errors would dominate the loss."""
    import torch
    X = normalise(feat, model.feat_mu, model.feat_sd, NIN_PAD, True)
    Y = normalise(targ, model.targ_mu, model.targ_sd, NOUT_PAD, False)
    ok = (Y.abs() <= zmax).all(1) & (X.float().abs() <= zmax).all(1)
    n0 = X.shape[0]
    idx = torch.nonzero(ok).squeeze(1)
    X, Y = X.index_select(0, idx), Y.index_select(0, idx)
    log("finetune: %d of %d samples within |z| <= %g" % (X.shape[0], n0, zmax))
    curves = []
    for k in range(model.nmembers):
        t0 = time.time()
        net = MLP(model.nhidden, model.nlayers, max_batch=batch, seed=seed * 100 + k + 1)
        net.set_params(model.params[k])
        losses = net.train(X, Y, batch=batch, epochs=epochs, lr0=lr0, seed=seed * 100 + k + 11,
                           chunk=10, log=log)
        model.params[k] = net.get_params()
        curves.append(losses.tolist())
        log("finetune: member %d, %d epochs, loss %.4g -> %.4g, %.1f s"
            % (k, epochs, losses[0], losses[-1], time.time() - t0))
        net.close()
    model.meta["finetune"] = {"samples": int(X.shape[0]), "epochs": epochs, "lr0": lr0, "loss_curves": curves}
    del X, Y
    return model


def trusted_cells(block_end, device):
    """Cells of the last RODAS3 step whose block neither failed nor accepted
    an error above FORCED_LIMIT (device bool tensor in cell order)."""
    import torch
    nob = len(block_end)
    _, _, failed, forced = abi.block_status(nob)
    ok = (failed == 0) & (forced <= FORCED_LIMIT)
    return torch.from_numpy(np.repeat(ok, np.asarray(block_end))).to(device)


def physical_cells(device_step, buf=None):
    """Cells of a DeviceStep's current state whose species are all finite and
    below 10**LOG10_C_MAX molec/cm^3 (device bool tensor in cell order).
    `buf`, when given, is a [ncell][68] f32 device scratch tensor."""
    import torch
    ds = device_step
    if buf is None:
        buf = torch.empty((ds.grid.ncell, NIN), dtype=torch.float32, device=ds.dev)
    abi.check(_bind().brams_sur_features_raw(C.byref(ds.args), buf.data_ptr(), FLOOR_C), "features")
    c = buf[:, :NOUT]
    return (torch.isfinite(c) & (c < LOG10_C_MAX)).all(1)


def fit_normaliser(x):
    """Column mean and standard deviation of a device [N][ncol] f32 tensor."""
    import torch
    L = _bind()
    n, ncol = x.shape
    s = torch.zeros(ncol, dtype=torch.float64, device=x.device)
    q = torch.zeros(ncol, dtype=torch.float64, device=x.device)
    abi.check(L.brams_sur_moments(x.data_ptr(), n, ncol, s.data_ptr(), q.data_ptr()), "moments")
    s, q = s.cpu().numpy(), q.cpu().numpy()
    mu = s / n
    var = np.maximum(q / n - mu * mu, 0.0)
    return mu.astype(np.float32), np.sqrt(var).astype(np.float32)


def normalise(raw, mu, sd, ldo, bf16):
    import torch
    L = _bind()
    n, ncol = raw.shape
    dt = torch.bfloat16 if bf16 else torch.float32
    out = torch.empty((n, ldo), dtype=dt, device=raw.device)
    mu_d = torch.tensor(mu, dtype=torch.float32, device=raw.device)
    sd_d = torch.tensor(sd, dtype=torch.float32, device=raw.device)
    abi.check(L.brams_sur_normalize(raw.data_ptr(), n, ncol, mu_d.data_ptr(), sd_d.data_ptr(),
                                    out.data_ptr(), ldo, 1 if bf16 else 0), "normalize")
    return out


# ------------------------------------------------------------------ model
class SurrogateModel:
    """Ensemble parameters + normalisation + check thresholds."""

    def __init__(self, nhidden=256, nlayers=4, nmembers=3):
        self.nhidden, self.nlayers, self.nmembers = nhidden, nlayers, nmembers
        self.params = [None] * nmembers
        self.feat_mu = self.feat_sd = self.targ_mu = self.targ_sd = None
        self.constant = np.zeros(NOUT, bool)      # species the mechanism never changes
        self.spread_tol, self.n_tol, self.c_tol = 0.02, 1e-3, 1e-3
        self.sig_c = 1.0
        self.meta = {}

    def save(self, path):
        d = dict(nhidden=self.nhidden, nlayers=self.nlayers, nmembers=self.nmembers,
                 feat_mu=self.feat_mu, feat_sd=self.feat_sd, targ_mu=self.targ_mu,
                 targ_sd=self.targ_sd, constant=self.constant,
                 thresholds=np.array([self.spread_tol, self.n_tol, self.c_tol, self.sig_c]),
                 meta=np.array(repr(self.meta)))
        for k, p in enumerate(self.params):
            d["member%d" % k] = p
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        m = cls(int(z["nhidden"]), int(z["nlayers"]), int(z["nmembers"]))
        m.feat_mu, m.feat_sd = z["feat_mu"], z["feat_sd"]
        m.targ_mu, m.targ_sd = z["targ_mu"], z["targ_sd"]
        m.constant = z["constant"]
        m.spread_tol, m.n_tol, m.c_tol, m.sig_c = [float(x) for x in z["thresholds"]]
        m.params = [z["member%d" % k] for k in range(m.nmembers)]
        m.meta = {"source": str(z["meta"])}
        return m


def train_ensemble(feat, targ, nmembers=3, nhidden=256, nlayers=4, epochs=50, batch=8192, lr0=1e-3,
                   log=print, seed=0):
    """Normalise a raw dataset and train `nmembers` networks on it."""
    model = SurrogateModel(nhidden, nlayers, nmembers)
    fmu, fsd = fit_normaliser(feat)
    fsd = np.where(fsd > 1e-12, fsd, 1.0).astype(np.float32)
    tmu, tsd = fit_normaliser(targ)
    model.constant = tsd <= 1e-12
    tsd = np.where(model.constant, 1.0, tsd).astype(np.float32)
    tmu = np.where(model.constant, 0.0, tmu).astype(np.float32)
    model.feat_mu, model.feat_sd, model.targ_mu, model.targ_sd = fmu, fsd, tmu, tsd
    X = normalise(feat, fmu, fsd, NIN_PAD, True)
    Y = normalise(targ, tmu, tsd, NOUT_PAD, False)
    n = X.shape[0] // batch * batch
    curves = []
    for k in range(nmembers):
        t0 = time.time()
        net = MLP(nhidden, nlayers, max_batch=batch, seed=seed * 100 + k + 1)
        losses = net.train(X[:n] if n == X.shape[0] else X, Y, batch=batch, epochs=epochs, lr0=lr0,
                           seed=seed * 100 + k + 11, chunk=10, log=log)
        model.params[k] = net.get_params()
        curves.append(losses.tolist())
        log("train: member %d, %d epochs x %d batches, loss %.4g -> %.4g, %.1f s"
            % (k, epochs, X.shape[0] // batch, losses[0], losses[-1], time.time() - t0))
        net.close()
    model.meta = {"samples": int(X.shape[0]), "epochs": epochs, "batch": batch, "lr0": lr0,
                  "loss_curves": curves}
    del X, Y
    return model


# ------------------------------------------------------------------ step
class SurrogateStep:
    """brams_sur_step on the resident state of a DeviceStep."""

    def __init__(self, model, device_step):
        import torch
        self.model = model
        self.ds = device_step
        dev = device_step.dev
        self.nets = []
        for p in model.params:
            net = MLP(model.nhidden, model.nlayers, max_batch=64, seed=0)
            net.set_params(p)
            self.nets.append(net)

        def t(a):
            return torch.tensor(np.asarray(a, np.float32), device=dev)

        self._keep = [t(model.feat_mu), t(model.feat_sd), t(model.targ_mu),
                      t(np.where(model.constant, 0.0, model.targ_sd))]
        n, c = atom_vectors()
        self._keep += [t(n), t(c)]
        m = SurModel()
        for k, net in enumerate(self.nets):
            m.members[k] = net.h.value
        m.nmembers = len(self.nets)
        m.feat_mu, m.feat_sd, m.targ_mu, m.targ_sd, m.natoms, m.catoms = [x.data_ptr() for x in self._keep]
        m.floor_c = FLOOR_C
        m.sig_c = model.sig_c
        m.spread_tol, m.n_tol, m.c_tol = model.spread_tol, model.n_tol, model.c_tol
        self.cmodel = m
        self.reset()

    def reset(self):
        """Forget the fallback's per-cell memory (after restoring a state)."""
        _bind().brams_sur_reset()

    def step(self):
        """One surrogate+fallback step; the stats carry the count of fallback
        blocks that failed numerically (also brams_sur_step's return value)."""
        s = SurStats()
        abi.check(_bind().brams_sur_step(C.byref(self.ds.args), C.byref(self.cmodel), C.byref(s)),
                  "brams_sur_step")
        return s


# ------------------------------------------------------------------ metrics
def species_errors(got, ref, grid, floor_ppb=1.0e-12, valid=None):
    """Per-species median and p99 relative error of device [47][n3d] states
    over the interior cells (only `valid` ones, a bool tensor in cell order,
    when given); the floor is 1e-12 ppb in each species' unit."""
    import torch
    mech = load_mechanism()
    m1, m2, m3 = grid.shape
    g = got.view(47, m3, m2, m1)[:, grid.ja - 1:grid.jz, grid.ia - 1:grid.iz, 1:m1 - 1].double()
    r = ref.view(47, m3, m2, m1)[:, grid.ja - 1:grid.jz, grid.ia - 1:grid.iz, 1:m1 - 1].double()
    out = {}
    for s, sp in enumerate(mech.species):
        floor = floor_ppb * sp.weight / 28.96 if sp.transported else floor_ppb * 1e-9 * 6.0e18
        e = ((g[s] - r[s]).abs() / (r[s].abs() + floor)).flatten()
        if valid is not None:
            e = e[valid]
        if e.numel() > 8_000_000:
            e = e[:: (e.numel() // 8_000_000 + 1)]
        q = torch.quantile(e.float(), torch.tensor([0.5, 0.99], device=e.device)).cpu().numpy()
        out[sp.name] = (float(q[0]), float(q[1]))
    return out
