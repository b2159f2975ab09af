"""What the trained network gains over the method without it.

The surrogate step with the committed trained ensemble
(brams_amd/models/relacs_tuv_surrogate.npz) against the same step with
untrained networks of the same shape and normalisation, on a warm
40x40x40 state of the benchmark's distributions (split SZA, plume, six FP64
spin-up steps).  Without training nearly every cell fails the checks and is
integrated by RODAS3, so the step costs what the FP64 solver costs; with
training most cells are accepted, and the accepted cells stay close to the
FP64 step.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "brams_amd", "models", "relacs_tuv_surrogate.npz")


def _warm_state():
    import torch
    from brams_amd.chem import DeviceStep
    from brams_amd.workload import Grid, jphoto_from_sza, make_workload, tolerances
    grid = Grid(40, 40, 40)
    wl = make_workload(grid, seed=0, sza="split", plume_sigma=4.0, plume_factor=50.0)
    atol, rtol = tolerances("baseline")
    ref = DeviceStep(wl, atol, rtol)
    jph = [torch.from_numpy(np.ascontiguousarray(jphoto_from_sza(grid, wl["sza_col"] + 0.5 * t)
                                                 .ravel(order="F"))).to(ref.dev) for t in range(7)]
    for t in range(6):
        ref.args.jphoto = jph[t].data_ptr()
        ref.step()
    ref.args.jphoto = jph[6].data_ptr()
    return grid, wl, atol, rtol, ref, jph[6]


def test_trained_ensemble_beats_untrained():
    import torch
    from brams_amd.chem import DeviceStep
    from brams_amd.mlp import MLP
    from brams_amd.surrogate import SurrogateModel, SurrogateStep, species_errors
    print("case surrogate gain, 40x40x40 warm state", flush=True)
    grid, wl, atol, rtol, ref, jph = _warm_state()
    warm_sc, warm_dt = ref.sc_p.clone(), ref.last_dt.clone()
    torch.cuda.synchronize()
    ref.step()                                            # the FP64 answer of the next step
    torch.cuda.synchronize()

    trained = SurrogateModel.load(MODEL)
    untrained = SurrogateModel.load(MODEL)
    for k in range(untrained.nmembers):
        net = MLP(untrained.nhidden, untrained.nlayers, max_batch=64, seed=100 + k)
        untrained.params[k] = net.get_params()
        net.close()

    results = {}
    for name, model in (("trained", trained), ("untrained", untrained)):
        ds = DeviceStep(wl, atol, rtol)
        ds.args.jphoto = jph.data_ptr()
        ds.sc_p.copy_(warm_sc)
        ds.last_dt.copy_(warm_dt)
        sur = SurrogateStep(model, ds)
        s = sur.step()
        err = species_errors(ds.sc_p, ref.sc_p, grid)
        results[name] = (s.fallback_cells / s.cells, s.total_ms, max(v[0] for v in err.values()))
        print("%s: fallback %.3f, %.1f ms, worst median error %.4f" % ((name,) + results[name]))
    f_t, _, med_t = results["trained"]
    f_u, _, _ = results["untrained"]
    assert f_u > 0.9                 # untrained: (almost) everything falls back
    assert f_t < 0.5 * f_u           # trained: most cells are accepted
    assert med_t < 0.05              # and the step stays close to FP64 after one step
