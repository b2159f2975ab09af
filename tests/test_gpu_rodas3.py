"""GPU parity of the RODAS3 step (HIP kernels through the C-ABI) vs the oracle."""
import numpy as np
import pytest

import oracle_lib
from brams_amd.workload import Grid, make_workload, tolerances
from parity import parity_report

pytestmark = pytest.mark.gpu


def _device_run(wl, atol, rtol, **kw):
    from brams_amd.chem import DeviceStep
    st = DeviceStep(wl, atol, rtol, **kw)
    rc = st.step()
    from brams_amd import abi
    return st.species_host(), st.last_dt.cpu().numpy(), rc, abi.stats()


@pytest.mark.parametrize("tol", ["baseline", "reference"])
@pytest.mark.parametrize("seed", [0, 1])
def test_16x16x10_device_matches_oracle(seed, tol):
    """16x16x10 block (2560 cells, 40 blocks), random SZA 0-120 deg."""
    print("case 16x16x10 seed=%d tol=%s" % (seed, tol), flush=True)
    grid = Grid(16, 16, 10)
    wl = make_workload(grid, seed=seed)
    atol, rtol = tolerances(tol)
    gpu, dt_gpu, rc, st = _device_run(wl, atol, rtol)
    hs, ost, mask, orc = oracle_lib.run_oracle(wl, atol, rtol)
    rep = parity_report(gpu, hs.sc_p, grid, wl)
    print("gpu attempts %d substeps %d | oracle %d %d | worst %.3g nfail %d | mask %d/%d"
          % (st.attempts, st.substeps, ost.attempts, ost.substeps, rep["worst"], rep["nfail"],
             st.mask_nnz, ost.mask_nnz))
    assert rc == 0 and orc == 0
    assert st.mask_nnz == ost.mask_nnz
    assert st.forced_blocks == ost.forced_blocks
    assert st.attempts == ost.attempts and st.substeps == ost.substeps
    assert st.threshold_pivots == oracle_lib.block_thresholds(len(wl["block_end"]))[0]
    assert st.order_threshold_pivots == ost.order_threshold_pivots == 0
    assert rep["nfail"] == 0, rep
    np.testing.assert_allclose(dt_gpu, hs.last_dt, rtol=1e-6)


def test_host_entry_matches_device_entry():
    """brams_chem_rodas3_dyndt (host arrays, Fortran layout) == device entry."""
    from brams_amd.chem import run_workload_host
    grid = Grid(12, 9, 7)
    wl = make_workload(grid, seed=3)
    atol, rtol = tolerances("baseline")
    sc_p, last_dt, rc, _ = run_workload_host(wl, atol, rtol)
    gpu, dt_gpu, rc2, _ = _device_run(wl, atol, rtol)
    assert rc == 0 and rc2 == 0
    for a, b in zip(sc_p, gpu):
        assert np.array_equal(a, b)
    assert np.array_equal(last_dt, dt_gpu)


def test_longest_first_block_order_changes_no_result():
    """The second call on a partition launches its blocks ordered by the
    first call's attempts; every block's result must stay bit-identical."""
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    print("case block order, 24x24x10 seed 5", flush=True)
    abi.lib().brams_chem_finalize()                     # no attempts history: partition order
    grid = Grid(24, 24, 10)
    wl = make_workload(grid, seed=5, sza="split", plume_sigma=4.0)
    atol, rtol = tolerances("baseline")
    ds = DeviceStep(wl, atol, rtol)
    ds.step()
    first = ds.sc_p.cpu().numpy().copy()
    first_dt = ds.last_dt.cpu().numpy().copy()
    att, _, _, _ = abi.block_status(len(wl["block_end"]))
    assert att.max() > att.min()                        # the order is not the identity
    ds.reset()
    ds.step()                                           # ordered by the first call's attempts
    assert np.array_equal(first, ds.sc_p.cpu().numpy())
    assert np.array_equal(first_dt, ds.last_dt.cpu().numpy())


@pytest.mark.parametrize("nx,nz,dt", [(98, 4, 60.0), (50, 6, 30.0)])
def test_nested_grid_shapes_match_oracle(nx, nz, dt):
    """The nested grids of the three-grid workload (98x98 and 50x50 columns,
    dt 60 s and 30 s) at reduced depth: AllocIndex gives them 65-cell blocks
    (16 and 24 of them), which run as 128-lane workgroups with a two-wave
    step-size reduction; split SZA and a plume, atol 1e-12 ppb."""
    print("case nested nx=%d nz=%d dt=%g" % (nx, nz, dt), flush=True)
    grid = Grid(nx, nx, nz, dtlt=dt)
    wl = make_workload(grid, seed=21, sza="split", plume_sigma=8.0)
    assert max(wl["block_end"]) == 65
    atol, rtol = tolerances("baseline")
    gpu, dt_gpu, rc, st = _device_run(wl, atol, rtol)
    hs, ost, mask, orc = oracle_lib.run_oracle(wl, atol, rtol)
    rep = parity_report(gpu, hs.sc_p, grid, wl)
    print("gpu attempts %d | oracle %d | worst %.3g nfail %d" % (st.attempts, ost.attempts, rep["worst"],
                                                                 rep["nfail"]))
    assert rc == 0 and orc == 0
    assert st.attempts == ost.attempts and st.substeps == ost.substeps
    assert rep["nfail"] == 0, rep
    np.testing.assert_allclose(dt_gpu, hs.last_dt, rtol=1e-6)


@pytest.mark.parametrize("method,ndyn", [("SEQUENTIAL", 1), ("PARALLEL", 1), ("PARALLEL", 2)])
def test_split_methods_match_oracle(method, ndyn):
    """The write-back forms of chem_spack_rodas3_dyndt.f90:713-798 through
    the host entry: SEQUENTIAL overwrites sc_p; PARALLEL with N_DYN_CHEM=1
    adds (new - old)/dtlt to sc_t and leaves sc_p; PARALLEL with N_DYN_CHEM>1
    starts from sc_p - sc_t_dyn*N*dtlt and writes sc_t_dyn*N*dtlt + new."""
    from brams_amd.chem import run_workload_host
    print("case split %s N=%d, 16x16x6 seed 4" % (method, ndyn), flush=True)
    grid = Grid(16, 16, 6)
    wl = make_workload(grid, seed=4, sza="split", plume_sigma=3.0)
    atol, rtol = tolerances("reference")
    code = {"SEQUENTIAL": 1, "PARALLEL": 2}[method]
    sc_t = sc_t_dyn = None
    if method == "PARALLEL" and ndyn == 1:
        rng = np.random.default_rng(4)
        sc_t = [np.asfortranarray((rng.standard_normal(a.shape) * 1e-6).astype(np.float32)) for a in wl["sc_p"]]
    if method == "PARALLEL" and ndyn > 1:
        # a dynamics tendency that leaves a positive start state sc_p - sc_t_dyn*N*dtlt
        sc_t_dyn = [np.asfortranarray((-0.01 * a / (ndyn * grid.dtlt)).astype(np.float32)) for a in wl["sc_p"]]
    gpu, dt_gpu, rc, gpu_t = run_workload_host(wl, atol, rtol, split_method=method, n_dyn_chem=ndyn,
                                               sc_t=sc_t, sc_t_dyn=sc_t_dyn)
    hs, ost, _, orc = oracle_lib.run_oracle(wl, atol, rtol, split_method=code, n_dyn_chem=ndyn,
                                            sc_t=sc_t, sc_t_dyn=sc_t_dyn)
    assert rc == 0 and orc == 0
    np.testing.assert_allclose(dt_gpu, hs.last_dt, rtol=1e-6)
    if method == "PARALLEL" and ndyn == 1:
        for s, (a, b) in enumerate(zip(gpu, wl["sc_p"])):      # sc_p untouched
            assert np.array_equal(a, b), s
        got, ref = gpu_t, hs.sc_t
    else:
        got, ref = gpu, hs.sc_p
    worst = 0.0
    for s, (a, b) in enumerate(zip(got, ref)):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        tol = 1e-6 * np.abs(b) + 1e-12 * max(np.abs(b).max(), 1e-30)
        bad = np.abs(a - b) > tol
        worst = max(worst, float((np.abs(a - b) / (np.abs(b) + 1e-30)).max()))
        assert not bad.any(), (s, int(bad.sum()))
    print("worst rel %.3g" % worst)

