"""GPU parity of the ROS2 solver (chem_spack_ros_dyndt.f90, CHEMISTRY=3)
against the oracle's ROS2 through the C-ABI.

Checked: FP32 outputs within rel 1e-6 / abs 1e-12 ppb, attempts and
accepted sub-steps equal, last_accepted_dt within rel 1e-9; the host entry
equal to the device entry bit for bit; the 65-cell blocks of the nested
shapes on 128 lanes; and the 1.6M-cell grid on 201 sampled blocks (ROS2 has
no value mask, so every block is independent of the others).
"""
import numpy as np
import pytest

import oracle_lib
from brams_amd.workload import Grid, make_workload, tolerances
from parity import parity_report

pytestmark = pytest.mark.gpu


def _gpu(wl, atol, rtol):
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    ds = DeviceStep(wl, atol, rtol, solver="ros2")
    rc = ds.step()
    nob = len(wl["block_end"])
    att, sub, fail, _ = abi.block_status(nob)
    out = (ds.species_host(), ds.last_dt.cpu().numpy(), rc, abi.stats(), att, sub)
    del ds
    return out


@pytest.mark.parametrize("tol", ["baseline", "reference"])
@pytest.mark.parametrize("seed,sza", [(0, "random"), (1, "split"), (2, "split_night_first")])
def test_ros2_16x16x10_matches_oracle(seed, sza, tol):
    print("case ros2 16x16x10 seed=%d sza=%s tol=%s" % (seed, sza, tol), flush=True)
    grid = Grid(16, 16, 10)
    wl = make_workload(grid, seed=seed, sza=sza)
    atol, rtol = tolerances(tol)
    gpu, dt, rc, st, att, sub = _gpu(wl, atol, rtol)
    hs, ost, mask, orc = oracle_lib.run_oracle(wl, atol, rtol, solver="ros2")
    oatt, osub = oracle_lib.block_counts(len(wl["block_end"]))
    rep = parity_report(gpu, hs.sc_p, grid, wl)
    print("  attempts %d/%d worst %.3g nfail %d mask %d" % (st.attempts, ost.attempts, rep["worst"], rep["nfail"],
                                                          st.mask_nnz), flush=True)
    assert rc == 0 and orc == 0
    assert st.mask_nnz == 469
    assert np.array_equal(att, oatt) and np.array_equal(sub, osub)
    assert rep["nfail"] == 0, rep
    np.testing.assert_allclose(dt, hs.last_dt, rtol=1e-9)


def test_ros2_host_entry_matches_device_entry():
    from brams_amd.chem import run_workload_host
    grid = Grid(12, 9, 7)
    wl = make_workload(grid, seed=3)
    atol, rtol = tolerances("baseline")
    sc_p, last_dt, rc, _ = run_workload_host(wl, atol, rtol, solver="ros2")
    gpu, dt, rc2, _, _, _ = _gpu(wl, atol, rtol)
    assert rc == 0 and rc2 == 0
    for a, b in zip(sc_p, gpu):
        assert np.array_equal(a, b)
    assert np.array_equal(last_dt, dt)


@pytest.mark.parametrize("nx,nz,dt", [(98, 4, 60.0), (50, 6, 30.0)])
def test_ros2_nested_shapes_match_oracle(nx, nz, dt):
    print("case ros2 nested nx=%d nz=%d dt=%g" % (nx, nz, dt), flush=True)
    grid = Grid(nx, nx, nz, dtlt=dt)
    wl = make_workload(grid, seed=21, sza="split")
    assert max(wl["block_end"]) == 65
    atol, rtol = tolerances("baseline")
    gpu, gdt, rc, st, att, sub = _gpu(wl, atol, rtol)
    hs, ost, mask, orc = oracle_lib.run_oracle(wl, atol, rtol, solver="ros2")
    oatt, osub = oracle_lib.block_counts(len(wl["block_end"]))
    rep = parity_report(gpu, hs.sc_p, grid, wl)
    print("  attempts %d/%d worst %.3g nfail %d" % (st.attempts, ost.attempts, rep["worst"], rep["nfail"]), flush=True)
    assert rc == 0 and orc == 0
    assert np.array_equal(att, oatt) and np.array_equal(sub, osub)
    assert rep["nfail"] == 0, rep
    np.testing.assert_allclose(gdt, hs.last_dt, rtol=1e-9)


def test_ros2_200x200x40_sampled_blocks_match_oracle():
    from test_gpu_scale_parity import compare, sample_blocks
    print("case ros2 200x200x40 split, 201 sampled blocks", flush=True)
    grid = Grid(200, 200, 40, dtlt=120.0)
    wl = make_workload(grid, seed=0, sza="split")
    atol, rtol = tolerances("baseline")
    gpu, gdt, rc, st, att, sub = _gpu(wl, atol, rtol)
    nob = len(wl["block_end"])
    blocks = sample_blocks(nob, 125)
    hs, ost, mask, orc = oracle_lib.run_oracle_blocks(wl, atol, rtol, blocks, solver="ros2")
    oatt, osub = oracle_lib.block_counts(nob)
    ids = np.concatenate([[0], blocks])
    worst, nfail = compare(grid, wl, gpu, hs.sc_p, ids)
    flips = int(np.sum(att[ids] != oatt[ids]))
    print("  %d blocks, gpu %.1f ms attempts/cell %.1f | sampled %d | flips %d worst %.3g nfail %d"
          % (nob, st.kernel_ms, st.cell_attempts / st.cells, len(ids), flips, worst, nfail), flush=True)
    assert rc == 0 and orc == 0
    assert flips == 0 and np.array_equal(sub[ids], osub[ids])
    assert nfail == 0, worst
    np.testing.assert_allclose(gdt[ids], hs.last_dt[ids], rtol=1e-9)


def test_ros2_longest_first_block_order_changes_no_result():
    """The second ROS2 call launches its 64-cell blocks ordered by the first
    call's attempts; every block's result must stay bit-identical."""
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    print("case ros2 block order, 24x24x10 seed 5", flush=True)
    abi.lib().brams_chem_finalize()
    grid = Grid(24, 24, 10)
    wl = make_workload(grid, seed=5, sza="split")
    atol, rtol = tolerances("baseline")
    ds = DeviceStep(wl, atol, rtol, solver="ros2")
    ds.step()
    first = ds.sc_p.cpu().numpy().copy()
    first_dt = ds.last_dt.cpu().numpy().copy()
    att, _, _, _ = abi.block_status(len(wl["block_end"]))
    assert att.max() > att.min()
    ds.reset()
    ds.step()
    assert np.array_equal(first, ds.sc_p.cpu().numpy())
    assert np.array_equal(first_dt, ds.last_dt.cpu().numpy())
