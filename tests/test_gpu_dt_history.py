"""Restart from the dt history record on the GPU: two chemistry steps in a
row against one step, a restart (the library's grid state dropped with
brams_chem_finalize, the species and the record's per-block step sizes
restored) and the second step.  The restarted run must reproduce the
uninterrupted one bit for bit (species, last_accepted_dt, attempts).  A
restart that starts the blocks afresh at dtlt*N_DYN_CHEM, as AllocIndex does
(chem_spack_utils.f90:170-172), is run beside it to show what the record
keeps."""
import numpy as np
import pytest

from brams_amd import abi
from brams_amd.chem import DeviceStep, dt_history_pack, dt_history_unpack
from brams_amd.workload import Grid, make_workload, tolerances


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["rodas3", "ros2"])
def test_restart_from_dt_history_reproduces_the_run(solver):
    grid = Grid(32, 32, 16)
    print("dt history restart: %s on 32x32x16" % solver, flush=True)
    wl = make_workload(grid, seed=11, sza="split")
    atol, rtol = tolerances("baseline")
    ds = DeviceStep(wl, atol, rtol, solver=solver)
    assert ds.step() == 0
    sc1 = ds.species_host()
    dt1 = ds.last_dt.cpu().numpy().copy()
    assert ds.step() == 0
    att2 = abi.stats().attempts
    sc2 = ds.species_host()
    dt2 = ds.last_dt.cpu().numpy().copy()
    del ds
    rec = dt_history_pack(grid, dt1)

    def restart(last_dt):
        abi.lib().brams_chem_finalize()
        wl_r = dict(wl)
        wl_r["sc_p"] = sc1
        wl_r["last_accepted_dt"] = last_dt
        ds = DeviceStep(wl_r, atol, rtol, solver=solver)
        assert ds.step() == 0
        out = (abi.stats().attempts, ds.species_host(), ds.last_dt.cpu().numpy().copy())
        del ds
        return out

    att, sc, dt = restart(dt_history_unpack(rec, grid))
    assert att == att2
    assert np.array_equal(dt, dt2)
    for s in range(len(sc2)):
        assert np.array_equal(sc[s], sc2[s]), s
    att_f, sc_f, dt_f = restart(np.full(dt1.size, np.float64(np.float32(grid.dtlt))))
    nsame = sum(np.array_equal(sc_f[s], sc2[s]) for s in range(len(sc2)))
    print("restart with the record: %d attempts = uninterrupted; fresh AllocIndex dt: %d attempts, "
          "%d of %d species bit-identical" % (att, att_f, nsame, len(sc2)))
