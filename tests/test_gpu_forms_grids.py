"""Kernel forms and the multi-grid pass on the GPU, through the C-ABI.

- The per-lane kernel at 1, 2 and 4 waves per SIMD (forms 3, 4, 5) gives
  bit-identical results; the autotuner times all three on the first call of
  a grid and keeps one.
- brams_chem_rodas3_dyndt_device_grids integrates several grids in one pass
  and must give, bit for bit, what one brams_chem_rodas3_dyndt_device call
  per grid gives (chemistry.f90:313-320 calls the step once per grid), on a
  first call and on a second call (longest-first order across the grids).
"""
import ctypes as C

import numpy as np
import pytest

from brams_amd import abi
from brams_amd.workload import Grid, make_workload, tolerances

pytestmark = pytest.mark.gpu


def test_register_budget_forms_are_bit_identical_and_tuned():
    import torch
    from brams_amd.chem import DeviceStep
    print("case forms 3/4/5, 24x24x10 seed 2", flush=True)
    grid = Grid(24, 24, 10)
    wl = make_workload(grid, seed=2, sza="split", plume_sigma=4.0)
    atol, rtol = tolerances("baseline")
    outs = []
    try:
        for form in (3, 4, 5):
            abi.set_form(form)
            ds = DeviceStep(wl, atol, rtol)
            rc = ds.step()
            st = abi.stats()
            assert rc == 0 and st.kernel_form == form
            outs.append((ds.sc_p.clone(), ds.last_dt.clone()))
            del ds
    finally:
        abi.set_form(-1)
    for sc, dt in outs[1:]:
        assert torch.equal(sc, outs[0][0]) and torch.equal(dt, outs[0][1])
    abi.lib().brams_chem_finalize()                     # forget the grid: the next call tunes
    ds = DeviceStep(wl, atol, rtol)
    assert ds.step() == 0
    form, ms = abi.tuning()
    print("  autotuned form %d, ms per form %s" % (form, list(np.round(ms, 2))), flush=True)
    assert form in (3, 4, 5)
    assert np.all(ms[3:6] > 0)
    assert torch.equal(ds.sc_p, outs[0][0]) and torch.equal(ds.last_dt, outs[0][1])


def test_ros2_forms_are_bit_identical_and_tuned():
    """The ROS2 step at one and four waves per SIMD (forms 3 and 5) gives
    bit-identical results; form 4 (two waves) does not exist for ROS2 and is
    refused; the first ROS2 call of a grid times both and keeps one."""
    import torch
    from brams_amd.chem import DeviceStep
    print("case ros2 forms 3/5, 24x24x10 seed 2", flush=True)
    grid = Grid(24, 24, 10)
    wl = make_workload(grid, seed=2, sza="split")
    atol, rtol = tolerances("baseline")
    outs = []
    try:
        for form in (3, 5):
            abi.set_form(form)
            ds = DeviceStep(wl, atol, rtol, solver="ros2")
            assert ds.step() == 0 and abi.stats().kernel_form == form
            outs.append((ds.sc_p.clone(), ds.last_dt.clone()))
            del ds
        abi.set_form(4)
        ds = DeviceStep(wl, atol, rtol, solver="ros2")
        with pytest.raises(abi.ChemError):
            ds.step()
        del ds
    finally:
        abi.set_form(-1)
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    abi.lib().brams_chem_finalize()                     # forget the grid: the next call tunes
    ds = DeviceStep(wl, atol, rtol, solver="ros2")
    assert ds.step() == 0 and abi.stats().kernel_form in (3, 5)
    assert torch.equal(ds.sc_p, outs[0][0]) and torch.equal(ds.last_dt, outs[0][1])


NESTED = [(Grid(40, 40, 10, dtlt=120.0), 0, "split"),
          (Grid(98, 98, 4, dtlt=60.0), 1, "split_night_first"),
          (Grid(50, 50, 6, dtlt=30.0), 2, "random")]


def test_grids_in_one_pass_match_one_call_per_grid():
    import torch
    from brams_amd.chem import DeviceStep
    L = abi.lib()
    L.brams_chem_rodas3_dyndt_device_grids.argtypes = [C.POINTER(abi.ChemDevArgs), C.c_int32]
    atol, rtol = tolerances("baseline")
    wls = [make_workload(g, seed=s, sza=z, plume_sigma=4.0) for g, s, z in NESTED]
    ref = [DeviceStep(wl, atol, rtol) for wl in wls]
    mul = [DeviceStep(wl, atol, rtol) for wl in wls]
    arr = (abi.ChemDevArgs * len(mul))(*[d.args for d in mul])
    for call in (1, 2):
        print("case grids in one pass, call %d" % call, flush=True)
        rc_ref = sum(d.step() for d in ref)
        rc = L.brams_chem_rodas3_dyndt_device_grids(arr, len(mul))
        torch.cuda.synchronize()
        assert rc >= 0, L.brams_chem_last_error()
        st = abi.stats()
        assert rc == rc_ref
        assert st.blocks == sum(len(wl["block_end"]) for wl in wls)
        for i, (r, m) in enumerate(zip(ref, mul)):
            assert torch.equal(r.sc_p, m.sc_p), "grid %d species" % i
            assert torch.equal(r.last_dt, m.last_dt), "grid %d last_accepted_dt" % i
        print("  rc %d, %d blocks over %d grids bit-identical" % (rc, st.blocks, len(mul)), flush=True)
    # two grids of the same extents would share one grid's state: refused
    twice = (abi.ChemDevArgs * 2)(mul[0].args, mul[0].args)
    assert L.brams_chem_rodas3_dyndt_device_grids(twice, 2) == -2
    assert b"same extents" in L.brams_chem_last_error()
