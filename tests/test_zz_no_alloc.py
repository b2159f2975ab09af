"""brams_chem_rodas3_dyndt_device_noalloc against brams_chem_rodas3_dyndt_device
on the same inputs, through the C-ABI: FP32 species, last_accepted_dt and
the failed-block count bit-identical.  Three calls in a row on shrinking
grids, so each call reuses what the previous one left in the entry's fixed
device array: 98x98x4 (600 blocks, the first 16 of 65 cells on 128 lanes),
16x16x10 (40 blocks of 64), 7x19x1 (two blocks of 67 and 66 cells, every
block on 128 lanes)."""
import ctypes as C

import numpy as np
import pytest

from brams_amd import abi
from brams_amd.workload import Grid, make_workload, tolerances

pytestmark = pytest.mark.gpu

CASES = [(Grid(98, 98, 4, dtlt=60.0), 21, "baseline"),
         (Grid(16, 16, 10), 3, "reference"),
         (Grid(7, 19, 1), 8, "baseline")]


def test_noalloc_entry_matches_device_entry():
    import torch
    from brams_amd.chem import DeviceStep
    L = abi.lib()
    L.brams_chem_rodas3_dyndt_device_noalloc.argtypes = [C.POINTER(abi.ChemDevArgs)]
    for grid, seed, tol in CASES:
        print("case noalloc %dx%dx%d seed %d tol %s" % (grid.nx, grid.ny, grid.nz, seed, tol), flush=True)
        wl = make_workload(grid, seed=seed, sza="random")
        atol, rtol = tolerances(tol)
        ref = DeviceStep(wl, atol, rtol)
        rc_ref = ref.step()
        na = DeviceStep(wl, atol, rtol)
        rc_na = L.brams_chem_rodas3_dyndt_device_noalloc(C.byref(na.args))
        torch.cuda.synchronize()
        assert rc_na >= 0, L.brams_chem_last_error()
        assert rc_na == rc_ref
        assert torch.equal(na.sc_p, ref.sc_p)
        assert torch.equal(na.last_dt, ref.last_dt)
        assert not torch.equal(na.sc_p, na.sc_p0)        # the step did change the state
        print("  rc %d, %d blocks bit-identical" % (rc_na, len(wl["block_end"])), flush=True)
        del ref, na
