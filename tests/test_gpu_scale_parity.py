"""GPU parity at full size: the 200x200x40 grid (1.6M cells, 25 000 blocks)
and the nested 98x98x40 / 50x50x40 grids, against the oracle on a
deterministic sample of blocks.

Blocks are independent once block 1 has fixed the sparsity mask
(chem_spack_rodas3_dyndt.f90:470-506), so the oracle integrates block 1 and
the sampled blocks only (ora_rodas3_dyndt_blocks) and each sampled block's
result equals what the full reference step gives it.  Checked per sampled
block: every cell's REAL outputs within rel 1e-6 / abs 1e-12 ppb, attempts
and accepted sub-steps equal, last_accepted_dt within rel 1e-9 (the
controller's err**(-1/3) carries the round-off of fused multiply-adds, about
1e-11 relative), threshold-rule counts equal on the blocks the GPU counts
them in (block 1 and the 65..127-cell blocks); globally the mask size.

Some blocks of the plume blow up: they accept steps at dt_min whose error
overflows (above 3.4e38), and some of them end in a non-finite state or a
zero pivot, where the reference stops the model
(chem_spack_rodas3_dyndt.f90 aborts in Sparse1.3).  Which of the blown-up
blocks ends that way depends on round-off (a different FMA contraction, a
last-bit difference of exp/pow), so the check there is on the set of
blown-up blocks (failed, or a forced error of 1e30 or more): every
block the GPU flags is added to the oracle's list, the two sets must be
equal on the blocks run, both steps must report a failure when either does,
and blown-up blocks are left out of the value checks.

Both orientations of the split day/night field are run, because cell 1
(k=2, i=ia, j=ja) sets the mask for the whole call: "split" puts it in
daylight, "split_night_first" at night (photolysis entries masked out).
"""
import numpy as np
import pytest

import oracle_lib
from brams_amd.mechanism import load_mechanism
from brams_amd.workload import Grid, make_workload, tolerances

pytestmark = pytest.mark.gpu

MECH = load_mechanism()
BLOWN = 1e30      # a forced acceptance at least this large marks a blown-up block


def sample_blocks(nob, every, head=0):
    ids = set(range(1, min(head, nob)))
    ids.update(range(every, nob, every))
    ids.add(nob - 1)
    ids.discard(0)
    return np.array(sorted(ids), dtype=np.int32)


def block_cells(grid, block_end, blocks):
    """(k, i, j) 0-based array indices of the cells of the listed blocks."""
    start = np.concatenate([[0], np.cumsum(block_end)])
    g = np.concatenate([np.arange(start[b], start[b + 1]) for b in blocks])
    nk, ni = grid.nz, grid.nx
    k = 1 + g % nk
    t = g // nk
    i = grid.ia - 1 + t % ni
    j = grid.ja - 1 + t // ni
    return k, i, j


def compare(grid, wl, gpu, ref, blocks):
    k, i, j = block_cells(grid, wl["block_end"], blocks)
    exner = (wl["pp"][k, i, j].astype(np.float64) + wl["pi0"][k, i, j]) / 1004.0
    M = (exner ** (1004.0 / 287.0) * 1e5) * 7.243e16 / (wl["theta"][k, i, j] * exner)
    worst, nfail = 0.0, 0
    for s, sp in enumerate(MECH.species):
        a = np.asarray(gpu[s], np.float64)[k, i, j]
        b = np.asarray(ref[s], np.float64)[k, i, j]
        floor = 1e-12 * sp.weight / 28.96 if sp.transported else 1e-12 * 1e-9 * M
        err = np.abs(a - b)
        nfail += int((err > 1e-6 * np.abs(b) + floor).sum())
        worst = max(worst, float(np.max(err / (np.abs(b) + floor))))
    return worst, nfail


def run_case(grid, sza, tol, every, head=0, seed=0, dt_rtol=1e-9, plume=True, relax=False):
    """relax: compare the second step, from the state the GPU's first step
    leaves (species and last_accepted_dt), instead of the cold step."""
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    print("case %dx%dx%d dt=%g sza=%s tol=%s plume=%s" % (grid.nx, grid.ny, grid.nz, grid.dtlt, sza, tol, plume),
          flush=True)
    wl = make_workload(grid, seed=seed, sza=sza, plume_sigma=20.0 * grid.nx / 200.0 if plume else None,
                       plume_factor=50.0)
    atol, rtol = tolerances(tol)
    nob = len(wl["block_end"])
    ds = DeviceStep(wl, atol, rtol)
    if relax:
        ds.step()
        wl = dict(wl)
        wl["sc_p"] = ds.species_host()
        wl["last_accepted_dt"] = ds.last_dt.cpu().numpy().copy()
    rc = ds.step()
    st = abi.stats()
    gpu = ds.species_host()
    gdt = ds.last_dt.cpu().numpy()
    gatt, gsub, gfail, gforced = abi.block_status(nob)
    gth = abi.block_thresholds(nob)
    del ds
    gblown = (gfail != 0) | (gforced >= BLOWN)
    gbad = np.nonzero(gblown)[0]
    blocks = np.union1d(sample_blocks(nob, every, head), gbad[gbad > 0]).astype(np.int32)
    hs, ost, mask, orc = oracle_lib.run_oracle_blocks(wl, atol, rtol, blocks)
    oatt, osub = oracle_lib.block_counts(nob)
    oth = oracle_lib.block_thresholds(nob)
    ofail = oracle_lib.block_failed(nob)
    oblown = (ofail != 0) | (oracle_lib.block_forced(nob) >= BLOWN)
    ran = np.concatenate([[0], blocks])
    ids = ran[~gblown[ran] & ~oblown[ran]]
    worst, nfail = compare(grid, wl, gpu, hs.sc_p, ids)
    flips = int(np.sum(gatt[ids] != oatt[ids]))
    print("  %d blocks, %d run on the oracle | gpu %.1f ms attempts/cell %.1f | mask %d/%d | "
          "blown up gpu %s (failed %s) oracle %s (failed %s) | attempt mismatches %d | worst %.3g "
          "nfail %d | threshold %d"
          % (nob, len(ran), st.kernel_ms, st.cell_attempts / st.cells, st.mask_nnz, ost.mask_nnz,
             gbad.tolist(), np.nonzero(gfail)[0].tolist(), ran[oblown[ran]].tolist(),
             np.nonzero(ofail)[0].tolist(), flips, worst, nfail, int(gth.sum())), flush=True)
    assert np.array_equal(gblown[ran], oblown[ran])
    assert (rc > 0) == (orc > 0)
    assert rc == int(np.count_nonzero(gfail)) and not gblown[0]
    assert st.mask_nnz == ost.mask_nnz
    assert flips == 0
    assert np.array_equal(gsub[ids], osub[ids])
    assert nfail == 0, worst
    np.testing.assert_allclose(gdt[ids], hs.last_dt[ids], rtol=dt_rtol)
    # the GPU counts threshold-rule failures in block 1 and the 65..127-cell
    # blocks (where Sparse1.3 orders, and the 128-lane kernel); the oracle in all
    sizes = np.asarray(wl["block_end"])
    counted = ids[(ids == 0) | (sizes[ids] > 64)]
    assert np.array_equal(gth[counted], oth[counted])
    assert st.order_threshold_pivots == ost.order_threshold_pivots == 0
    return st


@pytest.mark.parametrize("tol", ["baseline", "reference"])
@pytest.mark.parametrize("sza", ["split", "split_night_first"])
def test_200x200x40_matches_oracle(sza, tol):
    """The plain 1.6M-cell step (25 000 blocks, the autotuned per-lane
    form); 501 sampled blocks plus block 1."""
    st = run_case(Grid(200, 200, 40, dtlt=120.0), sza, tol, every=50)
    assert st.blocks == 25000


@pytest.mark.parametrize("nx,dt,sza,tol", [(98, 60.0, "split", "baseline"),
                                           (98, 60.0, "split_night_first", "reference"),
                                           (50, 30.0, "split", "baseline"),
                                           (50, 30.0, "split_night_first", "reference")])
def test_nested_grids_full_depth_match_oracle(nx, dt, sza, tol):
    """The nested grids of the three-grid workload at full depth (40
    levels): their first 32 blocks have 65 cells (128-lane workgroups) and
    are all sampled; 98x98x40 has 6002 blocks, 50x50x40 1562."""
    run_case(Grid(nx, nx, 40, dtlt=dt), sza, tol, every=25, head=40)
