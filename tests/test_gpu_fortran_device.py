"""The device-resident Fortran coupling (brams_amd/fortran/mod_brams_chem_device.f90)
driven like chemistry.f90:383-460 with the PARALLEL split and N_DYN_CHEM = 2:
every dynamics step chem_accum (ntask 1) adds half that step's tendency to
sc_t_dyn on the device, every second step chem_rodas3_dyndt runs on the
device arrays and chem_accum (ntask 2) clears sc_t_dyn.  Two consecutive
chemistry steps, all in device memory, must equal bit for bit the host entry
(brams_chem_rodas3_dyndt) fed the same sc_t_dyn accumulated on the host the
reference's way (REAL product and sum rounded separately,
chemistry.f90:560-565).  With TUV j-values in the input, the driver first
maps them into jphoto on the device (chem_tuv_jphoto_device); the host
entry then gets the oracle's mapping of the same values, so the chemistry
agrees bit for bit only if the two mappings do."""
import os
import subprocess

import numpy as np
import pytest

from brams_amd.chem import run_workload_host
from brams_amd.mechanism import load_mechanism
from brams_amd.workload import Grid, make_workload, tolerances

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "brams_amd", "fortran", "device_driver")
NDYN, NCHEM = 2, 2


def _flat(a):
    return np.asarray(a).ravel(order="F")


def _tuv_input(grid, wl, kj=30):
    """TUV j-values whose mapping differs from wl's J: reaction 2n+1 carries
    rate n scaled by 1.1 (level m of valJLoc from level m+1 of jphoto)."""
    ni, nj = grid.iz - grid.ia + 1, grid.jz - grid.ja + 1
    kz = grid.m1 + 2
    valj = np.zeros((ni, nj, kj, kz), order="F")
    xref = np.zeros(kj, np.int32)
    J = np.asarray(wl["jphoto"], np.float64)
    for n in range(min(17, kj // 2)):
        xref[2 * n + 1] = n + 1
        blk = J[n, 1:, grid.ia - 1:grid.iz, grid.ja - 1:grid.jz] * 1.1      # (m1-1, ni, nj)
        valj[:, :, 2 * n + 1, :grid.m1 - 1] = np.transpose(blk, (1, 2, 0))
    return valj, xref


@pytest.mark.parametrize("tuv", [False, True])
def test_fortran_device_coupling_matches_host_entry(tmp_path, tuv):
    import oracle_lib
    assert os.path.exists(DRIVER), "Fortran device driver not built"
    mech = load_mechanism()
    grid = Grid(10, 8, 9)
    wl = make_workload(grid, seed=12)
    valj, xref = _tuv_input(grid, wl) if tuv else (None, None)
    atol, rtol = tolerances("baseline")
    n3 = grid.m1 * grid.m2 * grid.m3
    tr = np.array([1 if s.transported else 0 for s in mech.species], np.int32)
    tidx = np.array([i + 1 for i, s in enumerate(mech.species) if s.transported], np.int32)
    rng = np.random.default_rng(7)
    # tendencies of each dynamics step: a few 1e-6 of the state per second
    tend = [[[(np.float32(1e-6) * rng.standard_normal(grid.shape).astype(np.float32)
               * np.asarray(wl["sc_p"][s], np.float32)).astype(np.float32) for s in range(47)]
             for _ in range(NDYN)] for _ in range(NCHEM)]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as fh:
        np.array([grid.m1, grid.m2, grid.m3, grid.ia, grid.iz, grid.ja, grid.jz, len(wl["block_end"]), 47,
                  len(tidx), NDYN, NCHEM, 0 if valj is None else valj.shape[2],
                  0 if valj is None else valj.shape[3]], np.int32).tofile(fh)
        np.array([grid.dtlt, 1004.0, 1004.0 / 287.0, 1.0e5], np.float32).tofile(fh)
        for k in ("pp", "pi0", "theta", "rv"):
            _flat(np.asarray(wl[k], np.float32)).tofile(fh)
        np.array([s.weight for s in mech.species], np.float32).tofile(fh)
        tr.tofile(fh)
        _flat(np.asarray(wl["jphoto"], np.float64)).tofile(fh)
        np.asarray(wl["last_accepted_dt"], np.float64).tofile(fh)
        np.asarray(atol, np.float64).tofile(fh)
        np.asarray(rtol, np.float64).tofile(fh)
        tidx.tofile(fh)
        for s in range(47):
            _flat(np.asarray(wl["sc_p"][s], np.float32)).tofile(fh)
        if valj is not None:
            _flat(valj).tofile(fh)
            xref.tofile(fh)
        for c in range(NCHEM):
            for d in range(NDYN):
                for s in range(47):
                    _flat(tend[c][d][s]).tofile(fh)
    print("case fortran device coupling 10x8x9 seed 12, PARALLEL N_DYN_CHEM=2, 2 chemistry steps, tuv %s"
          % tuv, flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    nob = len(wl["block_end"])
    dev_sc = np.frombuffer(raw[:47 * n3 * 4], np.float32).reshape(47, n3)
    dev_dt = np.frombuffer(raw[47 * n3 * 4:47 * n3 * 4 + 8 * nob], np.float64)
    off = 47 * n3 * 4 + 8 * nob
    dev_td = np.frombuffer(raw[off:off + 47 * n3 * 4], np.float32)
    dev_j = np.frombuffer(raw[off + 47 * n3 * 4:], np.float64).reshape(np.asarray(wl["jphoto"]).shape, order="F")
    assert not dev_td.any()                              # chem_accum ntask 2 cleared sc_t_dyn
    if valj is not None:
        mapped = oracle_lib.tuv_jphoto(grid, valj, xref, wl["jphoto"])
        assert np.array_equal(dev_j, mapped)
        assert not np.array_equal(mapped, np.asarray(wl["jphoto"]))
        wl = dict(wl, jphoto=mapped)

    # the host entry, sc_t_dyn accumulated in REAL as chem_accum does
    ni = np.float32(1.0) / np.float32(NDYN)
    sc = [np.array(x, np.float32, copy=True, order="F") for x in wl["sc_p"]]
    dt = np.array(wl["last_accepted_dt"], np.float64, copy=True)
    for c in range(NCHEM):
        td = [np.zeros(grid.shape, np.float32, order="F") for _ in range(47)]
        for d in range(NDYN):
            for s in np.flatnonzero(tr):
                td[s] = np.asfortranarray(td[s] + ni * tend[c][d][s])
        w = dict(wl, sc_p=sc, last_accepted_dt=dt)
        sc, dt, rc, _ = run_workload_host(w, atol, rtol, split_method="PARALLEL", n_dyn_chem=NDYN, sc_t_dyn=td)
        assert rc == 0
    for s in range(47):
        assert np.array_equal(dev_sc[s], _flat(sc[s])), mech.species[s].name
    assert np.array_equal(dev_dt, dt)
    # the tendencies reached the result: the same steps without them differ
    sc0 = [np.array(x, np.float32, copy=True, order="F") for x in wl["sc_p"]]
    dt0 = np.array(wl["last_accepted_dt"], np.float64, copy=True)
    for c in range(NCHEM):
        zeros = [np.zeros(grid.shape, np.float32, order="F") for _ in range(47)]
        sc0, dt0, _, _ = run_workload_host(dict(wl, sc_p=sc0, last_accepted_dt=dt0), atol, rtol,
                                           split_method="PARALLEL", n_dyn_chem=NDYN, sc_t_dyn=zeros)
    assert any(not np.array_equal(_flat(a), _flat(b)) for a, b in zip(sc, sc0))
