"""The Fortran drop-ins (brams_amd/fortran/mod_chem_spack_rodas3_dyndt.f90,
mod_chem_spack_ros_dyndt.f90) called like chemistry.f90:443-453 / 430-440
from a Fortran driver, through
ISO_C_BINDING into the HIP library: same result as the C-ABI called from
Python, and parity with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
from brams_amd.mechanism import default_name, load_mechanism
from brams_amd.workload import Grid, make_workload, tolerances
from parity import parity_report

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the driver linked to the process's mechanism library (shim_driver_<mech>
# for the mechanisms other than RELACS_TUV)
DRIVER = os.path.join(ROOT, "brams_amd", "fortran",
                      "shim_driver" if default_name() == "relacs_tuv" else "shim_driver_" + default_name())
MECH = load_mechanism()
NS = MECH.nspecies


def _block_index(grid, block_end):
    nob, mbs = len(block_end), int(block_end.max())
    ik = np.zeros((mbs, nob), np.int32, order="F"); ii = np.zeros_like(ik); ij = np.zeros_like(ik)
    g = 0
    for b in range(nob):
        for c in range(block_end[b]):
            ik[c, b] = 2 + g % grid.nz
            ii[c, b] = grid.ia + (g // grid.nz) % grid.nx
            ij[c, b] = grid.ja + g // (grid.nz * grid.nx)
            g += 1
    return ik, ii, ij


def _write_input(path, wl, atol, rtol):
    mech = load_mechanism()
    g = wl["grid"]
    be = np.asarray(wl["block_end"], np.int32)
    tidx = np.array([i + 1 for i, s in enumerate(mech.species) if s.transported], np.int32)
    nidx = np.array([i + 1 for i, s in enumerate(mech.species) if not s.transported], np.int32)
    ik, ii, ij = _block_index(g, be)
    with open(path, "wb") as fh:
        np.array([g.m1, g.m2, g.m3, len(be), int(be.max()), mech.nspecies, mech.nr, mech.nr_photo,
                  len(tidx), len(nidx), 1], np.int32).tofile(fh)
        np.array([g.dtlt, 1004.0, 1004.0 / 287.0, 1.0e5], np.float32).tofile(fh)
        for k in ("pp", "pi0", "theta", "rv"):
            fh.write(np.asfortranarray(wl[k], np.float32).tobytes(order="F"))
        np.array([s.weight for s in mech.species], np.float32).tofile(fh)
        fh.write(np.asfortranarray(wl["jphoto"]).tobytes(order="F"))
        np.asarray(wl["last_accepted_dt"], np.float64).tofile(fh)
        np.asarray(atol, np.float64).tofile(fh)
        np.asarray(rtol, np.float64).tofile(fh)
        tidx.tofile(fh); nidx.tofile(fh); be.tofile(fh)
        for a in (ik, ii, ij):
            fh.write(a.tobytes(order="F"))
        for x in wl["sc_p"]:
            fh.write(np.asfortranarray(x).tobytes(order="F"))


def run_driver(tmp_path, wl, atol, rtol, mode):
    """The driver on one call (mode rodas3, ros2 or qssa): the species (or
    QSSA's tendencies) as Fortran-ordered arrays and last_accepted_dt."""
    assert os.path.exists(DRIVER), "Fortran driver not built"
    grid = wl["grid"]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, wl, atol, rtol)
    r = subprocess.run([DRIVER, fin, fout, mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n3 = grid.m1 * grid.m2 * grid.m3
    raw = open(fout, "rb").read()
    sc = np.frombuffer(raw[:NS * n3 * 4], np.float32).reshape(NS, n3)
    fort = [np.asfortranarray(sc[s].reshape(grid.shape, order="F")) for s in range(NS)]
    return fort, np.frombuffer(raw[NS * n3 * 4:], np.float64)


@pytest.mark.parametrize("solver", ["rodas3", "ros2"])
def test_fortran_shim_matches_c_abi_and_oracle(tmp_path, solver):
    """chem_rodas3_dyndt (mod_chem_spack_rodas3_dyndt) and chem_ros_dyndt
    (mod_chem_spack_ros_dyndt) with the reference's argument lists."""
    grid = Grid(10, 8, 9)
    wl = make_workload(grid, seed=11)
    atol, rtol = tolerances("baseline")
    print("case fortran shim %s 10x8x9 seed=11" % solver, flush=True)
    fort, fort_dt = run_driver(tmp_path, wl, atol, rtol, solver)
    from brams_amd.chem import run_workload_host
    py, py_dt, rc, _ = run_workload_host(wl, atol, rtol, solver=solver)
    assert rc == 0
    for a, b in zip(fort, py):
        assert np.array_equal(a, b)
    assert np.array_equal(fort_dt, py_dt)
    hs, st, mask, orc = oracle_lib.run_oracle(wl, atol, rtol, solver=solver)
    rep = parity_report(fort, hs.sc_p, grid, wl)
    assert rep["nfail"] == 0, rep


def test_fortran_qssa_matches_c_abi_and_oracle(tmp_path):
    """chem_qssa (mod_chem_spack_qssa) with the reference's argument list:
    tendencies bit-identical to the C-ABI's, parity with the oracle."""
    from brams_amd.chem import run_workload_qssa_host
    from test_gpu_qssa import compare
    grid = Grid(10, 8, 9)
    wl = make_workload(grid, seed=11, sza="split", all_ppbm=True)
    atol, rtol = tolerances("baseline")
    print("case fortran shim qssa 10x8x9 seed=11", flush=True)
    fort, _ = run_driver(tmp_path, wl, atol, rtol, "qssa")
    py, rc = run_workload_qssa_host(wl)
    assert rc == 0
    for a, b in zip(fort, py):
        assert np.array_equal(a, b)
    hs, orc = oracle_lib.run_oracle_qssa(wl)
    assert orc == 0
    _, nfail = compare(grid, wl, fort, hs)
    assert nfail == 0
