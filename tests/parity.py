"""Parity check of species fields against the oracle.

Bar (from the task definition of this path): |x - ref| <= 1e-6 |ref| + abs,
with abs = 1e-12 ppb expressed in each species' storage unit (ppbm for
transported species, molec/cm^3 at the cell's air density otherwise).
"""
import numpy as np

from brams_amd.mechanism import load_mechanism

REL = 1.0e-6
ABS_PPB = 1.0e-12


def _air_density(wl):
    cp, rgas, p00 = 1004.0, 287.0, 1.0e5
    with np.errstate(divide="ignore", invalid="ignore"):     # halo columns are zero
        return _air_density_raw(wl, cp, rgas, p00)


def _air_density_raw(wl, cp, rgas, p00):
    exner = (wl["pp"].astype(np.float64) + wl["pi0"].astype(np.float64)) / cp
    T = wl["theta"].astype(np.float64) * exner
    p = exner ** (cp / rgas) * p00
    return p * 7.243e16 / T


def parity_report(got, ref, grid, wl=None, rel=REL):
    mech = load_mechanism()
    sl = (slice(1, grid.m1 - 1), slice(grid.ia - 1, grid.iz), slice(grid.ja - 1, grid.jz))
    M = _air_density(wl)[sl] if wl is not None else 6.0e18
    worst, nfail, per = 0.0, 0, {}
    for s, sp in enumerate(mech.species):
        a = np.asarray(got[s], dtype=np.float64)[sl]
        b = np.asarray(ref[s], dtype=np.float64)[sl]
        floor = ABS_PPB * sp.weight / 28.96 if sp.transported else ABS_PPB * 1e-9 * M
        err = np.abs(a - b)
        bad = err > rel * np.abs(b) + floor
        r = np.max(err / (np.abs(b) + floor)) if a.size else 0.0
        per[sp.name] = float(r)
        worst = max(worst, float(r))
        nfail += int(bad.sum())
    return {"worst": worst, "nfail": nfail, "per_species": per}
