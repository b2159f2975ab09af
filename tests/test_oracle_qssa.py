"""Pins of the oracle's QSSA solver (chem_spack_qssa.f90, CHEMISTRY=1).

- Production and loss: bit for bit against the reference's compiled rates,
  dratedc, fexprod and fexloss, and P - L y against the (pinned) net
  tendency f of fexchem.
- The sub-step recurrence of one species against the closed-form solution of
  y' = P - L y over itchim steps, in each branch: exponential, steady state
  (L delt >= 10), explicit (L = 0), L < 0 (ybid = 1), and the reference's
  overwritten explicit branch for 0 < 100 L delt <= 1 (qssa:404 is
  re-assigned by qssa:406, so the exponential form applies).
- The gather's unit conversion (every species from ppbm) against the ideal
  gas number density with CODATA k_B, and the Exner inversion.
- The whole step: species with neither production nor loss get a tendency
  of rounding size only (conversion and write-back are inverses), sc_p is
  unchanged, and the step equals the composition of the reference's
  compiled kinetic and P/L with the pinned per-species recurrence; a cell
  reaching 1e20 is counted and keeps its tendency.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from brams_amd.mechanism import load_mechanism
from brams_amd.workload import Grid, make_workload

MECH = load_mechanism()
NS = MECH.nspecies
KB = 1.380649e-23        # CODATA 2018, exact
ITCHIM = 75              # chemistry.f90:393


def _states(n, seed):
    rng = np.random.default_rng(seed)
    temp = rng.uniform(220.0, 310.0, n)
    press = rng.uniform(2.0e4, 1.0e5, n)
    vapp = np.exp(rng.uniform(math.log(1e-5), math.log(2e-2), n))
    jph = np.exp(rng.uniform(math.log(1e-8), math.log(1e-1), (n, 17)))
    y = np.exp(rng.uniform(math.log(1e2), math.log(1e13), (n, NS)))
    rk = O.oracle_kinetic(temp, press, vapp, jph)
    return temp, press, vapp, jph, y, rk


@pytest.mark.ref
def test_prod_loss_bit_identical_to_reference():
    *_, y, rk = _states(64, 3)
    P, L = O.oracle_prod_loss(rk, y)
    Pr, Lr = O.ref_prod_loss(rk, y)
    assert np.array_equal(P, Pr)
    assert np.array_equal(L, Lr)
    # rows the reference never assigns are 0 (6 without production, 8 without loss)
    assert int(np.sum(np.all(P == 0, axis=0))) == 6
    assert int(np.sum(np.all(L == 0, axis=0))) == 8


def test_prod_minus_loss_is_the_net_tendency():
    *_, y, rk = _states(32, 5)
    P, L = O.oracle_prod_loss(rk, y)
    f = O.oracle_fexchem(rk, y)
    scale = np.abs(P) + np.abs(L * y) + 1e-300
    assert np.all(np.abs(P - L * y - f) <= 1e-12 * scale)
    assert np.all(L >= 0.0) and np.all(P >= 0.0)


def _exact(y0, P, L, t):
    return P / L + (y0 - P / L) * math.exp(-L * t)


@pytest.mark.parametrize("y0,P,L,delt", [(5.0, 3.0, 0.02, 1.6), (1e9, 4e5, 3e-3, 1.6),
                                         (2.0, 0.0, 0.05, 0.8)])
def test_exponential_branch_matches_closed_form(y0, P, L, delt):
    got = O.qssa_species(y0, P, L, delt, ITCHIM)
    assert got == pytest.approx(_exact(y0, P, L, ITCHIM * delt), rel=1e-12)


def test_steady_state_branch():
    # L delt >= 10: the species is set to P / L at once
    assert O.qssa_species(7.0, 3.0, 20.0, 1.6, 1) == 3.0 / 20.0
    assert O.qssa_species(7.0, 3.0, 20.0, 1.6, ITCHIM) == 3.0 / 20.0


def test_explicit_branch_when_no_loss():
    got = O.qssa_species(2.0, 0.25, 0.0, 1.6, ITCHIM)
    assert got == pytest.approx(2.0 + ITCHIM * 1.6 * 0.25, rel=1e-14)


def test_negative_loss_uses_unit_rate():
    # cvmgm_dp(loss, 1, -loss) gives ybid = 1 for L < 0: y -> P + (y - P) e^-delt
    got = O.qssa_species(3.0, 0.5, -0.4, 0.2, ITCHIM)
    assert got == pytest.approx(0.5 + (3.0 - 0.5) * math.exp(-ITCHIM * 0.2), rel=1e-12)


def test_overwritten_explicit_branch():
    """100 L delt = 0.5 <= 1: qssa:404 would take the explicit step, but
    qssa:406 overwrites it; the exponential form results, which differs from
    75 explicit steps by about 1e-3 here."""
    y0, P, L, delt = 1.0, 0.0, 0.005 / 1.6, 1.6
    got = O.qssa_species(y0, P, L, delt, ITCHIM)
    assert got == pytest.approx(_exact(y0, P, L, ITCHIM * delt), rel=1e-12)
    explicit = y0 * (1.0 - L * delt) ** ITCHIM
    assert abs(got - explicit) > 5e-4 * got


def test_zero_substeps_leave_species():
    assert O.qssa_species(4.5, 1.0, 0.3, 1.6, 0) == 4.5


def _host(wl, sc_t=None):
    return O.HostStep(wl, np.zeros(NS), np.zeros(NS), 2, 1,
                      sc_t if sc_t is not None else [np.zeros_like(x) for x in wl["sc_p"]], None)


def test_gather_converts_every_species_from_ppbm():
    grid = Grid(6, 5, 4)
    wl = make_workload(grid, seed=2, all_ppbm=True)
    hs = _host(wl)
    lib = O.oracle()
    for cell in (0, 37, grid.ncell - 1):
        y = np.zeros(NS); T = np.zeros(1); p = np.zeros(1); vm = np.zeros(1)
        lib.ora_qssa_gather_cell(O.C.byref(hs.args), cell, O.ptr(y), O.ptr(T), O.ptr(p), O.ptr(vm))
        k = 1 + cell % grid.nz
        i = grid.ia - 1 + (cell // grid.nz) % grid.nx
        j = grid.ja - 1 + cell // (grid.nz * grid.nx)
        ex = (float(wl["pp"][k, i, j]) + float(wl["pi0"][k, i, j])) / 1004.0
        assert T[0] == pytest.approx(float(wl["theta"][k, i, j]) * ex, rel=1e-6)
        assert p[0] == pytest.approx(1.0e5 * ex ** (1004.0 / 287.0), rel=1e-5)
        nd = p[0] / (KB * T[0]) * 1e-6                       # molecules per cm^3
        for s, sp in enumerate(MECH.species):               # transported or not
            want = float(wl["sc_p"][s][k, i, j]) * 1e-9 * nd * 28.96 / sp.weight
            assert y[s] == pytest.approx(want, rel=1e-3), sp.name


def test_step_constant_species_round_trip_and_sc_p_unchanged():
    grid = Grid(8, 7, 5)
    wl = make_workload(grid, seed=4, all_ppbm=True)
    hs, rc = O.run_oracle_qssa(wl)
    assert rc == 0
    for s in range(NS):
        assert np.array_equal(hs.sc_p[s], wl["sc_p"][s])
    idx = (slice(1, grid.m1 - 1), slice(grid.ia - 1, grid.iz), slice(grid.ja - 1, grid.jz))
    rng = np.random.default_rng(0)
    *_, yy, rk = _states(4, 9)
    P, L = O.oracle_prod_loss(rk, yy)
    inert = [s for s in range(NS) if np.all(P[:, s] == 0) and np.all(L[:, s] == 0)]
    assert len(inert) == 3                                   # CO2, N2, H2O
    for s in range(NS):
        t = np.abs(hs.sc_t[s][idx]).astype(np.float64)
        x = np.abs(wl["sc_p"][s][idx]).astype(np.float64)
        if s in inert:
            assert np.all(t <= 4 * 2.0 ** -52 * x / grid.dtlt * 8), MECH.species[s].name
        else:
            assert np.any(t > 0.0), MECH.species[s].name


@pytest.mark.ref
def test_step_is_the_composition_of_reference_pieces():
    """Sampled cells of a 16x16x10 step against the reference's compiled
    kinetic and P/L fed to the pinned per-species recurrence: new mixing
    ratio sc_p + sc_t dtlt within rel 1e-6 (plus 1e-7 of the old value: the
    REAL tendency's rounding)."""
    grid = Grid(16, 16, 10)
    wl = make_workload(grid, seed=6, sza="split", plume_sigma=4.0, all_ppbm=True)
    hs, rc = O.run_oracle_qssa(wl)
    assert rc == 0
    lib = O.oracle()
    delt = float(np.float32(np.float32(grid.dtlt) / np.float32(ITCHIM)))
    for cell in range(0, grid.ncell, 97):
        y = np.zeros(NS); T = np.zeros(1); p = np.zeros(1); vm = np.zeros(1)
        lib.ora_qssa_gather_cell(O.C.byref(hs.args), cell, O.ptr(y), O.ptr(T), O.ptr(p), O.ptr(vm))
        k = 1 + cell % grid.nz
        i = grid.ia - 1 + (cell // grid.nz) % grid.nx
        j = grid.ja - 1 + cell // (grid.nz * grid.nx)
        jph = wl["jphoto"][:, k, i, j][None, :]
        rk = O.ref_kinetic(T, p, np.array([float(wl["rv"][k, i, j])]), jph)
        P, L = O.ref_prod_loss(rk, y[None, :])
        for s in range(NS):
            v = max(O.qssa_species(y[s], P[0, s], L[0, s], delt, ITCHIM), 0.0)
            old = float(wl["sc_p"][s][k, i, j])
            new = v * old / y[s] if y[s] != 0.0 else 0.0
            got = old + float(hs.sc_t[s][k, i, j]) * float(np.float32(grid.dtlt))
            assert abs(got - new) <= 1e-6 * abs(new) + 1e-7 * abs(old) + 1e-30, (cell, MECH.species[s].name)


def test_cell_reaching_1e20_is_counted_and_keeps_its_tendency():
    grid = Grid(5, 4, 3)
    wl = make_workload(grid, seed=8, all_ppbm=True)
    wl["sc_p"][0] = wl["sc_p"][0].copy()
    k, i, j = 2, grid.ia, grid.ja                            # interior cell 1 (0-based k index 2)
    wl["sc_p"][0][k - 1, i - 1, j - 1] = 1.0e25              # O3 far above 1e20 molec/cm^3
    sc_t = [np.full_like(x, 0.125) for x in wl["sc_p"]]
    hs, rc = O.run_oracle_qssa(wl, sc_t=sc_t)
    assert rc == 1
    for s in range(NS):
        assert hs.sc_t[s][k - 1, i - 1, j - 1] == np.float32(0.125)
    assert any(hs.sc_t[s][k - 1, i, j - 1] != np.float32(0.125) for s in range(NS))
