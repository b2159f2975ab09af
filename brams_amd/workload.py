"""Deterministic synthetic chemistry workloads in the reference's host layout.

Builds exactly what the Fortran caller hands chem_rodas3_dyndt
(src/brams/ccatt/chem_spack_rodas3_dyndt.f90:41-131): Fortran-ordered (k,i,j)
float32 fields pp, pi0, theta, rv, one float32 array per species (transported
species in ppbm, the others in molec/cm^3), jphoto (17,k,i,j) float64, and the
AllocIndex block partition.

Random numbers come from numpy's Philox counter-based generator, so the same
seed gives the same bytes on every machine.

Distributions (all drawn per cell unless said otherwise):
  species   log-uniform mixing ratios (ppbv): O3 10-100; NO and NO2 each
            0.005-25 (NOx 0.01-50); CO 50-500; VOC (ETH ALKA ALKE BIO ARO HCHO
            ALD KET CRBO) 0.01-20; radicals (HO HO2 O3P O1D MO2 AKAP AKEP BIOP
            PHO ADD AROP CBOP OLN XO2) 1e-9-1e-5 ppb (1e-6-1e-2 ppt); all
            other species 1e-4-10.  ppbv -> ppbm by weight/28.96, ppbv ->
            molec/cm^3 by 1e-9 * M with M = 7.243e16 p/T.
  T         300 K - 6.5 K/km * z + U(-5,5) K, z = H ln(1000 hPa / p),
            H = 287 * 260 / 9.81 m.
  p         log-linear from 1000 hPa (k=2) to 200 hPa (k=m1-1); handed over as
            Exner pi0 = cp (p/p00)^(R/cp), pp = 0, theta = T cp / pi0.
  rv        log-uniform 1e-5-2e-2 kg/kg.
  J         CLEAR_SKY_J (s^-1) times cos(SZA), 0 beyond 90 deg; then the
            RELACS relations J14 = 0.962055 J13 + 0.0106247 J9 and
            J16 = 12 J13 + 208 J15 (TUV/ModTuvDriver2.7.f90:1363-1368).
  SZA       "random": U(0,120) deg per column; "split": 30 -> 150 deg linear in
            i (half day, half night); "split_night_first": the same field
            mirrored, 150 -> 30 deg, so that the reference's mask cell
            (k=2, i=ia, j=ja, rodas3:470-482) is a night cell.
  plume     optional Gaussian patch (sigma in cells) centred in the domain in
            which NO, NO2 and the VOCs are multiplied by up to `plume_factor`.
"""
import math

import numpy as np

from .mechanism import load_mechanism

CP, RGAS, P00 = 1004.0, 287.0, 1.0e5
CPOR = CP / RGAS

# Representative clear-sky surface maxima at SZA = 0 for the 17 RELACS_TUV
# photolysis slots (chem1_list.f90:786-840 names them): NO2, O3->O1D, O3->O3P,
# HONO, HNO3, HNO4, NO3->NO, NO3->NO2, H2O2, HCHO radical, HCHO molecular,
# ALD, OP1, OP2 (from J13/J9), KET, CRBO (from J13/J15), ONIT.
CLEAR_SKY_J = np.array([
    8.0e-3, 3.0e-5, 4.0e-4, 1.6e-3, 6.0e-7, 5.0e-6, 2.0e-2, 1.7e-1, 7.0e-6,
    3.0e-5, 4.0e-5, 5.0e-6, 5.0e-6, 0.0, 1.0e-6, 0.0, 1.0e-6])

RADICALS = ["HO", "HO2", "O3P", "O1D", "MO2", "AKAP", "AKEP", "BIOP", "PHO", "ADD",
            "AROP", "CBOP", "OLN", "XO2"]
VOC = ["ETH", "ALKA", "ALKE", "BIO", "ARO", "HCHO", "ALD", "KET", "CRBO"]
RANGES = {"O3": (10.0, 100.0), "NO": (0.005, 25.0), "NO2": (0.005, 25.0), "CO": (50.0, 500.0)}
for _s in VOC:
    RANGES[_s] = (0.01, 20.0)
for _s in RADICALS:
    RANGES[_s] = (1.0e-9, 1.0e-5)
DEFAULT_RANGE = (1.0e-4, 10.0)


def atol_from_ppb(atol_ppb, p_pa=1.0e5, t_k=288.15):
    """atol given in ppb expressed in the solver's molec/cm^3 at (p, T)."""
    return atol_ppb * 1.0e-9 * (p_pa * 7.243e16 / t_k)


class Grid:
    """One BRAMS grid: Fortran extents (m1,m2,m3) and interior column bounds."""

    def __init__(self, nx, ny, nz, dtlt=120.0):
        self.m1, self.m2, self.m3 = nz + 2, nx + 2, ny + 2
        self.ia, self.iz, self.ja, self.jz = 2, nx + 1, 2, ny + 1
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dtlt = float(dtlt)

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        return (self.m1, self.m2, self.m3)


def alloc_index(grid, block_size=64):
    """AllocIndex (chem_spack_utils.f90:93-125): block sizes of one grid."""
    n = grid.ncell
    nob = 1 if block_size >= n else n // block_size
    be = np.full(nob, min(block_size, n), dtype=np.int32)
    rest = n - int(be.sum())
    b = 0
    while rest > 0:
        be[b % nob] += 1
        rest -= 1
        b += 1
    return be


def jphoto_from_sza(grid, sza_col):
    """jphoto (nr_photo,m1,m2,m3) float64 for per-column solar zenith angles
    (deg): CLEAR_SKY_J * cos(SZA), zero beyond 90 deg, with the RELACS
    relations for slots 14 and 16 (other mechanisms: RELACS_TUV's magnitudes
    repeated over their rates); levels 1 and m1 copy their neighbours."""
    m1, m2, m3 = grid.shape
    nk, ni, nj = grid.nz, grid.nx, grid.ny
    sza_col = np.asarray(sza_col, dtype=np.float64)
    cosz = np.where(sza_col < 90.0, np.cos(np.deg2rad(sza_col)), 0.0)
    mech = load_mechanism()
    nph = mech.nr_photo
    if mech.name == "RELACS_TUV":
        peak = CLEAR_SKY_J
    else:                               # synthetic magnitudes for other mechanisms' rates
        peak = CLEAR_SKY_J[np.arange(nph) % len(CLEAR_SKY_J)]
    J = peak[None, None, None, :] * cosz[None, :, :, None]              # (1, ni, nj, nph)
    J = np.broadcast_to(J, (nk, ni, nj, nph)).copy()
    if mech.name == "RELACS_TUV":
        J[..., 13] = 0.962055 * J[..., 12] + 0.0106247 * J[..., 8]
        J[..., 15] = 12.0 * J[..., 12] + 208.0 * J[..., 14]
    jph = np.zeros((nph, m1, m2, m3), dtype=np.float64, order="F")
    jph[:, 1:m1 - 1, grid.ia - 1:grid.iz, grid.ja - 1:grid.jz] = np.moveaxis(J, 3, 0)
    jph[:, 0] = jph[:, 1]
    jph[:, m1 - 1] = jph[:, m1 - 2]
    return jph


def make_workload(grid, seed=0, sza="random", plume_sigma=None, plume_factor=50.0,
                  sza_offset=0.0, all_ppbm=False):
    """Return a dict with the host arrays of one chemistry call on `grid`.

    Transported species are in ppbm and the others in molec/cm^3, as the
    Rosenbrock drivers keep them; all_ppbm=True gives every species in ppbm,
    the units the QSSA driver converts from (chem_spack_qssa.f90:184-191)."""
    mech = load_mechanism()
    rng = np.random.Generator(np.random.Philox(seed))
    m1, m2, m3 = grid.shape
    nk, ni, nj = grid.nz, grid.nx, grid.ny
    # pressure / temperature per (k,i,j) over the chemistry levels k=2..m1-1
    frac = np.arange(nk) / max(nk - 1, 1)
    p_lev = 1000.0e2 * (200.0 / 1000.0) ** frac                     # Pa
    z_lev = (RGAS * 260.0 / 9.81) * np.log(1000.0e2 / p_lev)        # m
    shape3 = (nk, ni, nj)
    T = 300.0 - 6.5e-3 * z_lev[:, None, None] + rng.uniform(-5.0, 5.0, shape3)
    p = np.broadcast_to(p_lev[:, None, None], shape3)
    rv = np.exp(rng.uniform(math.log(1e-5), math.log(2e-2), shape3))
    M = p * 7.243e16 / T
    # solar zenith angle per column
    if sza == "random":
        sza_col = rng.uniform(0.0, 120.0, (ni, nj))
    elif sza == "split":
        sza_col = np.broadcast_to((30.0 + 120.0 * np.arange(ni) / max(ni - 1, 1))[:, None], (ni, nj))
    elif sza == "split_night_first":
        sza_col = np.broadcast_to((150.0 - 120.0 * np.arange(ni) / max(ni - 1, 1))[:, None], (ni, nj))
    else:
        raise ValueError(sza)
    sza_col = sza_col + sza_offset
    # plume multiplier per column
    if plume_sigma:
        ii, jj = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
        d2 = (ii - (ni - 1) / 2.0) ** 2 + (jj - (nj - 1) / 2.0) ** 2
        plume = 1.0 + (plume_factor - 1.0) * np.exp(-d2 / (2.0 * plume_sigma ** 2))
    else:
        plume = np.ones((ni, nj))

    def full(a, dtype):
        out = np.zeros((m1, m2, m3), dtype=dtype, order="F")
        out[1:m1 - 1, grid.ia - 1:grid.iz, grid.ja - 1:grid.jz] = a
        out[0] = out[1]
        out[m1 - 1] = out[m1 - 2]
        return out

    pi0 = CP * (p / P00) ** (RGAS / CP)
    theta = T * CP / pi0
    wl = dict(grid=grid, seed=seed)
    wl["pp"] = full(np.zeros(shape3), np.float32)
    wl["pi0"] = full(pi0, np.float32)
    wl["theta"] = full(theta, np.float32)
    wl["rv"] = full(rv, np.float32)
    sc_p = []
    for s in mech.species:
        lo, hi = RANGES.get(s.name, DEFAULT_RANGE)
        ppbv = np.exp(rng.uniform(math.log(lo), math.log(hi), shape3))
        if s.name in ("NO", "NO2") or s.name in VOC:
            ppbv = ppbv * plume[None, :, :]
        if s.transported:
            val = ppbv * s.weight / 28.96
        else:
            val = ppbv * s.weight / 28.96 if all_ppbm else ppbv * 1.0e-9 * M
        sc_p.append(full(val, np.float32))
    wl["sc_p"] = sc_p
    jph = jphoto_from_sza(grid, sza_col)
    wl["jphoto"] = jph
    wl["block_end"] = alloc_index(grid)
    wl["last_accepted_dt"] = np.full(len(wl["block_end"]), float(np.float32(grid.dtlt)), dtype=np.float64)
    wl["sza_col"] = np.ascontiguousarray(sza_col)
    return wl


def tolerances(kind="baseline"):
    """(atol, rtol) per species in molec/cm^3.

    "reference": the reference's defaults, atol 1e7, rtol 1e-3
                 (src/brams/ccatt/mem_spack.f90:61-62).
    "baseline":  atol 1e-12 ppb (at 1000 hPa, 288.15 K), rtol 1e-3.
    """
    ns = load_mechanism().nspecies
    if kind == "reference":
        return np.full(ns, 1.0e7), np.full(ns, 1.0e-3)
    if kind == "baseline":
        return np.full(ns, atol_from_ppb(1.0e-12)), np.full(ns, 1.0e-3)
    raise ValueError(kind)
