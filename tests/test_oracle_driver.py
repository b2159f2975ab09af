"""The oracle's block partition, LU, RODAS3 stages, controller and whole step.

Pins: the reference's compiled AllocIndex and Sparse1.3 (oracle/_ref),
numpy.linalg.solve, the Rodas3 order of accuracy on y' = A y against
scipy.linalg.expm, the controller's closed-form special cases, and invariants
of the full step (sulphur conservation, split-method identities)."""
import numpy as np
import pytest
import scipy.linalg

import oracle_lib as O
from brams_amd.mechanism import load_mechanism
from brams_amd.workload import Grid, alloc_index, make_workload, tolerances

MECH = load_mechanism()
GRIDS = [(16, 16, 10), (1, 1, 5), (3, 3, 12), (4, 4, 9), (7, 5, 13), (98, 98, 4), (50, 50, 40)]


@pytest.mark.parametrize("nx,ny,nz", GRIDS)
def test_alloc_index_python_and_oracle_agree(nx, ny, nz):
    g = Grid(nx, ny, nz)
    be = alloc_index(g)
    nob = O.oracle().ora_alloc_index(64, g.m1, g.ia, g.iz, g.ja, g.jz, None, 0)
    out = np.zeros(nob, np.int32)
    O.oracle().ora_alloc_index(64, g.m1, g.ia, g.iz, g.ja, g.jz, O.ptr(out, O._ip), nob)
    assert np.array_equal(be, out)
    assert be.sum() == g.ncell and be.max() <= 127


@pytest.mark.ref
@pytest.mark.parametrize("nx,ny,nz", [g for g in GRIDS if g[0] * g[1] * g[2] >= 64])
def test_alloc_index_vs_reference(nx, ny, nz):
    """Grids with fewer than 64 cells are left out: there the reference's
    remainder (chem_spack_utils.f90:106) starts negative and its
    DO WHILE (resto/=0) only ends after 32-bit wrap-around, leaving a garbage
    block_end; this implementation defines that case as one block of N cells."""
    g = Grid(nx, ny, nz, dtlt=60.0)
    r = O.ref_alloc_index(64, g.m1, g.m2, g.m3, g.ia, g.iz, g.ja, g.jz, 60.0, 2)
    assert np.array_equal(r["block_end"], alloc_index(g))
    assert np.all(r["dt0"] == 120.0)
    # j -> i -> k order of the cells (chem_spack_utils.f90:144-160)
    seq = [(b, c) for b in range(r["nob"]) for c in range(r["block_end"][b])]
    ks = np.array([r["indexk"][b, c] for b, c in seq])
    is_ = np.array([r["indexi"][b, c] for b, c in seq])
    js = np.array([r["indexj"][b, c] for b, c in seq])
    gidx = np.arange(g.ncell)
    assert np.array_equal(ks, 2 + gidx % g.nz)
    assert np.array_equal(is_, g.ia + (gidx // g.nz) % g.nx)
    assert np.array_equal(js, g.ja + gidx // (g.nz * g.nx))


def _matrices(n, seed):
    rng = np.random.default_rng(seed)
    wl = make_workload(Grid(4, 4, 4), seed=seed)
    T = rng.uniform(200, 310, n); P = rng.uniform(2e4, 1e5, n); V = 10 ** rng.uniform(-5, -1.7, n)
    J = 10 ** rng.uniform(-8, -2, (n, 17)); J[::2] = 0
    y = 10 ** rng.uniform(0, 12, (n, 47))
    rk = O.oracle_kinetic(T, P, V, J)
    jac = O.oracle_jacobian(rk, y)
    dts = 10 ** rng.uniform(0, 2.1, n)
    mats = [np.eye(47) / (0.5 * dt) - jj for dt, jj in zip(dts, jac)]
    return mats


@pytest.mark.parametrize("seed", [0, 1])
def test_lu_solve_vs_numpy(seed):
    for A in _matrices(20, seed):
        mask = (A != 0).astype(np.uint8)
        b = np.random.default_rng(seed).normal(size=47) * 1e6
        x = np.zeros(47)
        Ac = np.ascontiguousarray(A); mr = np.ascontiguousarray(mask.ravel())
        rc = O.oracle().ora_lu_solve_dense(O.ptr(Ac), O.ptr(mr, O.C.POINTER(O.C.c_uint8)),
                                           O.ptr(b), O.ptr(x))
        assert rc == 0
        xr = np.linalg.solve(A, b)
        assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(A, 1) * np.linalg.norm(x) + 1e-300
        np.testing.assert_allclose(x, xr, rtol=1e-9, atol=1e-12 * np.abs(xr).max())


def test_lu_zero_pivot_is_reported():
    A = _matrices(1, 3)[0]
    A[5, :] = 0.0
    A[5, 5] = 0.0
    mask = (A != 0).astype(np.uint8)
    b = np.ones(47); x = np.zeros(47)
    Ac = np.ascontiguousarray(A); mr = np.ascontiguousarray(mask.ravel())
    rc = O.oracle().ora_lu_solve_dense(O.ptr(Ac), O.ptr(mr, O.C.POINTER(O.C.c_uint8)),
                                       O.ptr(b), O.ptr(x))
    assert rc != 0


@pytest.mark.ref
def test_lu_vs_reference_sparse():
    """Same masked matrices through the reference's Sparse1.3 (Markowitz order
    chosen on the values, reciprocal pivots): same solution to round-off."""
    r = O.ref()
    for A in _matrices(12, 7):
        ij = [(i + 1, j + 1) for i in range(47) for j in range(47) if A[i, j] != 0]
        ipos = np.array([a for a, _ in ij], np.int32); jpos = np.array([b for _, b in ij], np.int32)
        elem = (O.C.c_long * len(ij))()
        h = r.ref_sp_create(47, len(ij), O.ptr(ipos, O._ip), O.ptr(jpos, O._ip), elem)
        vals = np.array([A[i - 1, j - 1] for i, j in ij])
        assert r.ref_sp_load_factor(h, len(ij), elem, O.ptr(vals)) == 0
        b = np.linspace(1, 2, 47) * 1e5
        bs = b.copy()
        xs = np.zeros(47)
        r.ref_sp_solve(h, O.ptr(bs), O.ptr(xs))
        x = np.zeros(47)
        Ac = np.ascontiguousarray(A)
        mask = np.ascontiguousarray((A != 0).astype(np.uint8).ravel())
        O.oracle().ora_lu_solve_dense(O.ptr(Ac), O.ptr(mask, O.C.POINTER(O.C.c_uint8)), O.ptr(b), O.ptr(x))
        np.testing.assert_allclose(x, xs, rtol=1e-11, atol=1e-13 * np.abs(xs).max())


def test_rodas3_is_third_order_on_linear_problem():
    """Global error of the Rodas3 stages on y' = A y against expm: order 3.

    A is a decay chain (first-order 'reactions' 1->2->...->6), so the exact
    solution stays positive and conserves the total, like the chemistry;
    the stage clipping never triggers."""
    k = np.array([1.6, 0.8, 0.4, 0.2, 0.1])
    n = 6
    A = np.zeros((n, n))
    for i, ki in enumerate(k):
        A[i, i] -= ki
        A[i + 1, i] += ki
    y0 = np.array([10.0, 5.0, 2.0, 1.0, 0.5, 0.1])
    t_end = 4.0
    ex = scipy.linalg.expm(A * t_end) @ y0
    errs = []
    for nsteps in (8, 16, 32, 64):
        y, _ = O.rodas3_linear(A, y0, t_end / nsteps, nsteps)
        errs.append(np.linalg.norm(y - ex))
        assert abs(y.sum() - y0.sum()) < 1e-12 * y0.sum()     # linear invariant
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - 3.0) < 0.25), rates


def test_rodas3_error_estimate_is_third_order_local():
    """k4 (the embedded error estimate) of one step shrinks like h^3."""
    A = np.array([[-1.0, 0.2], [0.1, -0.5]])
    y0 = np.array([1.0, 2.0])
    ests = [O.rodas3_linear(A, y0, h, 1)[1] for h in (0.1, 0.05, 0.025, 0.0125)]
    rates = np.log2(np.array(ests[:-1]) / np.array(ests[1:]))
    assert np.all(np.abs(rates - 3.0) < 0.15) and np.all(np.diff(rates) > 0), rates


def test_rodas3_stiff_decay_is_l_stable():
    """One huge step on a very stiff decay damps instead of blowing up."""
    y, _ = O.rodas3_linear(np.array([[-1.0e6]]), np.array([1.0]), 100.0, 1)
    assert 0.0 <= y[0] < 1e-3


def test_controller_special_cases():
    # err = 1 -> factor 0.9; tiny err -> x6; huge err -> x0.2; clamps
    assert O.dt_new(10.0, 1.0, 1.2, 120.0) == pytest.approx(9.0, rel=1e-15)
    assert O.dt_new(10.0, 1e-30, 1.2, 120.0) == 60.0
    assert O.dt_new(10.0, 1e30, 1.2, 120.0) == 2.0
    assert O.dt_new(100.0, 1e-30, 1.2, 120.0) == 120.0
    assert O.dt_new(2.0, 1e30, 1.2, 120.0) == 1.2
    assert O.dt_new(10.0, 8.0, 1.2, 120.0) == pytest.approx(10.0 * 0.9 / 2.0, rel=1e-14)


def _step(seed, tol, **kw):
    wl = make_workload(Grid(16, 16, 10), seed=seed)
    atol, rtol = tolerances(tol)
    return wl, O.run_oracle(wl, atol, rtol, **kw)


STATES = [("random", None, 0), ("random", None, 1), ("split", 3.0, 0), ("split_night_first", 3.0, 0)]


@pytest.mark.ref
@pytest.mark.parametrize("tol", ["reference", "baseline"])
@pytest.mark.parametrize("sza,plume,seed", STATES)
def test_step_with_reference_sparse(sza, plume, seed, tol):
    """16x16x10 workloads (random SZA; split SZA with a plume and the mask
    cell in daylight; the same mirrored, mask cell at night): the oracle's
    fixed-order LU against the reference's own Sparse1.3, which re-derives a
    Markowitz order from cell 1 of block 1 on every block-1 attempt
    (chem_spack_rodas3_dyndt.f90:470-517).  Same mask, attempts and
    sub-steps; outputs within the parity bar (in practice bit-identical
    REAL values in all but a few cells)."""
    wl = make_workload(Grid(16, 16, 10), seed=seed, sza=sza, plume_sigma=plume)
    atol, rtol = tolerances(tol)
    hs, st, mask, rc = O.run_oracle(wl, atol, rtol)
    hr, sr, maskr, rcr = O.run_oracle(wl, atol, rtol, reference_sparse=True)
    assert rc == 0 and rcr == 0
    assert st.attempts == sr.attempts and st.substeps == sr.substeps
    assert np.array_equal(mask, maskr)
    a, b = np.stack(hs.sc_p).astype(np.float64), np.stack(hr.sc_p).astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    assert rel.max() <= 1e-6
    assert np.mean(a == b) > 0.9999
    np.testing.assert_allclose(hs.last_dt, hr.last_dt, rtol=1e-11)


@pytest.mark.parametrize("tol", ["reference", "baseline"])
@pytest.mark.parametrize("sza,plume,seed", STATES)
def test_fixed_order_passes_sparse_threshold(sza, plume, seed, tol):
    """Every pivot of the fixed order passes Sparse1.3's threshold test
    (|pivot| > 1e-3 x largest other entry of its reduced column,
    spFactor.c:202-203, spAllocate.c:191) in every cell and attempt, and in
    particular on cell 1 of block 1, the matrix the reference orders on."""
    wl = make_workload(Grid(16, 16, 10), seed=seed, sza=sza, plume_sigma=plume)
    atol, rtol = tolerances(tol)
    hs, st, mask, rc = O.run_oracle(wl, atol, rtol)
    assert rc == 0
    assert st.threshold_pivots == 0 and st.order_threshold_pivots == 0


def test_threshold_count_catches_a_small_pivot():
    """The count is live: pivot J of the fixed order set to 0.5 under column
    entries of 1000 fails the rule (1000 x 1e-3 >= 0.5) exactly once; 0.5
    under entries of 100 passes."""
    import ctypes as C
    perm, _ = O.oracle_lu_order()
    lib = O.oracle()
    lib.ora_lu_threshold_fails.argtypes = [O._dp, C.POINTER(C.c_uint8)]
    mask = np.ones(47 * 47, np.uint8)

    def count(J, below, pivot):
        a = np.eye(47) * 2.0
        a[perm[J], perm[J]] = pivot
        for I in range(J + 1, 47):
            a[perm[I], perm[J]] = below         # entries outside the filled pattern are ignored
        a = np.ascontiguousarray(a)
        return lib.ora_lu_threshold_fails(O.ptr(a), mask.ctypes.data_as(C.POINTER(C.c_uint8)))

    assert count(13, 0.0, 0.5) == 0
    assert count(13, 1000.0, 0.5) == 1
    assert count(13, 100.0, 0.5) == 0
    assert count(40, 1000.0, 0.5) == 1


@pytest.mark.ref
def test_reference_sparse_order_depends_on_pattern_only():
    """Sparse1.3's Markowitz order (spOrderAndFactor, spFactor.c:173-300) of
    cell-1 matrices taken from many cells, states and step sizes is one
    order per sparsity pattern: the daylight pattern and the night pattern
    (photolysis entries zero) each give a single order.  The order is thus a
    function of the mask, which the fixed-order LU exploits."""
    import ctypes as C
    R = O.ref()
    rng = np.random.default_rng(0)
    orders = {}
    for seed, sza in [(0, "random"), (1, "split"), (2, "split_night_first")]:
        wl = make_workload(Grid(16, 16, 10), seed=seed, sza=sza, plume_sigma=None if sza == "random" else 3.0)
        g = wl["grid"]
        for cell in rng.integers(0, g.ncell, 12):
            cell = int(cell)
            y, T, p, _ = O.gather_cell(wl, cell)
            k, r = 2 + cell % g.nz, cell // g.nz
            i, j = 2 + r % g.nx, 2 + r // g.nx
            jph = np.asarray(wl["jphoto"][:, k - 1, i - 1, j - 1])
            rk = O.oracle_kinetic([T], [p], [float(wl["rv"][k - 1, i - 1, j - 1])], [jph])[0]
            J = O.oracle_jacobian(rk[None], y[None])[0]
            for dt in (120.0, 3.0):
                A = -J + np.eye(47) * (1.0 / (0.5 * dt))
                ii, jj = np.nonzero(A)
                ipos = np.ascontiguousarray(ii + 1, np.int32)
                jpos = np.ascontiguousarray(jj + 1, np.int32)
                vals = np.ascontiguousarray(A[ii, jj])
                elem = (C.c_long * len(ii))()
                h = R.ref_sp_create(47, len(ii), O.ptr(ipos, O._ip), O.ptr(jpos, O._ip), elem)
                assert R.ref_sp_load_factor(h, len(ii), elem, O.ptr(vals)) == 0
                rm = np.zeros(47, np.int32); cm = np.zeros(47, np.int32); fi = np.zeros(1, np.int32)
                R.ref_sp_order(h, O.ptr(rm, O._ip), O.ptr(cm, O._ip), O.ptr(fi, O._ip))
                R.ref_sp_destroy(h)
                assert np.array_equal(rm, cm)                   # diagonal pivoting
                orders.setdefault((ii * 47 + jj).tobytes(), set()).add(tuple(rm))
    assert len(orders) >= 2                                     # day and night patterns seen
    assert all(len(v) == 1 for v in orders.values())


def _sulphur(wl, sc):
    g = wl["grid"]
    cp, rgas = 1004.0, 287.0
    ex = (wl["pp"].astype(np.float64) + wl["pi0"]) / cp
    T = wl["theta"] * ex
    p = ex ** (cp / rgas) * 1e5
    volmol = (6.02e23 * 1e-15 * 28.96) * p / (8.314 * T)
    s = sum(sc[MECH.index(n)].astype(np.float64) * volmol / MECH.species[MECH.index(n)].weight
            for n in ("SO2", "SULF"))
    return s[1:g.m1 - 1, g.ia - 1:g.iz, g.ja - 1:g.jz]


def test_step_conserves_sulphur():
    """SO2 + SULF is a linear invariant of the mechanism, which Rosenbrock
    stages keep to round-off; only the reference's non-negativity clipping
    (rodas3:542-549, 573-583, 617) can break it, in the few cells where SO2 is
    driven through zero."""
    wl, (hs, st, mask, rc) = _step(1, "baseline")
    assert rc == 0
    s0, s1 = _sulphur(wl, wl["sc_p"]), _sulphur(wl, hs.sc_p)
    rel = np.abs(s1 - s0) / s0
    assert np.mean(rel <= 3e-7) >= 0.99 and rel.max() < 1e-2, (np.mean(rel <= 3e-7), rel.max())


def test_step_untouched_outside_interior_and_constant_species():
    wl, (hs, st, mask, rc) = _step(2, "reference")
    g = wl["grid"]
    for s in range(47):
        a, b = hs.sc_p[s], wl["sc_p"][s]
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[-1], b[-1])
    for n in ("CO2", "N2", "H2O"):
        s = MECH.index(n)
        assert np.array_equal(hs.sc_p[s], np.maximum(wl["sc_p"][s], 0))


def test_parallel_split_tendency_identity():
    """PARALLEL, N_DYN_CHEM=1 adds (new - old)/dtlt to sc_t and leaves sc_p."""
    wl = make_workload(Grid(8, 8, 6), seed=4)
    atol, rtol = tolerances("reference")
    hsym, _, _, _ = O.run_oracle(wl, atol, rtol, split_method=0)
    zeros = [np.zeros_like(x) for x in wl["sc_p"]]
    hpar, _, _, _ = O.run_oracle(wl, atol, rtol, split_method=2, sc_t=zeros)
    g = wl["grid"]
    sl = (slice(1, g.m1 - 1), slice(g.ia - 1, g.iz), slice(g.ja - 1, g.jz))
    for s, sp in enumerate(MECH.species):
        assert np.array_equal(hpar.sc_p[s], wl["sc_p"][s]) or not sp.transported
        if sp.transported:
            old = wl["sc_p"][s][sl].astype(np.float64)
            rec = old + hpar.sc_t[s][sl].astype(np.float64) * g.dtlt
            new = hsym.sc_p[s][sl].astype(np.float64)
            # sc_t is float32 (new-old)/dtlt: the reconstruction carries ~1e-7 of |old|
            assert np.all(np.abs(rec - new) <= 1e-6 * (np.abs(old) + np.abs(new))), sp.name


def test_last_accepted_dt_carries_over():
    wl = make_workload(Grid(8, 8, 4), seed=6)
    atol, rtol = tolerances("reference")
    h1, s1, _, _ = O.run_oracle(wl, atol, rtol)
    wl2 = dict(wl)
    wl2["sc_p"] = h1.sc_p
    wl2["last_accepted_dt"] = h1.last_dt
    h2, s2, _, _ = O.run_oracle(wl2, atol, rtol)
    assert s2.attempts < s1.attempts            # warm start from the carried step
    assert np.all(h1.last_dt > 0) and np.all(h1.last_dt <= 120.0)
