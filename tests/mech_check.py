"""Checks of one mechanism other than RELACS_TUV, run as a child process with
BRAMS_MECH=<mech> (a process works with one mechanism, as a BRAMS build
does; tests/test_mechanisms.py and tests/test_gpu_mechanisms.py start it).

  --cpu  the oracle built for the mechanism against the reference's own
         kernels compiled from src/brams/ccatt/<MECH>/ (oracle/_ref):
         kinetic, fexchem, jacdchemdc and the P/L split bit for bit on
         random states, and the whole RODAS3 step through the reference's
         Sparse1.3 (its own Markowitz ordering) against the oracle's fixed
         pivot order: equal attempts, sub-steps and mask, outputs within
         rel 1e-6, and no pivot of the fixed order failing Sparse's
         threshold rule on the matrix Sparse orders on (cell 1 of block 1);
         the ROS2 step likewise, from a relaxed state (see gpu()).
  --gpu  the library built for the mechanism (libbrams_chem_<mech>.so)
         against the oracle: the RODAS3 step (both tolerance settings), the
         ROS2 step (from a relaxed state, see gpu()) and the QSSA step on
         16x16x10, the TUV mapping with the mechanism's J relations bit for
         bit, several grids in one pass (grid by grid in a lean library)
         against one call per grid and the no-allocation entry against the
         device entry, both bit for bit, the split write-backs through the
         host entry, and the Fortran drop-ins linked to the mechanism's
         library against the C-ABI, bit for bit.
  --gpu-scale  the RODAS3 step at full depth against the oracle on sampled
         blocks (tests/test_gpu_scale_parity.py's check): 50x50x40 at dt
         30 s, whose first 32 blocks have 65 cells (128-lane workgroups), and
         the 1.6M-cell 200x200x40 grid; and the ROS2 and QSSA steps on that
         grid.

Prints one line per check and exits non-zero on the first failure.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle_lib as O  # noqa: E402
from brams_amd.mechanism import load_mechanism  # noqa: E402
from brams_amd.workload import Grid, make_workload, tolerances  # noqa: E402
from parity import parity_report  # noqa: E402

MECH = load_mechanism()
NS, NP = MECH.nspecies, MECH.nr_photo


def say(msg):
    print("%s: %s" % (MECH.name, msg), flush=True)


def states(n, seed):
    rng = np.random.default_rng(seed)
    T = rng.uniform(180.0, 330.0, n)
    P = rng.uniform(1.0e4, 1.05e5, n)
    V = 10 ** rng.uniform(-6, np.log10(3e-2), n)
    J = 10 ** rng.uniform(-9, -1, (n, NP))
    J[::3] = 0.0
    y = 10 ** rng.uniform(-2, 13, (n, NS))
    return T, P, V, J, y


def cpu():
    for seed in (0, 1):
        T, P, V, J, y = states(128, seed)
        a, b = O.oracle_kinetic(T, P, V, J), O.ref_kinetic(T, P, V, J)
        assert np.array_equal(a, b), ("kinetic", np.argwhere(a != b)[:5])
        rk = b
        assert np.array_equal(O.oracle_fexchem(rk, y), O.ref_fexchem(rk, y)), "fexchem"
        assert np.array_equal(O.oracle_jacobian(rk, y), O.ref_jacobian(rk, y)), "jacobian"
        p1, l1 = O.oracle_prod_loss(rk, y)
        p2, l2 = O.ref_prod_loss(rk, y)
        assert np.array_equal(p1, p2) and np.array_equal(l1, l2), "prod/loss"
    say("kinetic, fexchem, jacdchemdc, fexprod/fexloss bit for bit against the compiled reference")
    for tol in ("reference", "baseline"):
        for sza, plume in (("random", None), ("split", 4.0)):
            wl = make_workload(Grid(16, 16, 10), seed=3, sza=sza, plume_sigma=plume)
            atol, rtol = tolerances(tol)
            hs, st, mask, rc = O.run_oracle(wl, atol, rtol)
            hr, sr, maskr, rcr = O.run_oracle(wl, atol, rtol, reference_sparse=True)
            assert rc == 0 and rcr == 0, (rc, rcr)
            assert st.attempts == sr.attempts and st.substeps == sr.substeps, (st.attempts, sr.attempts)
            assert np.array_equal(mask, maskr)
            a = np.stack(hs.sc_p).astype(np.float64)
            b = np.stack(hr.sc_p).astype(np.float64)
            rel = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))
            assert rel <= 1e-6, rel
            # Sparse1.3 applies its threshold rule where it orders: cell 1 of
            # block 1 (the other cells reuse that order)
            assert st.order_threshold_pivots == 0, st.order_threshold_pivots
            say("RODAS3 step %s/%s: oracle = reference Sparse1.3 path (attempts %d, mask %d, worst rel %.2g, "
                "threshold failures on the ordering matrix 0, elsewhere %d)"
                % (sza, tol, st.attempts, int(mask.sum()), rel, st.threshold_pivots))
    # ROS2 from the relaxed state of gpu(): the oracle's fixed order against
    # the reference's Sparse1.3 ordered on its first factorisation of the
    # full pattern (chem_spack_solve_sparse.f90); equal attempts per block,
    # outputs within rel 1e-6 or 4x the one-ulp sensitivity
    wr, sens = relaxed_state(Grid(16, 16, 10))
    atol, rtol = tolerances("reference")
    nob = len(wr["block_end"])
    hs, st, mask, rc = O.run_oracle(wr, atol, rtol, solver="ros2")
    a1, s1 = O.block_counts(nob)
    hr, sr, maskr, rcr = O.run_oracle(wr, atol, rtol, solver="ros2", reference_sparse=True)
    a2, s2 = O.block_counts(nob)
    assert rc == 0 and rcr == 0 and np.array_equal(a1, a2) and np.array_equal(s1, s2)
    rep = parity_report(hr.sc_p, hs.sc_p, Grid(16, 16, 10), wr, rel=max(1e-6, 4 * sens))
    assert rep["nfail"] == 0, (rep["worst"], sens)
    say("ROS2 step (relaxed state): oracle = reference Sparse1.3 path (attempts %d per block equal, worst %.2g, "
        "one-ulp sensitivity %.2g)" % (st.attempts, rep["worst"], sens))


def relaxed_state(grid):
    """The seeded split-SZA state after one RODAS3 step of the oracle at the
    reference tolerances, and the oracle ROS2 step's worst change under a
    one-ulp perturbation of it."""
    wl = make_workload(grid, seed=2, sza="split")
    atol, rtol = tolerances("reference")
    h0, _, _, _ = O.run_oracle(wl, atol, rtol)
    wr = dict(wl)
    wr["sc_p"] = [np.asfortranarray(np.asarray(x, np.float32)) for x in h0.sc_p]
    wr["last_accepted_dt"] = np.asarray(h0.last_dt).copy()
    hs, _, _, _ = O.run_oracle(wr, atol, rtol, solver="ros2")
    wp = dict(wr)
    wp["sc_p"] = [np.asfortranarray(np.nextafter(x, np.float32(np.inf))) for x in wr["sc_p"]]
    hp, _, _, _ = O.run_oracle(wp, atol, rtol, solver="ros2")
    return wr, parity_report(hp.sc_p, hs.sc_p, grid, wr)["worst"]


def gpu():
    import torch
    from brams_amd import abi
    from brams_amd.chem import DeviceStep, tuv_jphoto_device
    grid = Grid(16, 16, 10)
    assert abi.mechanism() == (MECH.name, NS, MECH.nr, NP), abi.mechanism()
    for tol in ("baseline", "reference"):
        wl = make_workload(grid, seed=1, sza="split")
        atol, rtol = tolerances(tol)
        ds = DeviceStep(wl, atol, rtol)
        rc = ds.step()
        st = abi.stats()
        gpu_sc = ds.species_host()
        gdt = ds.last_dt.cpu().numpy()
        del ds
        hs, ost, mask, orc = O.run_oracle(wl, atol, rtol)
        rep = parity_report(gpu_sc, hs.sc_p, grid, wl)
        assert rc == 0 and orc == 0
        assert st.attempts == ost.attempts and st.substeps == ost.substeps, (st.attempts, ost.attempts)
        assert st.mask_nnz == ost.mask_nnz, (st.mask_nnz, ost.mask_nnz)
        assert rep["nfail"] == 0, rep["worst"]
        # the parity bar for last_accepted_dt (rel 1e-6): FMA contraction in the
        # error norm moves RACM_TUV's controller output by about 1e-9
        np.testing.assert_allclose(gdt, hs.last_dt, rtol=1e-6)
        say("GPU RODAS3 16x16x10 %s: attempts %d = oracle, mask %d, worst %.2g, %.2f ms"
            % (tol, st.attempts, st.mask_nnz, rep["worst"], st.kernel_ms))
    # ROS2 from a relaxed state: one RODAS3 step of the seeded state (oracle,
    # reference tolerances) first.  Out of the seeded state, ROS2 force-accepts
    # every block of CB07_TUV and RACM_TUV, and there a one-ulp change of the
    # input moves the reference algorithm's own output by up to 1e20 (the
    # values are then defined by round-off only); from the relaxed state no
    # block is forced.  The bar: rel 1e-6, or 4x the oracle's own change under
    # a one-ulp perturbation of the same input if that is larger.
    wr, sens = relaxed_state(grid)
    atol, rtol = tolerances("reference")
    ds = DeviceStep(wr, atol, rtol, solver="ros2")
    rc = ds.step()
    st = abi.stats()
    gpu_sc = ds.species_host()
    del ds
    hs, ost, mask, orc = O.run_oracle(wr, atol, rtol, solver="ros2")
    rep = parity_report(gpu_sc, hs.sc_p, grid, wr, rel=max(1e-6, 4 * sens))
    assert rc == 0 and orc == 0 and st.attempts == ost.attempts and ost.forced_blocks == 0 and rep["nfail"] == 0, \
        (rc, orc, st.attempts, ost.attempts, ost.forced_blocks, rep["worst"], sens)
    say("GPU ROS2 16x16x10 (relaxed state, reference tolerances): attempts %d = oracle, worst %.2g "
        "(oracle's one-ulp sensitivity %.2g)" % (st.attempts, rep["worst"], sens))
    wl = make_workload(grid, seed=4, sza="split", all_ppbm=True)
    ds = DeviceStep(wl, np.zeros(NS), np.zeros(NS))
    rc = ds.qssa(75)
    tend = ds.tendency_host()
    del ds
    hq, orc = O.run_oracle_qssa(wl)
    # the QSSA parity bar of tests/test_gpu_qssa.py: rel 1e-6 of the implied
    # mixing ratio, plus the conditioning of the reference's recurrence
    from test_gpu_qssa import compare
    worst, nleft = compare(grid, wl, tend, hq)
    assert rc == 0 and orc == 0 and nleft == 0, (rc, orc, nleft)
    say("GPU QSSA 16x16x10: worst %.2g, none beyond the conditioning bound" % worst)
    kj, kz = 40, grid.m1 + 1
    rng = np.random.default_rng(5)
    ni, nj = grid.iz - grid.ia + 1, grid.jz - grid.ja + 1
    valj = np.asfortranarray(rng.uniform(0.0, 1e-2, (ni, nj, kj, kz)))
    xref = np.zeros(kj, np.int32)
    xref[1::2][:NP] = np.arange(1, min(NP, kj // 2) + 1)
    before = np.asfortranarray(rng.uniform(1e-6, 1e-3, (NP, grid.m1, grid.m2, grid.m3)))
    dv = torch.from_numpy(np.ascontiguousarray(valj.ravel(order="F"))).cuda()
    dj = torch.from_numpy(np.ascontiguousarray(before.ravel(order="F"))).cuda()
    tuv_jphoto_device(grid, dv, kj, kz, xref, dj)
    got = dj.cpu().numpy().reshape(before.shape, order="F")
    assert np.array_equal(got, O.tuv_jphoto(grid, valj, xref, before)), "tuv mapping"
    say("GPU TUV mapping with the J relations: bit-identical")
    from test_gpu_forms_grids import test_grids_in_one_pass_match_one_call_per_grid
    test_grids_in_one_pass_match_one_call_per_grid()
    say("GPU several grids in one pass = one call per grid, bit for bit (two calls)")
    from test_zz_no_alloc import test_noalloc_entry_matches_device_entry
    test_noalloc_entry_matches_device_entry()
    say("GPU no-allocation entry = device entry, bit for bit (three grids in a row, shrinking)")
    for method, ndyn in (("SEQUENTIAL", 1), ("PARALLEL", 1), ("PARALLEL", 2)):
        worst, sens = split_case(method, ndyn)
        say("GPU split write-back %s N_DYN_CHEM=%d through the host entry: worst %.2g (oracle's one-ulp "
            "sensitivity %.2g)" % (method, ndyn, worst, sens))
    fortran_drop_ins()


def fortran_drop_ins():
    """The Fortran drop-ins (chem_rodas3_dyndt, chem_ros_dyndt, chem_qssa)
    linked to this mechanism's library (brams_amd/fortran/shim_driver_<mech>)
    against the C-ABI called from Python: bit for bit."""
    import pathlib
    import tempfile
    import test_gpu_fortran_shim as FS
    from brams_amd.chem import run_workload_host, run_workload_qssa_host
    atol, rtol = tolerances("baseline")
    with tempfile.TemporaryDirectory() as td:
        tp = pathlib.Path(td)
        wl = make_workload(Grid(10, 8, 9), seed=11)
        for solver in ("rodas3", "ros2"):
            fort, fdt = FS.run_driver(tp, wl, atol, rtol, solver)
            py, pdt, rc, _ = run_workload_host(wl, atol, rtol, solver=solver)
            assert rc == 0 and all(np.array_equal(a, b) for a, b in zip(fort, py)) and np.array_equal(fdt, pdt), solver
        wq = make_workload(Grid(10, 8, 9), seed=11, sza="split", all_ppbm=True)
        fort, _ = FS.run_driver(tp, wq, atol, rtol, "qssa")
        py, rc = run_workload_qssa_host(wq)
        assert rc == 0 and all(np.array_equal(a, b) for a, b in zip(fort, py)), "qssa"
    say("Fortran drop-ins chem_rodas3_dyndt, chem_ros_dyndt and chem_qssa linked to %s = the C-ABI, bit for bit"
        % os.path.basename(FS.DRIVER))


def split_case(method, ndyn):
    """The write-back forms of chem_spack_rodas3_dyndt.f90:713-798 through the
    host entry on the 16x16x6 workload of tests/test_gpu_rodas3.py (SEQUENTIAL
    overwrites sc_p; PARALLEL with N_DYN_CHEM=1 adds (new - old)/dtlt to sc_t;
    N_DYN_CHEM>1 starts from sc_p - sc_t_dyn*N*dtlt).  Every block of it
    accepts at dt_min on these mechanisms, so the bar on the written array is
    rel 1e-6 (+1e-12 of its largest value), or four times the oracle's own
    change under a one-ulp perturbation of the input, whichever is larger."""
    from brams_amd.chem import run_workload_host
    grid = Grid(16, 16, 6)
    wl = make_workload(grid, seed=4, sza="split", plume_sigma=3.0)
    atol, rtol = tolerances("reference")
    code = {"SEQUENTIAL": 1, "PARALLEL": 2}[method]
    sc_t = sc_t_dyn = None
    if method == "PARALLEL" and ndyn == 1:
        rng = np.random.default_rng(4)
        sc_t = [np.asfortranarray((rng.standard_normal(a.shape) * 1e-6).astype(np.float32)) for a in wl["sc_p"]]
    if method == "PARALLEL" and ndyn > 1:
        sc_t_dyn = [np.asfortranarray((-0.01 * a / (ndyn * grid.dtlt)).astype(np.float32)) for a in wl["sc_p"]]
    gpu, dt_gpu, rc, gpu_t = run_workload_host(wl, atol, rtol, split_method=method, n_dyn_chem=ndyn,
                                               sc_t=sc_t, sc_t_dyn=sc_t_dyn)
    hs, _, _, orc = O.run_oracle(wl, atol, rtol, split_method=code, n_dyn_chem=ndyn, sc_t=sc_t, sc_t_dyn=sc_t_dyn)
    wp = dict(wl)
    wp["sc_p"] = [np.asfortranarray(np.nextafter(np.asarray(x, np.float32), np.float32(np.inf))) for x in wl["sc_p"]]
    hp, _, _, _ = O.run_oracle(wp, atol, rtol, split_method=code, n_dyn_chem=ndyn, sc_t=sc_t, sc_t_dyn=sc_t_dyn)
    assert rc == 0 and orc == 0, (rc, orc)
    np.testing.assert_allclose(dt_gpu, hs.last_dt, rtol=1e-6)
    t1 = method == "PARALLEL" and ndyn == 1
    if t1:
        for s, (a, b) in enumerate(zip(gpu, wl["sc_p"])):         # sc_p untouched
            assert np.array_equal(a, b), s
    got, ref, pert = (gpu_t, hs.sc_t, hp.sc_t) if t1 else (gpu, hs.sc_p, hp.sc_p)

    def rel(x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        # rel <= 1e-6 is |x - y| <= 1e-6 |y| + 1e-12 max|y|
        return np.abs(x - y) / (np.abs(y) + 1e-6 * max(np.abs(y).max(), 1e-30))

    sens = max(float(rel(c, r).max()) for c, r in zip(pert, ref))
    worst = max(float(rel(g, r).max()) for g, r in zip(got, ref))
    assert worst <= max(1e-6, 4 * sens), (method, ndyn, worst, sens)
    return worst, sens


def gpu_scale():
    # last_accepted_dt at the general bar (rel 1e-6): FMA contraction in the
    # error norm moves RACM_TUV's controller output by about 1e-9
    from test_gpu_scale_parity import run_case
    st = run_case(Grid(50, 50, 40, dtlt=30.0), "split", "baseline", every=25, head=40, dt_rtol=1e-6)
    say("GPU RODAS3 50x50x40 (dt 30 s, 65-cell blocks on 128 lanes): sampled blocks = oracle, %.1f ms"
        % st.kernel_ms)
    # without the x50 NOx/VOC plume: on CB07_TUV and RACM_TUV it blows up
    # hundreds of blocks (forced errors past 1e30), and which of them end that
    # way is decided by round-off.  CB07_TUV's cold step out of the seeded
    # state does so even without it (about 200 blocks blown up, 1400 more
    # forced past 1e6); its check takes the second step, where the blocks
    # that blew up are the only ones forced that far.
    relax = MECH.name == "CB07_TUV"
    st = run_case(Grid(200, 200, 40), "split", "reference", every=125, dt_rtol=1e-6, plume=False, relax=relax)
    say("GPU RODAS3 200x200x40 (1.6M cells, no plume%s): sampled blocks = oracle, %.1f ms"
        % (", second step" if relax else "", st.kernel_ms))
    gpu_scale_ros2()
    # QSSA on every cell of the 1.6M-cell grid with the plume, against the
    # oracle's whole-grid QSSA (tests/test_gpu_qssa.py's check and bar)
    from test_gpu_qssa import test_qssa_200x200x40_matches_oracle
    test_qssa_200x200x40_matches_oracle()
    say("GPU QSSA 200x200x40 (every cell, plume) = oracle within the QSSA bar")


def gpu_scale_ros2():
    """ROS2 on the 1.6M-cell grid (the four-wave kernel from 4096 blocks on)
    against the oracle on block 1 and 200 sampled blocks, from the state one
    RODAS3 step of the GPU leaves (reference tolerances; out of the seeded
    state ROS2 accepts at dt_min nearly everywhere and a one-ulp change of
    the input moves its output by up to 1e27).  Equal attempts and sub-steps
    per block, the same blocks failed (those a blown-up RODAS3 step left
    non-finite), and outputs within rel 1e-6 or four times the oracle's own
    change under a one-ulp perturbation of the input, whichever is larger."""
    from brams_amd import abi
    from brams_amd.chem import DeviceStep
    from test_gpu_scale_parity import compare, sample_blocks
    grid = Grid(200, 200, 40)
    print("case ros2 200x200x40 split reference, relaxed state, 201 sampled blocks", flush=True)
    wl = make_workload(grid, seed=0, sza="split")
    atol, rtol = tolerances("reference")
    nob = len(wl["block_end"])
    ds = DeviceStep(wl, atol, rtol)
    ds.step()
    wr = dict(wl)
    wr["sc_p"] = ds.species_host()
    wr["last_accepted_dt"] = ds.last_dt.cpu().numpy().copy()
    del ds
    ds = DeviceStep(wr, atol, rtol, solver="ros2")
    rc = ds.step()
    st = abi.stats()
    gpu = ds.species_host()
    gatt, gsub, gfail, _ = abi.block_status(nob)
    del ds
    blocks = np.union1d(sample_blocks(nob, 125), np.nonzero(gfail)[0][np.nonzero(gfail)[0] > 0]).astype(np.int32)
    ran = np.concatenate([[0], blocks])
    hs, ost, _, orc = O.run_oracle_blocks(wr, atol, rtol, blocks, solver="ros2")
    oatt, osub = O.block_counts(nob)
    ofail = O.block_failed(nob)
    wp = dict(wr)
    wp["sc_p"] = [np.asfortranarray(np.nextafter(np.asarray(x, np.float32), np.float32(np.inf))) for x in wr["sc_p"]]
    hp, _, _, _ = O.run_oracle_blocks(wp, atol, rtol, blocks, solver="ros2")
    ids = ran[(gfail[ran] == 0) & (ofail[ran] == 0)]
    sens, _ = compare(grid, wr, hp.sc_p, hs.sc_p, ids)
    worst, _ = compare(grid, wr, gpu, hs.sc_p, ids)
    assert np.array_equal(gfail[ran] != 0, ofail[ran] != 0) and (rc > 0) == (orc > 0), (rc, orc)
    assert np.array_equal(gatt[ids], oatt[ids]) and np.array_equal(gsub[ids], osub[ids])
    assert worst <= max(1e-6, 4 * sens), (worst, sens)
    say("GPU ROS2 200x200x40 (four waves per SIMD, relaxed state): attempts per sampled block = oracle, "
        "%d failed blocks on both sides, worst %.2g (oracle's one-ulp sensitivity %.2g), %.1f ms"
        % (int(np.count_nonzero(gfail)), worst, sens, st.kernel_ms))

if __name__ == "__main__":
    if MECH.name == "RELACS_TUV":
        print("mech_check.py checks the other mechanisms; set BRAMS_MECH", file=sys.stderr)
        sys.exit(2)
    if "--cpu" in sys.argv:
        cpu()
    if "--gpu" in sys.argv:
        gpu()
    if "--gpu-scale" in sys.argv:
        gpu_scale()
    say("all checks passed")
