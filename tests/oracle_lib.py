"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference
pieces built from the reference's own sources (oracle/_ref/libbrams_ref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use these, as the checker.
"""
import ctypes as C
import os

import numpy as np

from brams_amd.mechanism import default_name, load_mechanism

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MECH = default_name()
_SUFFIX = "" if MECH == "relacs_tuv" else "_" + MECH
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle%s.so" % _SUFFIX)
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libbrams_ref%s.so" % _SUFFIX)

_M = load_mechanism()
NS, NR, NP = _M.nspecies, _M.nr, _M.nr_photo

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


def ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class ChemArgs(C.Structure):
    """Mirror of brams_chem_args in include/brams_chem.h."""
    _fields_ = [
        ("m1", C.c_int32), ("m2", C.c_int32), ("m3", C.c_int32),
        ("ia", C.c_int32), ("iz", C.c_int32), ("ja", C.c_int32), ("jz", C.c_int32),
        ("dtlt", C.c_float), ("n_dyn_chem", C.c_int32), ("split_method", C.c_int32),
        ("photoj_method", C.c_int32),
        ("cp", C.c_float), ("cpor", C.c_float), ("p00", C.c_float),
        ("pp", _fp), ("pi0", _fp), ("theta", _fp), ("rv", _fp),
        ("jphoto", _dp),
        ("nspecies", C.c_int32), ("nr", C.c_int32), ("nr_photo", C.c_int32),
        ("weight", _fp),
        ("n_transported", C.c_int32), ("n_not_transported", C.c_int32),
        ("transp_chem_index", _ip), ("no_transp_chem_index", _ip),
        ("sc_p", C.POINTER(_fp)), ("sc_t", C.POINTER(_fp)), ("sc_t_dyn", C.POINTER(_fp)),
        ("nob", C.c_int32), ("maxblock_size", C.c_int32),
        ("block_end", _ip), ("last_accepted_dt", _dp),
        ("atol", _dp), ("rtol", _dp),
    ]


class ChemStats(C.Structure):
    _fields_ = [("cells", C.c_int64), ("blocks", C.c_int64), ("substeps", C.c_int64),
                ("attempts", C.c_int64), ("failed_blocks", C.c_int64),
                ("mask_nnz", C.c_int32), ("kernel_ms", C.c_float),
                ("cell_substeps", C.c_int64), ("cell_attempts", C.c_int64),
                ("block1_ms", C.c_float), ("main_ms", C.c_float), ("main_cells", C.c_int64),
                ("forced_blocks", C.c_int64), ("max_forced_err", C.c_float),
                ("threshold_pivots", C.c_int64), ("order_threshold_pivots", C.c_int64),
                ("kernel_form", C.c_int32)]


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        lib.ora_kinetic.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        lib.ora_fexchem.argtypes = [_dp, _dp, _dp]
        lib.ora_jacobian.argtypes = [_dp, _dp, _dp]
        lib.ora_lu_order.argtypes = [_ip, _ip]
        lib.ora_lu_solve_dense.argtypes = [_dp, C.POINTER(C.c_uint8), _dp, _dp]
        lib.ora_alloc_index.argtypes = [C.c_int] * 6 + [_ip, C.c_int]
        lib.ora_rodas3_dyndt.argtypes = [C.POINTER(ChemArgs), C.c_int, C.POINTER(ChemStats),
                                         C.POINTER(C.c_uint8)]
        lib.ora_ros2_dyndt.argtypes = [C.POINTER(ChemArgs), C.c_int, C.POINTER(ChemStats),
                                       C.POINTER(C.c_uint8)]
        lib.ora_block_counts.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int]
        lib.ora_use_external_sparse.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ora_qssa.argtypes = [C.POINTER(ChemArgs), C.c_int, C.c_int]
        lib.ora_qssa.restype = C.c_long
        lib.ora_prod_loss.argtypes = [_dp, _dp, _dp, _dp]
        lib.ora_qssa_species.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        lib.ora_qssa_species.restype = C.c_double
        lib.ora_qssa_gather_cell.argtypes = [C.POINTER(ChemArgs), C.c_long, _dp, _dp, _dp, _dp]
        lib.ora_tuv_jphoto.argtypes = [C.c_int] * 9 + [_dp, _ip, C.c_int, _dp]
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref_kinetic.argtypes = [C.c_int] * 3 + [_dp] * 5
        lib.ref_fexchem.argtypes = [C.c_int] * 3 + [_dp] * 3
        lib.ref_jacdchemdc.argtypes = [C.c_int] * 3 + [_dp] * 3
        lib.ref_prod_loss.argtypes = [C.c_int] * 3 + [_dp] * 4
        lib.ref_allocindex.argtypes = ([C.c_int] * 8 + [C.c_float] + [C.c_int] * 3 +
                                       [_ip, _ip, _ip, _ip, _ip, _ip, _dp])
        lib.ref_sp_create.argtypes = [C.c_int, C.c_int, _ip, _ip, C.POINTER(C.c_long)]
        lib.ref_sp_create.restype = C.c_long
        lib.ref_sp_load_factor.argtypes = [C.c_long, C.c_int, C.POINTER(C.c_long), _dp]
        lib.ref_sp_solve.argtypes = [C.c_long, _dp, _dp]
        lib.ref_sp_order.argtypes = [C.c_long, _ip, _ip, _ip]
        lib.ref_sp_destroy.argtypes = [C.c_long]
        _ref = lib
    return _ref


# ------------------------------------------------------------- kernels
def oracle_kinetic(temp, press, vapp, jph):
    """Vector of cells -> rk (n, 128)."""
    lib = oracle()
    n = len(temp)
    rk = np.zeros((n, NR))
    for c in range(n):
        j = np.ascontiguousarray(jph[c], dtype=np.float64)
        out = np.zeros(NR)
        lib.ora_kinetic(float(temp[c]), float(press[c]), float(vapp[c]), ptr(j), ptr(out))
        rk[c] = out
    return rk


def oracle_fexchem(rk, y):
    lib = oracle()
    out = np.zeros_like(y)
    for c in range(len(y)):
        a = np.ascontiguousarray(rk[c]); b = np.ascontiguousarray(y[c]); o = np.zeros(NS)
        lib.ora_fexchem(ptr(a), ptr(b), ptr(o))
        out[c] = o
    return out


def oracle_jacobian(rk, y):
    lib = oracle()
    out = np.zeros((len(y), NS, NS))
    for c in range(len(y)):
        a = np.ascontiguousarray(rk[c]); b = np.ascontiguousarray(y[c]); o = np.zeros(NS * NS)
        lib.ora_jacobian(ptr(a), ptr(b), ptr(o))
        out[c] = o.reshape(NS, NS)
    return out


def ref_kinetic(temp, press, vapp, jph):
    lib = ref()
    n = len(temp)
    jf = np.asfortranarray(np.asarray(jph, dtype=np.float64))      # (n, 17) column-major
    rk = np.zeros((n, NR), order="F")
    t = np.ascontiguousarray(temp, dtype=np.float64)
    v = np.ascontiguousarray(vapp, dtype=np.float64)
    p = np.ascontiguousarray(press, dtype=np.float64)
    lib.ref_kinetic(n, NP, NR, ptr(jf), ptr(rk), ptr(t), ptr(v), ptr(p))
    return np.ascontiguousarray(rk)


def ref_fexchem(rk, y):
    lib = ref()
    n = len(y)
    yf = np.asfortranarray(y); rf = np.asfortranarray(rk); out = np.zeros((n, NS), order="F")
    lib.ref_fexchem(n, NS, NR, ptr(yf), ptr(rf), ptr(out))
    return np.ascontiguousarray(out)


def ref_jacobian(rk, y, chunk=16):
    """The reference's jacdchemdc keeps an automatic dw(n,128,47) array on the
    stack (chem_spack_jacdchemdc.f90:64), so it is called on small chunks."""
    lib = ref()
    outs = []
    for s in range(0, len(y), chunk):
        yc, rc = y[s:s + chunk], rk[s:s + chunk]
        n = len(yc)
        yf = np.asfortranarray(yc); rf = np.asfortranarray(rc)
        out = np.zeros((n, NS, NS), order="F")
        lib.ref_jacdchemdc(n, NS, NR, ptr(yf), ptr(rf), ptr(out))
        outs.append(np.ascontiguousarray(out))
    return np.concatenate(outs)


def ref_prod_loss(rk, y, chunk=16):
    """The reference's rates, dratedc, fexprod and fexloss (the P - L c split
    of chem_spack_qssa.f90:247-265) on cells (n, 47) -> (prod, loss)."""
    lib = ref()
    P, L = [], []
    for s in range(0, len(y), chunk):
        yf = np.asfortranarray(y[s:s + chunk]); rf = np.asfortranarray(rk[s:s + chunk])
        n = len(yf)
        p = np.zeros((n, NS), order="F"); l = np.zeros((n, NS), order="F")
        lib.ref_prod_loss(n, NS, NR, ptr(yf), ptr(rf), ptr(p), ptr(l))
        P.append(np.ascontiguousarray(p)); L.append(np.ascontiguousarray(l))
    return np.concatenate(P), np.concatenate(L)


def oracle_prod_loss(rk, y):
    lib = oracle()
    P = np.zeros_like(y); L = np.zeros_like(y)
    for c in range(len(y)):
        r = np.ascontiguousarray(rk[c]); yc = np.ascontiguousarray(y[c])
        p = np.zeros(NS); l = np.zeros(NS)
        lib.ora_prod_loss(ptr(r), ptr(yc), ptr(p), ptr(l))
        P[c], L[c] = p, l
    return P, L


def tuv_jphoto(grid, valj, xref, jphoto):
    """The oracle's FAST-TUV mapping and RELACS relations: valj a Fortran
    (ni, nj, kj, kz) float64 array, jphoto a Fortran (17, m1, m2, m3) float64
    array (a mapped copy is returned)."""
    v = np.asfortranarray(valj, dtype=np.float64)
    out = np.array(jphoto, dtype=np.float64, order="F", copy=True)
    x = np.ascontiguousarray(xref, dtype=np.int32)
    rc = oracle().ora_tuv_jphoto(grid.m1, grid.m2, grid.m3, grid.ia, grid.iz, grid.ja, grid.jz,
                                 v.shape[2], v.shape[3], ptr(v), ptr(x, _ip), NP, ptr(out))
    if rc != 0:
        raise ValueError("ora_tuv_jphoto: invalid arguments")
    return out


def qssa_species(y, prod, loss, delt, itchim):
    return oracle().ora_qssa_species(y, prod, loss, delt, itchim)


def run_oracle_qssa(wl, itchim=75, nthreads=0, sc_t=None):
    """The oracle's QSSA step on a workload: returns (HostStep, rc) with
    HostStep.sc_t the tendencies (every species), starting from sc_t (zeros
    by default); sc_p is not changed."""
    if sc_t is None:
        sc_t = [np.zeros_like(x) for x in wl["sc_p"]]
    hs = HostStep(wl, np.zeros(NS), np.zeros(NS), 2, 1, sc_t, None)
    rc = oracle().ora_qssa(C.byref(hs.args), itchim, nthreads)
    return hs, rc


def ref_alloc_index(block_size, m1, m2, m3, ia, iz, ja, jz, dtlong, ndyn):
    lib = ref()
    ncell = (m1 - 2) * (iz - ia + 1) * (jz - ja + 1)
    maxnob = max(1, ncell // block_size) + 1
    maxbs = 2 * block_size + 2
    nob = np.zeros(1, np.int32); mbs = np.zeros(1, np.int32)
    be = np.zeros(maxnob, np.int32)
    ik = np.zeros((maxnob, maxbs), np.int32); ii = np.zeros_like(ik); ij = np.zeros_like(ik)
    dt0 = np.zeros(maxnob)
    lib.ref_allocindex(block_size, m1, m2, m3, ia, iz, ja, jz, C.c_float(dtlong), ndyn,
                       maxnob, maxbs, ptr(nob, _ip), ptr(mbs, _ip), ptr(be, _ip),
                       ptr(ik, _ip), ptr(ii, _ip), ptr(ij, _ip), ptr(dt0))
    n = int(nob[0])
    return dict(nob=n, maxblock_size=int(mbs[0]), block_end=be[:n].copy(),
                indexk=ik[:n].copy(), indexi=ii[:n].copy(), indexj=ij[:n].copy(), dt0=dt0[:n].copy())


def use_reference_sparse(enable=True):
    """Route the oracle's LU through the reference's compiled Sparse1.3."""
    lib = oracle()
    if enable:
        r = ref()
        f = lambda fn: C.cast(fn, C.c_void_p).value
        lib.ora_use_external_sparse(f(r.ref_sp_create), f(r.ref_sp_load_factor), f(r.ref_sp_solve))
    else:
        lib.ora_use_external_sparse(None, None, None)


def oracle_lu_order():
    perm = np.zeros(NS, np.int32); nnz = np.zeros(1, np.int32)
    oracle().ora_lu_order(ptr(perm, _ip), ptr(nnz, _ip))
    return perm, int(nnz[0])


# ------------------------------------------------------------- whole step
class HostStep:
    """Owns the numpy arrays behind one brams_chem_args and keeps them alive."""

    def __init__(self, wl, atol, rtol, split_method=0, n_dyn_chem=1, sc_t=None, sc_t_dyn=None):
        from brams_amd.mechanism import load_mechanism
        mech = load_mechanism()
        g = wl["grid"]
        self.wl = wl
        self.sc_p = [np.array(a, copy=True, order="F") for a in wl["sc_p"]]
        self.last_dt = np.array(wl["last_accepted_dt"], dtype=np.float64, copy=True)
        self.weight = np.array([s.weight for s in mech.species], dtype=np.float32)
        self.tidx = np.array([i + 1 for i, s in enumerate(mech.species) if s.transported], np.int32)
        self.nidx = np.array([i + 1 for i, s in enumerate(mech.species) if not s.transported], np.int32)
        self.block_end = np.ascontiguousarray(wl["block_end"], dtype=np.int32)
        self.atol = np.ascontiguousarray(atol, dtype=np.float64)
        self.rtol = np.ascontiguousarray(rtol, dtype=np.float64)
        self.sc_t = None if sc_t is None else [np.array(a, copy=True, order="F") for a in sc_t]
        self.sc_t_dyn = None if sc_t_dyn is None else [np.array(a, copy=True, order="F") for a in sc_t_dyn]
        a = ChemArgs()
        a.m1, a.m2, a.m3 = g.m1, g.m2, g.m3
        a.ia, a.iz, a.ja, a.jz = g.ia, g.iz, g.ja, g.jz
        a.dtlt = g.dtlt
        a.n_dyn_chem = n_dyn_chem
        a.split_method = split_method
        a.photoj_method = 0
        a.cp, a.cpor, a.p00 = 1004.0, 1004.0 / 287.0, 1.0e5
        a.pp, a.pi0 = ptr(wl["pp"], _fp), ptr(wl["pi0"], _fp)
        a.theta, a.rv = ptr(wl["theta"], _fp), ptr(wl["rv"], _fp)
        a.jphoto = ptr(wl["jphoto"])
        a.nspecies, a.nr, a.nr_photo = NS, NR, NP
        a.weight = ptr(self.weight, _fp)
        a.n_transported, a.n_not_transported = len(self.tidx), len(self.nidx)
        a.transp_chem_index = ptr(self.tidx, _ip)
        a.no_transp_chem_index = ptr(self.nidx, _ip)
        self._scp_tab = (_fp * NS)(*[ptr(x, _fp) for x in self.sc_p])
        a.sc_p = C.cast(self._scp_tab, C.POINTER(_fp))
        if self.sc_t is not None:
            self._sct_tab = (_fp * NS)(*[ptr(x, _fp) for x in self.sc_t])
            a.sc_t = C.cast(self._sct_tab, C.POINTER(_fp))
        if self.sc_t_dyn is not None:
            self._sctd_tab = (_fp * NS)(*[ptr(x, _fp) for x in self.sc_t_dyn])
            a.sc_t_dyn = C.cast(self._sctd_tab, C.POINTER(_fp))
        a.nob = len(self.block_end)
        a.maxblock_size = int(self.block_end.max())
        a.block_end = ptr(self.block_end, _ip)
        a.last_accepted_dt = ptr(self.last_dt)
        a.atol, a.rtol = ptr(self.atol), ptr(self.rtol)
        self.args = a


def run_oracle(wl, atol, rtol, nthreads=0, split_method=0, n_dyn_chem=1, sc_t=None, sc_t_dyn=None,
               reference_sparse=False, solver="rodas3"):
    """Run the oracle step on a workload; returns (HostStep, stats, mask, rc).
    solver: "rodas3" (chem_rodas3_dyndt) or "ros2" (chem_ros_dyndt)."""
    hs = HostStep(wl, atol, rtol, split_method, n_dyn_chem, sc_t, sc_t_dyn)
    st = ChemStats()
    mask = np.zeros(NS * NS, np.uint8)
    lib = oracle()
    fn = {"rodas3": lib.ora_rodas3_dyndt, "ros2": lib.ora_ros2_dyndt}[solver]
    if reference_sparse:
        use_reference_sparse(True)
    try:
        rc = fn(C.byref(hs.args), nthreads, C.byref(st), ptr(mask, C.POINTER(C.c_uint8)))
    finally:
        if reference_sparse:
            use_reference_sparse(False)
    return hs, st, mask.reshape(NS, NS), rc


def run_oracle_blocks(wl, atol, rtol, blocks, nthreads=0, solver="rodas3"):
    """The oracle step on block 1 plus the listed blocks (0-based); the other
    blocks' cells are left as they were.  Returns (HostStep, stats, mask, rc)."""
    hs = HostStep(wl, atol, rtol)
    st = ChemStats()
    mask = np.zeros(NS * NS, np.uint8)
    lib = oracle()
    fn = {"rodas3": lib.ora_rodas3_dyndt_blocks, "ros2": lib.ora_ros2_dyndt_blocks}[solver]
    fn.argtypes = [C.POINTER(ChemArgs), C.c_int, C.POINTER(ChemStats), C.POINTER(C.c_uint8), _ip, C.c_int]
    ids = np.ascontiguousarray(blocks, dtype=np.int32)
    rc = fn(C.byref(hs.args), nthreads, C.byref(st), ptr(mask, C.POINTER(C.c_uint8)), ptr(ids, _ip), len(ids))
    return hs, st, mask.reshape(NS, NS), rc


def block_thresholds(nob):
    c = np.zeros(nob, np.int32)
    lib = oracle()
    lib.ora_block_thresholds.argtypes = [_ip, C.c_int]
    lib.ora_block_thresholds(ptr(c, _ip), nob)
    return c


def block_forced(nob):
    """Per-block largest error accepted at dt_min of the last oracle step
    (float; inf when it overflowed)."""
    c = np.zeros(nob, np.float32)
    lib = oracle()
    lib.ora_block_forced.argtypes = [C.POINTER(C.c_float), C.c_int]
    lib.ora_block_forced(c.ctypes.data_as(C.POINTER(C.c_float)), nob)
    return c


def block_failed(nob):
    """Per-block failure flags of the last oracle step."""
    c = np.zeros(nob, np.int32)
    lib = oracle()
    lib.ora_block_failed.argtypes = [_ip, C.c_int]
    lib.ora_block_failed(ptr(c, _ip), nob)
    return c


def gather_cell(wl, cell, atol=None, rtol=None):
    """One cell's gathered state: (y[47] molec/cm^3, T, p, volmol)."""
    hs = HostStep(wl, np.ones(NS) if atol is None else atol, np.ones(NS) if rtol is None else rtol)
    lib = oracle()
    lib.ora_gather_cell.argtypes = [C.POINTER(ChemArgs), C.c_long, _dp, _dp, _dp, _dp]
    y = np.zeros(NS); t = np.zeros(1); p = np.zeros(1); v = np.zeros(1)
    assert lib.ora_gather_cell(C.byref(hs.args), cell, ptr(y), ptr(t), ptr(p), ptr(v)) == 0
    return y, float(t[0]), float(p[0]), float(v[0])


def scatter_cell(wl, cell, y):
    """Write one cell's state y[47] (molec/cm^3) back; returns the species
    arrays after the write-back."""
    hs = HostStep(wl, np.ones(NS), np.ones(NS))
    lib = oracle()
    lib.ora_scatter_cell.argtypes = [C.POINTER(ChemArgs), C.c_long, _dp]
    yy = np.ascontiguousarray(y, dtype=np.float64)
    assert lib.ora_scatter_cell(C.byref(hs.args), cell, ptr(yy)) == 0
    return hs.sc_p


def set_error_script(errs):
    """Scripted block-1 errors for the controller hooks (None switches off)."""
    global _script_keep
    lib = oracle()
    lib.ora_set_error_script.argtypes = [_dp, C.c_int]
    if errs is None:
        _script_keep = None
        lib.ora_set_error_script(None, 0)
    else:
        _script_keep = np.ascontiguousarray(errs, dtype=np.float64)
        lib.ora_set_error_script(ptr(_script_keep), len(_script_keep))


_script_keep = None


def controller_trace(max_n=4096):
    lib = oracle()
    lib.ora_controller_trace.argtypes = [_dp, _ip, _dp, C.c_int]
    dt = np.zeros(max_n); acc = np.zeros(max_n, np.int32); t = np.zeros(max_n)
    n = lib.ora_controller_trace(ptr(dt), ptr(acc, _ip), ptr(t), max_n)
    n = min(n, max_n)
    return dt[:n], acc[:n], t[:n]


def call_counts():
    lib = oracle()
    lib.ora_call_counts.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    j = C.c_int64(); f = C.c_int64()
    lib.ora_call_counts(C.byref(j), C.byref(f))
    return j.value, f.value


def block_forced(nob):
    f = np.zeros(nob, np.float32)
    lib = oracle()
    lib.ora_block_forced.argtypes = [C.POINTER(C.c_float), C.c_int]
    lib.ora_block_forced(f.ctypes.data_as(C.POINTER(C.c_float)), nob)
    return f


def block_counts(nob):
    a = np.zeros(nob, np.int64); s = np.zeros(nob, np.int64)
    oracle().ora_block_counts(a.ctypes.data_as(C.POINTER(C.c_int64)),
                              s.ctypes.data_as(C.POINTER(C.c_int64)), nob)
    return a, s


def rodas3_linear(A, y0, h, nsteps):
    lib = oracle()
    lib.ora_rodas3_linear.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
    A = np.ascontiguousarray(A, np.float64); y0 = np.ascontiguousarray(y0, np.float64)
    out = np.zeros(len(y0)); err = np.zeros(1)
    rc = lib.ora_rodas3_linear(len(y0), ptr(A), ptr(y0), h, nsteps, ptr(out), ptr(err))
    assert rc == 0
    return out, float(err[0])


def dt_new(dt, err, dt_min, dt_max):
    lib = oracle()
    lib.ora_dt_new.argtypes = [C.c_double] * 4
    lib.ora_dt_new.restype = C.c_double
    return lib.ora_dt_new(dt, err, dt_min, dt_max)


def ros2_linear(A, y0, h, nsteps):
    lib = oracle()
    lib.ora_ros2_linear.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
    A = np.ascontiguousarray(A, np.float64); y0 = np.ascontiguousarray(y0, np.float64)
    out = np.zeros(len(y0)); err = np.zeros(1)
    rc = lib.ora_ros2_linear(len(y0), ptr(A), ptr(y0), h, nsteps, ptr(out), ptr(err))
    assert rc == 0
    return out, float(err[0])


def ros2_dt_new(dt, err, dt_min, dt_max):
    lib = oracle()
    lib.ora_ros2_dt_new.argtypes = [C.c_double] * 4
    lib.ora_ros2_dt_new.restype = C.c_double
    return lib.ora_ros2_dt_new(dt, err, dt_min, dt_max)
