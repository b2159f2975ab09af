"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference
pieces built from the reference's own sources (oracle/_ref/libbrams_ref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use these, as the checker.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libbrams_ref.so")

NS, NR, NP = 47, 128, 17

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
                ("mask_nnz", C.c_int32), ("kernel_ms", C.c_float)]


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
        lib.ora_block_counts.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int]
        lib.ora_use_external_sparse.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
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
