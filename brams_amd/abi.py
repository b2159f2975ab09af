"""ctypes view of the C-ABI in include/brams_chem.h (libbrams_chem.so).

Loading fails loudly when the HIP library has not been built: there is no CPU
fallback anywhere in the product path.
"""
import ctypes as C
import os

from .mechanism import default_name, load_mechanism

PKG = os.path.dirname(os.path.abspath(__file__))


def lib_path(mech=None):
    """libbrams_chem.so for RELACS_TUV, libbrams_chem_<mech>.so for the others."""
    mech = mech or default_name()
    return os.path.join(PKG, "libbrams_chem.so" if mech == "relacs_tuv" else "libbrams_chem_%s.so" % mech)


LIB_PATH = lib_path()
_MECH = load_mechanism()
NSPECIES, NREACT, NPHOTO = _MECH.nspecies, _MECH.nr, _MECH.nr_photo
SPLIT = {"SYMMETRIC": 0, "SEQUENTIAL": 1, "PARALLEL": 2}
PHOTOJ = {"FAST-TUV": 0, "FAST-JX": 1, "LUT": 2}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)

# symbols include/brams_chem.h declares
EXPORTS = ["brams_chem_init", "brams_chem_rodas3_dyndt", "brams_chem_rodas3_dyndt_device",
           "brams_chem_rodas3_dyndt_device_noalloc", "brams_chem_rodas3_dyndt_device_grids",
           "brams_chem_accum_device", "brams_chem_ros_dyndt", "brams_chem_ros_dyndt_device",
           "brams_chem_qssa", "brams_chem_qssa_device", "brams_chem_tuv_jphoto_device",
           "brams_chem_alloc_index", "brams_chem_get_stats", "brams_chem_last_error",
           "brams_chem_finalize", "brams_chem_block_status", "brams_chem_block_thresholds",
           "brams_chem_set_form", "brams_chem_get_tuning", "brams_chem_mechanism",
           "brams_chem_dt_history_bytes", "brams_chem_dt_history_pack", "brams_chem_dt_history_unpack"]


class ChemArgs(C.Structure):
    """brams_chem_args"""
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


class TuvMapArgs(C.Structure):
    """brams_tuv_map_args"""
    _fields_ = [
        ("m1", C.c_int32), ("m2", C.c_int32), ("m3", C.c_int32),
        ("ia", C.c_int32), ("iz", C.c_int32), ("ja", C.c_int32), ("jz", C.c_int32),
        ("kj", C.c_int32), ("kz", C.c_int32),
        ("valj", C.c_void_p), ("xref", C.POINTER(C.c_int32)), ("nr_photo", C.c_int32),
        ("jphoto", C.c_void_p), ("stream", C.c_void_p),
    ]


class DtHistoryGrid(C.Structure):
    """brams_dt_history_grid"""
    _fields_ = [(k, C.c_int32) for k in ("m1", "m2", "m3", "ia", "iz", "ja", "jz", "block_size")]


class ChemDevArgs(C.Structure):
    """brams_chem_dev_args"""
    _fields_ = [
        ("m1", C.c_int32), ("m2", C.c_int32), ("m3", C.c_int32),
        ("ia", C.c_int32), ("iz", C.c_int32), ("ja", C.c_int32), ("jz", C.c_int32),
        ("dtlt", C.c_float), ("n_dyn_chem", C.c_int32), ("split_method", C.c_int32),
        ("photoj_method", C.c_int32),
        ("cp", C.c_float), ("cpor", C.c_float), ("p00", C.c_float),
        ("pp", C.c_void_p), ("pi0", C.c_void_p), ("theta", C.c_void_p), ("rv", C.c_void_p),
        ("jphoto", C.c_void_p),
        ("weight", C.c_void_p), ("transported", C.c_void_p),
        ("sc_p", C.c_void_p), ("sc_t", C.c_void_p), ("sc_t_dyn", C.c_void_p),
        ("nob", C.c_int32),
        ("last_accepted_dt", C.c_void_p),
        ("atol", C.c_void_p), ("rtol", C.c_void_p),
        ("stream", C.c_void_p),
    ]


class ChemStats(C.Structure):
    """brams_chem_stats"""
    _fields_ = [("cells", C.c_int64), ("blocks", C.c_int64), ("substeps", C.c_int64),
                ("attempts", C.c_int64), ("failed_blocks", C.c_int64),
                ("mask_nnz", C.c_int32), ("kernel_ms", C.c_float),
                ("cell_substeps", C.c_int64), ("cell_attempts", C.c_int64),
                ("block1_ms", C.c_float), ("main_ms", C.c_float), ("main_cells", C.c_int64),
                ("forced_blocks", C.c_int64), ("max_forced_err", C.c_float),
                ("threshold_pivots", C.c_int64), ("order_threshold_pivots", C.c_int64),
                ("kernel_form", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ChemError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libbrams_chem.so; raise if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChemError("%s is not built (run __graft_entry__.build()); "
                            "there is no CPU fallback" % os.path.basename(LIB_PATH))
        # torch ships its own HIP runtime under the soname libamdhip64.so.7.
        # Loaded first, it also serves this library (one runtime, one HSA
        # instance); loaded after /opt/rocm's copy, torch would start a
        # second HIP/HSA runtime that sees no devices.  So torch goes first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.brams_chem_init.argtypes = [C.c_int]
        L.brams_chem_rodas3_dyndt.argtypes = [C.POINTER(ChemArgs)]
        L.brams_chem_rodas3_dyndt_device.argtypes = [C.POINTER(ChemDevArgs)]
        L.brams_chem_ros_dyndt.argtypes = [C.POINTER(ChemArgs)]
        L.brams_chem_ros_dyndt_device.argtypes = [C.POINTER(ChemDevArgs)]
        L.brams_chem_qssa.argtypes = [C.POINTER(ChemArgs), C.c_int32]
        L.brams_chem_qssa_device.argtypes = [C.POINTER(ChemDevArgs), C.c_int32]
        L.brams_chem_tuv_jphoto_device.argtypes = [C.POINTER(TuvMapArgs)]
        L.brams_chem_alloc_index.argtypes = [C.c_int32] * 6 + [_ip, C.c_int32]
        L.brams_chem_get_stats.argtypes = [C.POINTER(ChemStats)]
        L.brams_chem_last_error.restype = C.c_char_p
        L.brams_chem_finalize.argtypes = []
        L.brams_chem_block_status.argtypes = [_ip, _ip, _ip, C.POINTER(C.c_float), C.c_int32]
        L.brams_chem_block_thresholds.argtypes = [_ip, C.c_int32]
        L.brams_chem_set_form.argtypes = [C.c_int32]
        L.brams_chem_get_tuning.argtypes = [_ip, C.POINTER(C.c_float), C.c_int32]
        L.brams_chem_mechanism.argtypes = [_ip, _ip, _ip]
        L.brams_chem_mechanism.restype = C.c_char_p
        L.brams_chem_dt_history_bytes.argtypes = [C.c_int32]
        L.brams_chem_dt_history_bytes.restype = C.c_int64
        L.brams_chem_dt_history_pack.argtypes = [C.POINTER(DtHistoryGrid), _dp, C.c_int32, C.c_void_p, C.c_int64]
        L.brams_chem_dt_history_pack.restype = C.c_int64
        L.brams_chem_dt_history_unpack.argtypes = [C.c_void_p, C.c_int64, C.POINTER(DtHistoryGrid), _dp,
                                                   C.c_int32]
        _lib = L
    return _lib


def check(rc, what="call"):
    """Raise on a fatal status; return the count of failed blocks otherwise."""
    if rc < 0:
        raise ChemError("%s failed (%d): %s" % (what, rc, lib().brams_chem_last_error().decode()))
    return rc


def block_status(nob):
    """Per-block (attempts, substeps, failed, forced_err) of the last step."""
    import numpy as np
    a = np.zeros(nob, np.int32); s = np.zeros(nob, np.int32); f = np.zeros(nob, np.int32)
    e = np.zeros(nob, np.float32)
    check(lib().brams_chem_block_status(a.ctypes.data_as(_ip), s.ctypes.data_as(_ip),
                                        f.ctypes.data_as(_ip), e.ctypes.data_as(C.POINTER(C.c_float)),
                                        nob), "brams_chem_block_status")
    return a, s, f, e


def block_thresholds(nob):
    """Per-block count of pivots failing Sparse1.3's threshold rule, last step."""
    import numpy as np
    c = np.zeros(nob, np.int32)
    check(lib().brams_chem_block_thresholds(c.ctypes.data_as(_ip), nob), "brams_chem_block_thresholds")
    return c


def set_form(form):
    """Kernel form of the 64-cell blocks (-1: autotune per grid); see brams_chem_set_form."""
    check(lib().brams_chem_set_form(form), "brams_chem_set_form")


def tuning():
    """(chosen form of the current grid, autotuner timings in ms per form)."""
    import numpy as np
    f = np.zeros(1, np.int32)
    ms = np.zeros(8, np.float32)
    n = check(lib().brams_chem_get_tuning(f.ctypes.data_as(_ip), ms.ctypes.data_as(C.POINTER(C.c_float)), 8),
              "brams_chem_get_tuning")
    return int(f[0]), ms[:n]


def mechanism():
    """(name, nspecies, nr, nr_photo) of the mechanism the library holds."""
    n = [C.c_int32() for _ in range(3)]
    name = lib().brams_chem_mechanism(*[C.byref(x) for x in n]).decode()
    return (name,) + tuple(x.value for x in n)


def stats():
    s = ChemStats()
    check(lib().brams_chem_get_stats(C.byref(s)), "brams_chem_get_stats")
    return s
