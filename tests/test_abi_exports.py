"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares; host-only entry points behave like the reference."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from brams_amd import abi
from brams_amd.workload import Grid, alloc_index

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(brams_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_expected_entry_points():
    names = _declared()
    for n in abi.EXPORTS + ["brams_mlp_create", "brams_mlp_train", "brams_mlp_forward",
                            "brams_mlp_gradients"]:
        assert n in names, n


@pytest.mark.parametrize("name", _declared())
def test_library_exports(name):
    lib = C.CDLL(abi.LIB_PATH)
    assert hasattr(lib, name), name


@pytest.mark.parametrize("mech", ["relacs_mx", "cb07_tuv", "racm_tuv"])
def test_mechanism_libraries_export_the_same_entry_points(mech):
    """libbrams_chem_<mech>.so (the lean build of another mechanism) has the
    same C-ABI, and brams_chem_mechanism names its mechanism."""
    path = abi.lib_path(mech)
    if not os.path.exists(path):
        pytest.skip("%s not built" % os.path.basename(path))
    lib = C.CDLL(path)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.brams_chem_mechanism.restype = C.c_char_p
    lib.brams_chem_mechanism.argtypes = [C.c_void_p] * 3
    assert lib.brams_chem_mechanism(None, None, None).decode() == mech.upper()


@pytest.mark.parametrize("nx,ny,nz", [(16, 16, 10), (98, 98, 40), (50, 50, 40), (3, 3, 12), (1, 1, 5)])
def test_alloc_index_through_c_abi(nx, ny, nz):
    from brams_amd.chem import alloc_index as c_alloc
    g = Grid(nx, ny, nz)
    assert np.array_equal(c_alloc(g.m1, g.ia, g.iz, g.ja, g.jz), alloc_index(g))


def test_invalid_arguments_are_reported_not_aborted():
    L = abi.lib()
    # a host step with inconsistent block partition must return an error code
    a = abi.ChemArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = 12, 18, 18, 2, 17, 2, 17
    rc = L.brams_chem_alloc_index(64, 2, 2, 1, 2, 17, None, 0)
    assert rc < 0                      # m1=2 -> no interior levels
    assert b"empty" in L.brams_chem_last_error()


def test_surrogate_step_rejects_parallel_split():
    """brams_sur_step writes accepted cells straight into sc_p (the
    SYMMETRIC/SEQUENTIAL write-back, rodas3:774-798), so a PARALLEL call is
    refused with an error before any device work."""
    from brams_amd.surrogate import SurModel, SurStats
    L = abi.lib()
    L.brams_sur_step.argtypes = [C.POINTER(abi.ChemDevArgs), C.POINTER(SurModel), C.POINTER(SurStats)]
    a = abi.ChemDevArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = 12, 18, 18, 2, 17, 2, 17
    a.split_method = abi.SPLIT["PARALLEL"]
    a.n_dyn_chem = 1
    m = SurModel()
    m.nmembers = 3
    rc = L.brams_sur_step(C.byref(a), C.byref(m), None)
    assert rc < 0
    assert b"SYMMETRIC and SEQUENTIAL" in L.brams_chem_last_error()


def test_slab_forms_are_refused_while_switched_off():
    """The slab kernel forms (0..2) are switched off in this build: the
    per-lane forms (3, 4, 5: 1, 2, 4 waves per SIMD) and autotune (-1) are
    accepted, a slab form is refused with an error and no device work."""
    L = abi.lib()
    assert L.brams_chem_set_form(0) == -6
    assert b"switched off" in L.brams_chem_last_error()
    assert L.brams_chem_set_form(6) == -2
    for form in (3, 4, 5):
        assert L.brams_chem_set_form(form) == 0
    assert L.brams_chem_set_form(-1) == 0


def test_grids_entry_checks_its_arguments():
    """brams_chem_rodas3_dyndt_device_grids refuses a bad grid count before
    any device work."""
    L = abi.lib()
    L.brams_chem_rodas3_dyndt_device_grids.argtypes = [C.POINTER(abi.ChemDevArgs), C.c_int32]
    arr = (abi.ChemDevArgs * 17)()
    assert L.brams_chem_rodas3_dyndt_device_grids(arr, 0) == -2
    assert L.brams_chem_rodas3_dyndt_device_grids(arr, 17) == -2
    assert b"ngrids" in L.brams_chem_last_error()


def test_noalloc_entry_refuses_grids_beyond_its_limit():
    """brams_chem_rodas3_dyndt_device_noalloc serves at most 32768 blocks;
    a larger grid (300x300x40: 56 250 blocks) is refused with -7 before any
    device work (the pointers are never dereferenced)."""
    L = abi.lib()
    L.brams_chem_rodas3_dyndt_device_noalloc.argtypes = [C.POINTER(abi.ChemDevArgs)]
    g = Grid(300, 300, 40)
    a = abi.ChemDevArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = g.m1, g.m2, g.m3, g.ia, g.iz, g.ja, g.jz
    a.dtlt, a.n_dyn_chem = 120.0, 1
    dummy = 0x1000
    a.pp = a.pi0 = a.theta = a.rv = a.jphoto = a.weight = a.transported = dummy
    a.sc_p = a.last_accepted_dt = a.atol = a.rtol = dummy
    a.nob = g.ncell // 64
    assert a.nob == 56250
    assert L.brams_chem_rodas3_dyndt_device_noalloc(C.byref(a)) == -7
    assert b"no-allocation limit" in L.brams_chem_last_error()


def test_qssa_arguments_are_checked_before_device_work():
    """brams_chem_qssa(_device) refuses itchim < 1 and a missing sc_t with an
    error, before any device work."""
    L = abi.lib()
    a = abi.ChemDevArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = 12, 18, 18, 2, 17, 2, 17
    a.dtlt = 60.0
    assert L.brams_chem_qssa_device(C.byref(a), 0) == -2
    assert b"itchim" in L.brams_chem_last_error()
    assert L.brams_chem_qssa_device(C.byref(a), 75) == -2
    assert b"sc_t" in L.brams_chem_last_error()
    h = abi.ChemArgs()
    h.m1, h.m2, h.m3, h.ia, h.iz, h.ja, h.jz = 12, 18, 18, 2, 17, 2, 17
    h.dtlt = 60.0
    h.nspecies = 30
    assert L.brams_chem_qssa(C.byref(h), 75) == -2
    assert b"nspecies" in L.brams_chem_last_error()
