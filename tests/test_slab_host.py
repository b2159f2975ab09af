"""The slab kernel's block integration (brams_amd/csrc/rodas3_slab.h) emulated
on the host for one-cell grids (tests/host_kernels/slab_host.cpp), against
the oracle's step: FP32 outputs, attempts, accepted sub-steps and
last_accepted_dt identical.  With one cell every lane of the wave holds the
same values, so one host lane is the whole wave.  The slab kernels are not
launched by the library in this build; this keeps their arithmetic checked."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
from brams_amd.mechanism import load_mechanism
from brams_amd.workload import Grid, make_workload, tolerances

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
MECH = load_mechanism()
# clang: the slab header's buffer accessors use its ext_vector_type extension
CLANG = os.environ.get("HOST_CLANG", "/opt/rocm/llvm/bin/clang++")


@pytest.fixture(scope="module")
def slab_host(tmp_path_factory):
    lib = str(tmp_path_factory.mktemp("slab") / "libslabhost.so")
    subprocess.run([CLANG, "-O1", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I", os.path.join(HERE, "host_kernels"), "-I", os.path.join(ROOT, "brams_amd", "csrc"),
                    "-o", lib, os.path.join(HERE, "host_kernels", "slab_host.cpp")], check=True)
    return C.CDLL(lib)


@pytest.mark.parametrize("tol", ["baseline", "reference"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_slab_block_matches_oracle_on_one_cell(slab_host, seed, tol):
    wl = make_workload(Grid(1, 1, 1), seed=seed, sza="random")
    atol, rtol = tolerances(tol)
    hs, st, mask, rc = oracle_lib.run_oracle(wl, atol, rtol)
    oatt, osub = oracle_lib.block_counts(1)
    g = wl["grid"]

    def f(a):
        return np.ascontiguousarray(np.asarray(a).ravel(order="F"))

    sc = np.ascontiguousarray(np.stack([f(x) for x in wl["sc_p"]]))
    w = np.array([s.weight for s in MECH.species], np.float32)
    tr = np.array([1 if s.transported else 0 for s in MECH.species], np.int32)
    dt = np.array(wl["last_accepted_dt"], np.float64)
    counts = np.zeros(5, np.int32)
    mw = np.zeros(15, np.uint32)
    pp, pi0, th, rv, jp = f(wl["pp"]), f(wl["pi0"]), f(wl["theta"]), f(wl["rv"]), f(wl["jphoto"])
    at, rt = np.ascontiguousarray(atol, np.float64), np.ascontiguousarray(rtol, np.float64)

    def P(a):
        return a.ctypes.data_as(C.c_void_p)

    slab_host.slab_step_host(g.m1, g.m2, g.m3, C.c_float(g.dtlt), P(pp), P(pi0), P(th), P(rv), P(jp),
                             P(w), P(tr), P(sc), P(dt), P(at), P(rt), P(counts), P(mw))
    ref = np.stack([f(x) for x in hs.sc_p])
    assert rc == 0 and counts[2] == 0
    assert counts[0] == oatt[0] and counts[1] == osub[0]
    assert np.array_equal(sc, ref)
    assert dt[0] == hs.last_dt[0]
