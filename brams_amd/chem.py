"""Host-side mirror of the reference's chemistry call, over the C-ABI.

chem_rodas3_dyndt() keeps the argument names, meaning and units of
mod_chem_spack_rodas3_dyndt::chem_rodas3_dyndt
(src/brams/ccatt/chem_spack_rodas3_dyndt.f90:41-131) for the arguments the
RELACS_TUV path reads; it updates chem1_sc_p (or sc_t) and last_accepted_dt in
place like the Fortran routine.  chem_ros_dyndt() is the same for the ROS2
solver (chem_spack_ros_dyndt.f90:39-45, CHEMISTRY=3), and chem_qssa() for
the QSSA solver (chem_spack_qssa.f90:37-43, CHEMISTRY=1), which adds every
species' tendency to sc_t.  DeviceStep runs any of them on arrays that
already live in HBM (torch tensors are used only as device memory).
"""
import ctypes as C

import numpy as np

from . import abi
from .mechanism import load_mechanism

_fp = C.POINTER(C.c_float)


def _f32(a):
    a = np.asarray(a)
    if a.dtype != np.float32 or not a.flags["F_CONTIGUOUS"]:
        raise abi.ChemError("expected a Fortran-ordered float32 array")
    return a.ctypes.data_as(_fp)


def alloc_index(m1, ia, iz, ja, jz, block_size=64):
    """AllocIndex block sizes (chem_spack_utils.f90:39-183), through the C-ABI."""
    L = abi.lib()
    nob = abi.check(L.brams_chem_alloc_index(block_size, m1, ia, iz, ja, jz, None, 0), "alloc_index")
    out = np.zeros(nob, np.int32)
    abi.check(L.brams_chem_alloc_index(block_size, m1, ia, iz, ja, jz,
                                       out.ctypes.data_as(C.POINTER(C.c_int32)), nob), "alloc_index")
    return out


def _dth_grid(grid, block_size):
    return abi.DtHistoryGrid(grid.m1, grid.m2, grid.m3, grid.ia, grid.iz, grid.ja, grid.jz, block_size)


def dt_history_pack(grid, last_accepted_dt, block_size=64):
    """The dt history record of one grid (brams_chem_dt_history_pack): the
    bytes a history (restart) file keeps so that a restarted run takes the
    same per-block step sizes as the uninterrupted one."""
    dt = np.ascontiguousarray(np.asarray(last_accepted_dt, np.float64))
    L = abi.lib()
    n = abi.check(L.brams_chem_dt_history_bytes(len(dt)), "brams_chem_dt_history_bytes")
    buf = C.create_string_buffer(n)
    gr = _dth_grid(grid, block_size)
    got = abi.check(L.brams_chem_dt_history_pack(C.byref(gr), dt.ctypes.data_as(abi._dp), len(dt), buf, n),
                    "brams_chem_dt_history_pack")
    return buf.raw[:got]


def dt_history_unpack(record, grid, block_size=64):
    """last_accepted_dt from a dt history record, checked against the grid it
    is restored into (brams_chem_dt_history_unpack)."""
    nob = alloc_index(grid.m1, grid.ia, grid.iz, grid.ja, grid.jz, block_size).size
    out = np.zeros(nob, np.float64)
    gr = _dth_grid(grid, block_size)
    abi.check(abi.lib().brams_chem_dt_history_unpack(record, len(record), C.byref(gr),
                                                     out.ctypes.data_as(abi._dp), nob),
              "brams_chem_dt_history_unpack")
    return out


SOLVERS = {"rodas3": "brams_chem_rodas3_dyndt", "ros2": "brams_chem_ros_dyndt"}


def chem_rodas3_dyndt(m1, m2, m3, ia, iz, ja, jz, dtlt, pp, pi0, theta, rv, nspecies, nr, nr_photo,
                      weight, PhotojMethod, transp_chem_index, no_transp_chem_index, chem1_sc_p,
                      nob, maxblock_size, block_end, last_accepted_dt, atol, rtol, N_DYN_CHEM,
                      split_method, cp, cpor, p00, jphoto, sc_t=None, sc_t_dyn=None, device=0,
                      solver="rodas3"):
    """One chemistry step on one grid; arrays are updated in place.

    Returns the number of blocks with a numerical failure (the reference would
    abort there); raises abi.ChemError on invalid arguments or a HIP error.
    solver "ros2" runs chem_ros_dyndt's method instead (chem_ros_dyndt())."""
    L = abi.lib()
    abi.check(L.brams_chem_init(device), "brams_chem_init")
    keep = []
    a = abi.ChemArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = m1, m2, m3, ia, iz, ja, jz
    a.dtlt = dtlt
    a.n_dyn_chem = N_DYN_CHEM
    a.split_method = abi.SPLIT[split_method.strip()]
    a.photoj_method = abi.PHOTOJ[PhotojMethod.strip()]
    a.cp, a.cpor, a.p00 = cp, cpor, p00
    a.pp, a.pi0, a.theta, a.rv = _f32(pp), _f32(pi0), _f32(theta), _f32(rv)
    jph = np.asarray(jphoto)
    if jph.dtype != np.float64 or not jph.flags["F_CONTIGUOUS"]:
        raise abi.ChemError("jphoto must be Fortran-ordered float64")
    a.jphoto = jph.ctypes.data_as(C.POINTER(C.c_double))
    a.nspecies, a.nr, a.nr_photo = nspecies, nr, nr_photo
    w = np.ascontiguousarray(weight, np.float32); keep.append(w)
    a.weight = w.ctypes.data_as(_fp)
    ti = np.ascontiguousarray(transp_chem_index, np.int32); keep.append(ti)
    ni = np.ascontiguousarray(no_transp_chem_index, np.int32); keep.append(ni)
    a.n_transported, a.n_not_transported = len(ti), len(ni)
    a.transp_chem_index = ti.ctypes.data_as(C.POINTER(C.c_int32))
    a.no_transp_chem_index = ni.ctypes.data_as(C.POINTER(C.c_int32))
    tab = (_fp * nspecies)(*[_f32(x) for x in chem1_sc_p]); keep.append(tab)
    a.sc_p = C.cast(tab, C.POINTER(_fp))
    if sc_t is not None:
        tt = (_fp * nspecies)(*[_f32(x) for x in sc_t]); keep.append(tt)
        a.sc_t = C.cast(tt, C.POINTER(_fp))
    if sc_t_dyn is not None:
        td = (_fp * nspecies)(*[_f32(x) for x in sc_t_dyn]); keep.append(td)
        a.sc_t_dyn = C.cast(td, C.POINTER(_fp))
    be = np.ascontiguousarray(block_end, np.int32); keep.append(be)
    a.nob, a.maxblock_size = nob, maxblock_size
    a.block_end = be.ctypes.data_as(C.POINTER(C.c_int32))
    if last_accepted_dt.dtype != np.float64 or not last_accepted_dt.flags["C_CONTIGUOUS"]:
        raise abi.ChemError("last_accepted_dt must be a contiguous float64 array")
    a.last_accepted_dt = last_accepted_dt.ctypes.data_as(C.POINTER(C.c_double))
    at = np.ascontiguousarray(atol, np.float64); keep.append(at)
    rt = np.ascontiguousarray(rtol, np.float64); keep.append(rt)
    a.atol = at.ctypes.data_as(C.POINTER(C.c_double))
    a.rtol = rt.ctypes.data_as(C.POINTER(C.c_double))
    name = SOLVERS[solver]
    return abi.check(getattr(L, name)(C.byref(a)), name)


def chem_ros_dyndt(*args, **kw):
    """chem_ros_dyndt (chem_spack_ros_dyndt.f90:39-45): the arguments of
    chem_rodas3_dyndt(), integrated with the ROS2 method."""
    return chem_rodas3_dyndt(*args, solver="ros2", **kw)


def chem_qssa(m1, m2, m3, ia, iz, ja, jz, itchim, dtlt, pp, pi0, theta, rv, jphoto, nspecies, nr,
              nr_photo, weight, PhotojMethod, chem1_sc_p, chem1_sc_t, cp, cpor, p00, device=0):
    """chem_qssa (chem_spack_qssa.f90:37-43) on one grid: every species of
    chem1_sc_p (ppbm) is read, chem1_sc_t += (new - old)/dtlt in place for
    every species; chem1_sc_p is not changed.  Returns the number of cells
    where a species reached 1e20 (the reference stops there)."""
    L = abi.lib()
    abi.check(L.brams_chem_init(device), "brams_chem_init")
    a = abi.ChemArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = m1, m2, m3, ia, iz, ja, jz
    a.dtlt = dtlt
    a.n_dyn_chem = 1
    a.photoj_method = abi.PHOTOJ[PhotojMethod.strip()]
    a.cp, a.cpor, a.p00 = cp, cpor, p00
    a.pp, a.pi0, a.theta, a.rv = _f32(pp), _f32(pi0), _f32(theta), _f32(rv)
    jph = np.asarray(jphoto)
    if jph.dtype != np.float64 or not jph.flags["F_CONTIGUOUS"]:
        raise abi.ChemError("jphoto must be a Fortran-ordered float64 array")
    a.jphoto = jph.ctypes.data_as(C.POINTER(C.c_double))
    a.nspecies, a.nr, a.nr_photo = nspecies, nr, nr_photo
    w = np.ascontiguousarray(weight, dtype=np.float32)
    a.weight = w.ctypes.data_as(_fp)
    tab_p = (_fp * nspecies)(*[_f32(x) for x in chem1_sc_p])
    tab_t = (_fp * nspecies)(*[_f32(x) for x in chem1_sc_t])
    a.sc_p = C.cast(tab_p, C.POINTER(_fp))
    a.sc_t = C.cast(tab_t, C.POINTER(_fp))
    return abi.check(L.brams_chem_qssa(C.byref(a), itchim), "brams_chem_qssa")


def run_workload_qssa_host(wl, itchim=75, sc_t=None):
    """Convenience: chem_qssa on a brams_amd.workload dict; returns the
    tendencies (copies of sc_t, zeros by default) and the status."""
    mech = load_mechanism()
    g = wl["grid"]
    sct = ([np.zeros_like(x, order="F") for x in wl["sc_p"]] if sc_t is None
           else [np.array(x, copy=True, order="F") for x in sc_t])
    rc = chem_qssa(g.m1, g.m2, g.m3, g.ia, g.iz, g.ja, g.jz, itchim, g.dtlt, wl["pp"], wl["pi0"],
                   wl["theta"], wl["rv"], wl["jphoto"], mech.nspecies, mech.nr, mech.nr_photo,
                   [s.weight for s in mech.species], "FAST-TUV", wl["sc_p"], sct, 1004.0,
                   1004.0 / 287.0, 1.0e5)
    return sct, rc


def tuv_jphoto_device(grid, valj, kj, kz, xref, jphoto, stream=None):
    """The FAST-TUV output mapping and RELACS's J relations
    (ModTuvDriver2.7.f90:1302-1332, 1363-1368) on device memory: valj a
    float64 tensor holding valJLoc(ia:iz, ja:jz, kj, kz) in Fortran order,
    xref the TUV cross-reference (kj ints: 1-based jphoto index, 0 unused),
    jphoto a float64 tensor holding (nr_photo, m1, m2, m3) in Fortran order,
    updated in place."""
    L = abi.lib()
    x = np.ascontiguousarray(xref, dtype=np.int32)
    if x.size != kj:
        raise abi.ChemError("xref must hold kj entries")
    a = abi.TuvMapArgs()
    a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = grid.m1, grid.m2, grid.m3, grid.ia, grid.iz, grid.ja, grid.jz
    a.kj, a.kz = kj, kz
    a.valj = valj.data_ptr()
    a.xref = x.ctypes.data_as(C.POINTER(C.c_int32))
    a.nr_photo = load_mechanism().nr_photo
    a.jphoto = jphoto.data_ptr()
    a.stream = stream
    return abi.check(L.brams_chem_tuv_jphoto_device(C.byref(a)), "brams_chem_tuv_jphoto_device")


def run_workload_host(wl, atol, rtol, split_method="SYMMETRIC", n_dyn_chem=1, sc_t=None,
                      sc_t_dyn=None, solver="rodas3"):
    """Convenience: run chem_rodas3_dyndt (or chem_ros_dyndt) on a
    brams_amd.workload dict (copies)."""
    mech = load_mechanism()
    g = wl["grid"]
    sc_p = [np.array(x, copy=True, order="F") for x in wl["sc_p"]]
    last_dt = np.array(wl["last_accepted_dt"], dtype=np.float64, copy=True)
    tidx = [i + 1 for i, s in enumerate(mech.species) if s.transported]
    nidx = [i + 1 for i, s in enumerate(mech.species) if not s.transported]
    sct = None if sc_t is None else [np.array(x, copy=True, order="F") for x in sc_t]
    rc = chem_rodas3_dyndt(g.m1, g.m2, g.m3, g.ia, g.iz, g.ja, g.jz, g.dtlt, wl["pp"], wl["pi0"],
                           wl["theta"], wl["rv"], mech.nspecies, mech.nr, mech.nr_photo,
                           [s.weight for s in mech.species], "FAST-TUV", tidx, nidx, sc_p,
                           len(wl["block_end"]), int(max(wl["block_end"])), wl["block_end"],
                           last_dt, atol, rtol, n_dyn_chem, split_method, 1004.0, 1004.0 / 287.0,
                           1.0e5, wl["jphoto"], sc_t=sct, sc_t_dyn=sc_t_dyn, solver=solver)
    return sc_p, last_dt, rc, sct


class DeviceStep:
    """A chemistry step whose arrays stay resident in HBM across calls.

    Inputs are uploaded once; step() runs brams_chem_rodas3_dyndt_device on the
    resident state (sc_p and last_accepted_dt are updated in place on the
    device, as a time loop would)."""

    def __init__(self, wl, atol, rtol, device=0, split_method="SYMMETRIC", n_dyn_chem=1, solver="rodas3"):
        import torch
        self.torch = torch
        L = abi.lib()
        abi.check(L.brams_chem_init(device), "brams_chem_init")
        mech = load_mechanism()
        g = wl["grid"]
        self.grid = g
        dev = torch.device("cuda", device)
        self.dev = dev

        def up(a, dt):
            return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dt).ravel(order="F"))).to(dev)

        self.pp = up(wl["pp"], np.float32)
        self.pi0 = up(wl["pi0"], np.float32)
        self.theta = up(wl["theta"], np.float32)
        self.rv = up(wl["rv"], np.float32)
        self.jphoto = up(wl["jphoto"], np.float64)
        self.sc_p0 = torch.stack([up(x, np.float32) for x in wl["sc_p"]])       # [nspecies, n3d]
        self.sc_p = self.sc_p0.clone()
        self.weight = torch.tensor([s.weight for s in mech.species], dtype=torch.float32, device=dev)
        self.transported = torch.tensor([1 if s.transported else 0 for s in mech.species],
                                        dtype=torch.int32, device=dev)
        self.last_dt0 = torch.tensor(np.asarray(wl["last_accepted_dt"], np.float64), device=dev)
        self.last_dt = self.last_dt0.clone()
        self.atol = torch.tensor(np.asarray(atol, np.float64), device=dev)
        self.rtol = torch.tensor(np.asarray(rtol, np.float64), device=dev)
        a = abi.ChemDevArgs()
        a.m1, a.m2, a.m3, a.ia, a.iz, a.ja, a.jz = g.m1, g.m2, g.m3, g.ia, g.iz, g.ja, g.jz
        a.dtlt = g.dtlt
        a.n_dyn_chem = n_dyn_chem
        a.split_method = abi.SPLIT[split_method]
        a.photoj_method = 0
        a.cp, a.cpor, a.p00 = 1004.0, 1004.0 / 287.0, 1.0e5
        a.pp, a.pi0, a.theta, a.rv = (self.pp.data_ptr(), self.pi0.data_ptr(),
                                      self.theta.data_ptr(), self.rv.data_ptr())
        a.jphoto = self.jphoto.data_ptr()
        a.weight = self.weight.data_ptr()
        a.transported = self.transported.data_ptr()
        a.sc_p = self.sc_p.data_ptr()
        a.nob = len(wl["block_end"])
        a.last_accepted_dt = self.last_dt.data_ptr()
        a.atol, a.rtol = self.atol.data_ptr(), self.rtol.data_ptr()
        a.stream = torch.cuda.current_stream(dev).cuda_stream
        self.args = a
        self.ncell = g.ncell
        self.entry = SOLVERS[solver] + "_device"

    def reset(self):
        """Restore the initial state (for repeated timing of the same step)."""
        self.sc_p.copy_(self.sc_p0)
        self.last_dt.copy_(self.last_dt0)

    def qssa(self, itchim=75):
        """brams_chem_qssa_device on the resident state: adds every species'
        tendency to self.sc_t ([nspecies, n3d], zeros at first use); sc_p is not
        changed.  Returns the number of cells that reached 1e20."""
        if getattr(self, "sc_t", None) is None:
            self.sc_t = self.torch.zeros_like(self.sc_p)
            self.args.sc_t = self.sc_t.data_ptr()
        L = abi.lib()
        return abi.check(L.brams_chem_qssa_device(C.byref(self.args), itchim), "brams_chem_qssa_device")

    def tendency_host(self):
        """sc_t back on the host as a list of Fortran-ordered (m1,m2,m3) arrays."""
        g = self.grid
        arr = self.sc_t.cpu().numpy()
        return [np.asfortranarray(arr[s].reshape((g.m1, g.m2, g.m3), order="F")) for s in range(arr.shape[0])]

    def step(self):
        return abi.check(getattr(abi.lib(), self.entry)(C.byref(self.args)), self.entry)

    def species_host(self):
        """sc_p back on the host as a list of Fortran-ordered (m1,m2,m3) arrays."""
        g = self.grid
        arr = self.sc_p.cpu().numpy()
        return [np.asfortranarray(arr[s].reshape((g.m1, g.m2, g.m3), order="F")) for s in range(arr.shape[0])]
