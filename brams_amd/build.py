"""Build every native piece in-tree.

  1. brams_amd/codegen/gen_mech.py  -> brams_amd/csrc/generated/<mech>_gen.h
                                       (+ _slab.h, _flops.json) per mechanism
  2. hipcc (gfx950)                 -> brams_amd/libbrams_chem.so (RELACS_TUV)
                                       and libbrams_chem_<mech>.so for the
                                       other mechanisms (lean kernel set)
  3. flang                          -> brams_amd/fortran/ shim objects + test driver
  4. make -C oracle                 -> oracle/liboracle*.so (+ oracle/_ref when the
                                       reference tree is present)
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
# Floating-point contraction stays at the compiler default (FMA formation
# allowed).  Building the RODAS3 units with -ffp-contract=off made the
# block-1 kernel run away on gfx950 (20000 attempts where the oracle takes
# 49) although the same source built by default matches the oracle exactly
# in outputs and attempt counts; parity at full size is checked by the
# 1.6M-cell tests, which report accept/reject flips separately.
OBJ = os.path.join(PKG, "_build")
# translation units of libbrams_chem.so.  The per-lane RODAS3 and ROS2 kernels
# take about two minutes of compile each, so they are spread over four units
# (rodas3_launch*.hip, rodas3_grids.hip, ros2_launch.hip) that compile in
# parallel; the slab kernels (two minutes) live in rodas3_slab.hip.
HIP_UNITS = ["rodas3_launch.hip", "rodas3_launch_b.hip", "rodas3_grids.hip", "ros2_launch.hip",
             "rodas3_slab.hip", "qssa_launch.hip", "brams_chem.hip", "surrogate.hip"]
# the mechanisms and their libraries: RELACS_TUV with every kernel form, the
# others lean (BRAMS_MECH_LEAN: the per-lane kernels at one wave per SIMD,
# no slab, list or several-grid forms), since one instantiation of RACM's
# kernel takes about three minutes to compile
DEFAULT_MECH = "relacs_tuv"
MECHS = ["racm_tuv", "relacs_tuv", "relacs_mx", "cb07_tuv"]   # longest compile first
_KERNEL_UNITS = {"rodas3_launch.hip", "rodas3_launch_b.hip", "rodas3_grids.hip", "ros2_launch.hip",
                 "qssa_launch.hip"}
_KERNEL_DEPS = ["rodas3_kernel.h", "rodas3_launch.h", "mech_families.h", "mechanism.h"]
# parallel compiles, one per available CPU (at most 16; each needs a few GB
# of host memory)
HIP_JOBS = int(os.environ.get("BRAMS_HIP_JOBS", str(min(16, len(os.sched_getaffinity(0))))))
REF_ROOT = os.environ.get("BRAMS_REFERENCE", "/root/reference")


def lib_file(mech):
    return os.path.join(PKG, "libbrams_chem.so" if mech == DEFAULT_MECH else "libbrams_chem_%s.so" % mech)


def mech_flags(mech):
    if mech == DEFAULT_MECH:
        return []
    return ['-DBRAMS_MECH_GEN="generated/%s_gen.h"' % mech,
            '-DBRAMS_MECH_SLAB="generated/%s_slab.h"' % mech, "-DBRAMS_MECH_LEAN"]


def unit_deps(unit, mech):
    gen_h = ["generated/%s_gen.h" % mech]
    if unit in _KERNEL_UNITS:
        return _KERNEL_DEPS + gen_h
    if unit == "rodas3_slab.hip":
        return ["rodas3_slab.h", "rodas3_slab_launch.h"] + _KERNEL_DEPS + gen_h + ["generated/%s_slab.h" % mech]
    return None     # every header


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def gen():
    for mech in MECHS:
        run([sys.executable, os.path.join(PKG, "codegen", "gen_mech.py"), "--mech", mech])
        run([sys.executable, os.path.join(ROOT, "oracle", "gen_tables.py"), "--mech", mech])


def _headers():
    out = []
    for d in (CSRC, os.path.join(CSRC, "generated"), os.path.join(ROOT, "include")):
        out += [os.path.join(d, f) for f in os.listdir(d) if f.endswith(".h")]
    return out


def build_hip(force=True, mechs=None):
    """Compile each unit of each mechanism's library to an object under
    _build/ (_build/<mech>/ for the lean ones) and link the libraries.
    force=False recompiles only units older than their source or a header."""
    mechs = mechs or MECHS
    jobs, links = [], []
    for mech in mechs:
        odir = OBJ if mech == DEFAULT_MECH else os.path.join(OBJ, mech)
        os.makedirs(odir, exist_ok=True)
        objs = []
        for unit in HIP_UNITS:
            src = os.path.join(CSRC, unit)
            obj = os.path.join(odir, unit.replace(".hip", ".o"))
            d = unit_deps(unit, mech)
            deps = [os.path.join(CSRC, h) for h in d] if d is not None else _headers()
            stale = (force or not os.path.exists(obj)
                     or os.path.getmtime(obj) < max(os.path.getmtime(f) for f in deps + [src]))
            if stale:
                jobs.append((src, obj, mech_flags(mech)))
            objs.append(obj)
        links.append((lib_file(mech), objs))

    def compile_unit(job):
        src, obj, flags = job
        run([HIPCC] + HIP_FLAGS + flags + ["-I", CSRC, "-c", "-o", obj + ".tmp", src])
        os.replace(obj + ".tmp", obj)

    # the kernel units first: they take minutes, the others seconds
    jobs.sort(key=lambda j: os.path.basename(j[0]) not in _KERNEL_UNITS)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, HIP_JOBS)) as ex:
        for f in [ex.submit(compile_unit, j) for j in jobs]:
            f.result()
    for lib, objs in links:
        tmp = lib + ".tmp"
        run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
        os.replace(tmp, lib)


def build_fortran():
    fdir = os.path.join(PKG, "fortran")
    if not shutil.which("flang") or not os.path.exists(os.path.join(fdir, "Makefile")):
        return
    run(["make", "-C", fdir, "-j4"])


def build_oracle():
    targets = ["oracle"]
    if os.path.isdir(os.path.join(REF_ROOT, "src", "brams", "ccatt")):
        targets.append("ref")
    run(["make", "-C", os.path.join(ROOT, "oracle"), "REF=" + REF_ROOT] + targets)


def build_all():
    gen()
    build_oracle()
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build_all()
