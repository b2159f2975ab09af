"""Build every native piece in-tree.

  1. brams_amd/codegen/gen_mech.py  -> brams_amd/csrc/generated/relacs_tuv_gen.h
  2. hipcc (gfx950)                 -> brams_amd/libbrams_chem.so
  3. flang                          -> brams_amd/fortran/ shim objects + test driver
  4. make -C oracle                 -> oracle/liboracle.so (+ oracle/_ref when the
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
LIB = os.path.join(PKG, "libbrams_chem.so")
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
_RODAS3_DEPS = ["rodas3_kernel.h", "rodas3_launch.h", "mech_families.h", "generated/relacs_tuv_gen.h"]
UNIT_DEPS = {
    "rodas3_launch.hip": _RODAS3_DEPS,
    "rodas3_launch_b.hip": _RODAS3_DEPS,
    "rodas3_grids.hip": _RODAS3_DEPS,
    "ros2_launch.hip": _RODAS3_DEPS,
    "qssa_launch.hip": _RODAS3_DEPS,
    "rodas3_slab.hip": ["rodas3_slab.h", "rodas3_slab_launch.h", "rodas3_kernel.h", "mech_families.h",
                        "generated/relacs_tuv_gen.h", "generated/relacs_tuv_slab.h"],
}
# parallel compiles (each needs a few GB of host memory)
HIP_JOBS = int(os.environ.get("BRAMS_HIP_JOBS", "6"))
REF_ROOT = os.environ.get("BRAMS_REFERENCE", "/root/reference")


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def gen():
    run([sys.executable, os.path.join(PKG, "codegen", "gen_mech.py")])
    run([sys.executable, os.path.join(ROOT, "oracle", "gen_tables.py")])


def _headers():
    out = []
    for d in (CSRC, os.path.join(CSRC, "generated"), os.path.join(ROOT, "include")):
        out += [os.path.join(d, f) for f in os.listdir(d) if f.endswith(".h")]
    return out


def build_hip(force=True):
    """Compile each unit to an object under _build/ and link the library.
    force=False recompiles only units older than their source or a header."""
    os.makedirs(OBJ, exist_ok=True)
    objs, jobs = [], []
    for unit in HIP_UNITS:
        src = os.path.join(CSRC, unit)
        obj = os.path.join(OBJ, unit.replace(".hip", ".o"))
        deps = [os.path.join(CSRC, h) for h in UNIT_DEPS[unit]] if unit in UNIT_DEPS else _headers()
        stale = (force or not os.path.exists(obj)
                 or os.path.getmtime(obj) < max(os.path.getmtime(f) for f in deps + [src]))
        if stale:
            jobs.append((src, obj))
        objs.append(obj)

    def compile_unit(job):
        src, obj = job
        run([HIPCC] + HIP_FLAGS + ["-I", CSRC, "-c", "-o", obj + ".tmp", src])
        os.replace(obj + ".tmp", obj)

    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, HIP_JOBS)) as ex:
        for f in [ex.submit(compile_unit, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)


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
