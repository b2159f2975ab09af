"""Build every native piece in-tree.

  1. brams_amd/codegen/gen_mech.py  -> brams_amd/csrc/generated/relacs_tuv_gen.h
  2. hipcc (gfx950)                 -> brams_amd/libbrams_chem.so
  3. flang                          -> brams_amd/fortran/ shim objects + test driver
  4. make -C oracle                 -> oracle/liboracle.so (+ oracle/_ref when the
                                       reference tree is present)
"""
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
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
REF_ROOT = os.environ.get("BRAMS_REFERENCE", "/root/reference")


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def gen():
    run([sys.executable, os.path.join(PKG, "codegen", "gen_mech.py")])
    run([sys.executable, os.path.join(ROOT, "oracle", "gen_tables.py")])


def build_hip():
    srcs = [os.path.join(CSRC, "brams_chem.hip"), os.path.join(CSRC, "surrogate.hip")]
    tmp = LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-I", CSRC, "-o", tmp] + srcs)
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
