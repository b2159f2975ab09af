"""The generated code of each mechanism compiled for the host (g++, no
contraction) and checked bit for bit: the slab kernel's accessor-templated functions
(right-looking factorisation with lazy assembly, Jacobian entries, solve,
rates, kinetics) against the plain generated functions and the oracle.
Source: tests/host_kernels/check_generated.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
MECHS = ["relacs_tuv", "relacs_mx", "cb07_tuv", "racm_tuv"]


@pytest.mark.parametrize("mech", MECHS)
def test_generated_kernels_match_on_host(tmp_path, mech):
    exe = str(tmp_path / "check_generated")
    oracle_dir = os.path.join(ROOT, "oracle")
    lib = "oracle" if mech == "relacs_tuv" else "oracle_" + mech
    if not os.path.exists(os.path.join(oracle_dir, "lib%s.so" % lib)):
        pytest.skip("oracle/lib%s.so not built" % lib)
    defs = [] if mech == "relacs_tuv" else ['-DBRAMS_MECH_GEN="generated/%s_gen.h"' % mech,
                                            '-DBRAMS_MECH_SLAB="generated/%s_slab.h"' % mech]
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off"] + defs +
                   ["-I", os.path.join(HERE, "host_kernels"), "-I", os.path.join(ROOT, "brams_amd", "csrc"),
                    "-o", exe, os.path.join(HERE, "host_kernels", "check_generated.cpp"),
                    "-L", oracle_dir, "-l" + lib, "-Wl,-rpath," + oracle_dir], check=True)
    out = subprocess.run([exe, "400", "7"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout
